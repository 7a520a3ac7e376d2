"""Loading libchll_mod.so in either order with torch.

The library and torch's ROCm build both need libamdhip64.so.7; a process that
maps two HIP runtimes leaves torch without a GPU.  chllmodel_amd._lib loads
torch's runtime (by path, without importing torch) before the library, so the
order of `import torch` and the library load does not matter; a process that
already holds another HIP runtime gets a clear error instead.  Each case runs
in a fresh child process.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "chllmodel_amd", "libchll_mod.so")

needs_lib = pytest.mark.skipif(not os.path.exists(LIB), reason="libchll_mod.so not built")


def child(code: str) -> subprocess.CompletedProcess:
    env = dict(os.environ)
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT,
                          timeout=300)


def torch_runtime():
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None:
        return None
    p = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    return p if os.path.exists(p) else None


@needs_lib
@pytest.mark.parametrize("torch_first", [False, True])
def test_one_hip_runtime_either_order(torch_first):
    if torch_runtime() is None:
        pytest.skip("torch with a bundled HIP runtime is not installed")
    code = f"""
import os
torch_first = {torch_first}
if torch_first:
    import torch
from chllmodel_amd import _lib
_lib.lib()
import torch
held = _lib._mapped("libamdhip64")
assert len(held) == 1, held
assert os.path.realpath(held[0]) == os.path.realpath({torch_runtime()!r}), held
print("ok", held[0])
"""
    r = child(code)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout


@needs_lib
def test_foreign_runtime_first_fails_clearly():
    if torch_runtime() is None:
        pytest.skip("torch with a bundled HIP runtime is not installed")
    code = f"""
import ctypes
ctypes.CDLL({LIB!r})  # as chll/chll.py does, with no torch in the process
from chllmodel_amd import _lib
try:
    _lib.lib()
except _lib.ChllError as e:
    print("clear error:", e)
else:
    raise SystemExit("expected ChllError")
"""
    r = child(code)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "torch imported afterwards would find no GPU" in r.stdout


@needs_lib
@pytest.mark.gpu
def test_library_before_torch_runs_on_gpu():
    """Fresh process: library first, then torch; both see the GPU, a sweep
    runs through the library and a torch kernel after it."""
    code = """
from chllmodel_amd import _lib
_lib.lib()
from chllmodel_amd.lattice import Lattice
from chllmodel_amd.states import isotropic_state
import torch
assert torch.cuda.is_available(), "torch sees no GPU"
with Lattice(8, 8, 8) as lat:
    lat.upload_state(isotropic_state(8, 8, 8, 3))
    lat.set_step(0.3)
    out = lat.sweeps(1, 4, 2, g=0.5, KK=1.0, kbt=0.7, arith="fp64", want_counts=True)
    assert out["counts"].shape == (4, 2) and out["counts"].sum() > 0
x = torch.ones(1000, device="cuda").sum().item()
assert x == 1000.0
print("ok")
"""
    print("child: library before torch, 8^3 sweeps then a torch kernel", flush=True)
    r = child(code)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
