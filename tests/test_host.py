"""Host-side logic that needs no GPU: ensemble validation (rejected before any
device call), observable estimators against closed forms, ChLLSim's output
folders and parameter file (chll/chll.py:117-151), and the vectorised
initialisers against the reference scripts' loop order."""
import os

import numpy as np
import pytest

from chllmodel_amd import observables as ob
from chllmodel_amd.ensemble import Ensemble, phase_scan_replicas
from chllmodel_amd.states import aligned_directors, isotropic_directors, random_spins


def rep(**kw):
    r = {"ni": 4, "nj": 4, "nk": 4, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": 0.5, "d0": 0.3}
    r.update(kw)
    return r


@pytest.mark.parametrize("reps,msg", [
    ([rep(kbt=0.5), rep(kbt=0.4)], "ascend"),
    ([rep(kbt=0.5), rep(kbt=0.9, g=0.7)], "one lattice size"),
    ([rep(ladder=0), rep(ladder=1), rep(ladder=0, kbt=0.9)], "not contiguous"),
    ([rep(ni=5)], "ni even"),
    ([rep(kbt=0.0)], "positive"),
])
def test_ensemble_validation(reps, msg):
    with pytest.raises(RuntimeError, match=msg):
        Ensemble(reps)


def test_phase_scan_shape():
    reps = phase_scan_replicas()
    assert len(reps) == 64
    assert sum(r["ni"] ** 3 for r in reps) == 41938944
    kts = [r["kbt"] for r in reps if r["ladder"] == 0]
    assert abs(kts[0] - 0.2) < 1e-12 and abs(kts[-1] - 2.0) < 1e-12 and np.all(np.diff(kts) > 0)


def test_p2_closed_forms():
    assert abs(ob.p2_from_q([1, 0, 0, 0, 0, 0]) - 1.0) < 1e-15      # all along x
    assert abs(ob.p2_from_q([0, 0, 1, 0, 0, 0]) - 1.0) < 1e-15      # all along z
    assert abs(ob.p2_from_q([1 / 3, 1 / 3, 1 / 3, 0, 0, 0])) < 1e-15  # isotropic
    # uniform director v: <n n> = v v^T -> P2 = 1 for any unit v
    v = np.array([0.3, -0.4, 0.5])
    v /= np.linalg.norm(v)
    q = [v[0] ** 2, v[1] ** 2, v[2] ** 2, v[0] * v[1], v[0] * v[2], v[1] * v[2]]
    assert abs(ob.p2_from_q(q) - 1.0) < 1e-14
    # random isotropic sample: P2 ~ O(1/sqrt(N))
    rs = np.random.RandomState(1)
    n = rs.normal(size=(200000, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    q = [np.mean(n[:, 0] ** 2), np.mean(n[:, 1] ** 2), np.mean(n[:, 2] ** 2), np.mean(n[:, 0] * n[:, 1]),
         np.mean(n[:, 0] * n[:, 2]), np.mean(n[:, 1] * n[:, 2])]
    assert ob.p2_from_q(q) < 0.02


def test_blocking_error_iid_and_correlated():
    rs = np.random.RandomState(2)
    x = rs.normal(size=1 << 14)
    e = ob.blocking_error(x)
    assert 0.8 < e / (1 / np.sqrt(x.size)) < 1.3
    # AR(1) with tau: the blocking estimate approaches sigma*sqrt((1+a)/(1-a)/n)
    a = 0.9
    y = np.zeros(1 << 16)
    for t in range(1, y.size):
        y[t] = a * y[t - 1] + rs.normal()
    want = np.sqrt(1 / (1 - a * a)) * np.sqrt((1 + a) / (1 - a) / y.size)
    assert 0.7 < ob.blocking_error(y) / want < 1.3


def test_fluctuation_estimators():
    rs = np.random.RandomState(3)
    e = 0.01 * rs.normal(size=4096) - 1.0
    cv, err = ob.cv_per_site(e, kT=0.5, N=1000)
    assert abs(cv - 1000 * e.var() / 0.25) < 1e-9 and err > 0
    assert abs(ob.cv_notebook(e, 0.5) - e[-100:].var() / 0.5) < 1e-15
    assert abs(ob.chi_notebook(e, 0.5, 1000) - 1000 * e[-500:].var() / 0.5) < 1e-12


def test_initialisers_match_reference_loop_order():
    """scripts/pdanneal.py:5-16 draws (costh, phi) per site in i, j, k loops;
    chll/chll.py:85-90 one draw per site for the spin."""
    ni, nj, nk = 3, 4, 2
    np.random.seed(7)
    nx, ny, nz = isotropic_directors(ni, nj, nk)
    s = random_spins(ni, nj, nk)
    np.random.seed(7)
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                costh = 2.0 * (np.random.rand() - 0.5)
                sinth = np.sqrt(1.0 - costh * costh)
                phi = np.random.rand() * 2.0 * np.pi
                assert nx[i, j, k] == sinth * np.cos(phi) and ny[i, j, k] == sinth * np.sin(phi)
                assert nz[i, j, k] == costh
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                want = -1 if np.random.rand() < 0.5 else 1
                assert s[i, j, k] == want
    ax, ay, az = aligned_directors(2, 2, 2, "xz")
    assert np.all(ax == 1 / np.sqrt(2)) and np.all(ay == 0) and np.all(az == 1 / np.sqrt(2))


def test_chllsim_folders_and_params(tmp_path, monkeypatch):
    from chllmodel_amd.chll import ChLLSim

    monkeypatch.chdir(tmp_path)
    os.makedirs("data")
    sim = ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="t1")
    assert sim.simpath == os.path.join("./data", "t1") and os.path.isdir(sim.simpath)
    sim.save_params(100, 10)
    txt = open(os.path.join(sim.simpath, "t1t1_params.txt")).read().splitlines()
    assert txt == ["t1", "nsteps = 100 nout = 10", "ni,nj,nk = 4 4 4", "kbt = 0.7", "d = 0.3", "KK = 1.0",
                   "g = 0.5", "rho = 0.0"]
    # a second sim of the same name without overwrite exits, as the reference does
    with pytest.raises(SystemExit):
        ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="t1")
    ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="t1", overwrite=True)
    # name=None: nothing saved, no error
    ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0).save_params(1, 1)


def test_chllsim_csv_and_npz(tmp_path, monkeypatch):
    """The CSV and npz written after a run (chll/chll.py:200-262): header,
    str() of each value, the npz keys, and exit() when a file exists and
    overwrite is False."""
    from chllmodel_amd.chll import ChLLSim

    monkeypatch.chdir(tmp_path)
    os.makedirs("data")
    sim = ChLLSim(2, 2, 2, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="io")
    sim.msteps = np.array([10, 20], dtype=np.int32)
    sim.m = np.asfortranarray([[1.5, 0.25, 1.0, 0.5, 0.1 + 0.2], [-2.0, 0.0, 0.875, 0.125, 1e-20]])
    sim.s[:] = -1
    sim.nx[:] = 1.0
    sim.save_macros()
    sim.save_config()
    assert open("data/io/io.csv").read() == ("Step,TotalEnergy,Eexcess,Paccept,Faccept,D\n"
                                             "10,1.5,0.25,1.0,0.5,0.30000000000000004\n"
                                             "20,-2.0,0.0,0.875,0.125,1e-20\n")
    z = np.load("data/io/io_data.npz", allow_pickle=False)
    assert sorted(z.files) == ["dope", "nx", "ny", "nz", "s"]
    assert np.all(z["nx"] == 1.0) and np.all(z["s"] == -1) and z["s"].dtype == np.int32
    with pytest.raises(SystemExit):
        sim.save_macros()
    with pytest.raises(SystemExit):
        sim.save_config()
    sim.overwrite = True
    sim.m[0, 0] = 3.0
    sim.save_macros()
    assert open("data/io/io.csv").read().splitlines()[1] == "10,3.0,0.25,1.0,0.5,0.30000000000000004"
    # no name: nothing is written and nothing fails
    anon = ChLLSim(2, 2, 2, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0)
    anon.msteps, anon.m = sim.msteps, sim.m
    anon.save_macros()
    anon.save_config()


def test_chllsim_plot_config(tmp_path, monkeypatch):
    pytest.importorskip("matplotlib")
    from chllmodel_amd.chll import ChLLSim

    monkeypatch.chdir(tmp_path)
    os.makedirs("data")
    sim = ChLLSim(6, 4, 5, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="pic")
    sim.nx[:] = 1.0
    sim.s[:] = np.where(np.arange(6)[:, None, None] % 2 == 0, 1, -1)
    out = sim.plot_config()
    assert out == os.path.join("./data", "pic", "pic.png")
    with open(out, "rb") as f:
        assert f.read(8) == b"\x89PNG\r\n\x1a\n"
    # no output folder: an explicit path is required
    with pytest.raises(ValueError):
        ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0).plot_config()


def test_chllsim_missing_data_folder_exits(tmp_path, monkeypatch):
    from chllmodel_amd.chll import ChLLSim

    monkeypatch.chdir(tmp_path)
    with pytest.raises(SystemExit):
        ChLLSim(4, 4, 4, kbt=0.7, d=0.3, KK=1.0, g=0.5, rho=0.0, name="x")



def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 (or a WORLD_SIZE above 1) exits with
    a message before any GPU work."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
    env = dict(os.environ, WORLD_SIZE="2")
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py")], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
