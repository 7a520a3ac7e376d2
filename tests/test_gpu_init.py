"""Initial states generated on the device (chll_lattice_init_state,
chll_ensemble_init_state) against the oracle's restatement of the same keyed
draws: spins exact, directors within 1e-15 (device and host sin/cos may differ
by an ulp), aligned states exact; and the sphere moments at L = 128."""
import numpy as np
import pytest

from chllmodel_amd.ensemble import Ensemble
from chllmodel_amd.lattice import Lattice
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

KINDS = {"isotropic": 0, "aligned_x": 1, "aligned_xz": 2}


@pytest.mark.parametrize("dims", [(8, 6, 4), (5, 4, 6), (32, 32, 32), (130, 6, 9)])
@pytest.mark.parametrize("kind", list(KINDS))
def test_lattice_init_matches_oracle(dims, kind):
    print(f"device init {kind} {dims}", flush=True)
    with Lattice(*dims) as lat:
        lat.init_state(kind, seed=99, replica=2, p_flip=0.4)
        got = lat.download()
    want = Oracle().init_state(*dims, kind=KINDS[kind], seed=99, replica=2, p_flip=0.4)
    assert np.array_equal(got["s"], want["s"])
    for f in ("nx", "ny", "nz"):
        if kind == "isotropic":
            assert np.abs(got[f] - want[f]).max() <= 1e-15, f
        else:
            assert np.array_equal(got[f], want[f]), f


def test_isotropic_moments_l128():
    L = 128
    N = L ** 3
    with Lattice(L, L, L) as lat:
        lat.init_state("isotropic", seed=20261203, p_flip=0.5)
        row = lat.measure(0.0, 0.0)
    se2 = np.sqrt(4 / (45 * N))
    for k in ("Qxx", "Qyy", "Qzz"):
        assert abs(row[k] - 1 / 3) < 5 * se2, (k, row[k])
    for k in ("Qxy", "Qxz", "Qyz"):
        assert abs(row[k]) < 5 * np.sqrt(1 / (15 * N)), (k, row[k])
    assert abs(row["eta_exact"]) < 5 / np.sqrt(N)


def test_ensemble_init_matches_oracle():
    reps = [{"ni": 8, "nj": 6, "nk": 4, "ladder": 0, "g": 0.5, "KK": 1.0, "kbt": kT, "d0": 0.3} for kT in (0.5, 0.7)]
    reps.append({"ni": 12, "nj": 4, "nk": 6, "ladder": 1, "g": 0.0, "KK": 1.0, "kbt": 1.0, "d0": 0.3})
    orc = Oracle()
    with Ensemble(reps, arith="fp64") as ens:
        ens.init_state("isotropic", seed=1234, p_flip=0.5)
        for r, rr in enumerate(reps):
            got = ens.download(r)
            want = orc.init_state(rr["ni"], rr["nj"], rr["nk"], kind=0, seed=1234, replica=r, p_flip=0.5)
            assert np.array_equal(got["s"], want["s"]), r
            for f in ("nx", "ny", "nz"):
                assert np.abs(got[f] - want[f]).max() <= 1e-15, (r, f)
