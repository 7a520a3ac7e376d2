"""ctypes bindings for the test oracle (``libchll_oracle.so``) and, where it was
built, the reference module compiled from source (``_ref/libchll_ref*.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  The product library
(``chllmodel_amd``) never imports this module.

The oracle restates ``chll/chll_mod.f90`` of mdeutsc2/ChLLModel in C
(``oracle/chll_oracle.c``); each entry point names the reference lines it
follows.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libchll_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libchll_ref.so")
REF_STREAM_SO = os.path.join(HERE, "_ref", "libchll_ref_stream.so")

RNG_STREAM, RNG_BUFFER, RNG_PHILOX = 0, 1, 2
BC_OPEN, BC_PERIODIC = 0, 1
ARITH_COMPAT, ARITH_FP64 = 0, 1

_f64p = ct.POINTER(ct.c_double)
_i32p = ct.POINTER(ct.c_int32)
_i64p = ct.POINTER(ct.c_int64)
_u64p = ct.POINTER(ct.c_uint64)
_u32p = ct.POINTER(ct.c_uint32)


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference build when its source is present)."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", HERE, "-B" if force else "-s", "libchll_oracle.so"], check=True)
    if os.path.exists("/root/reference/chll/chll_mod.f90"):
        args = ["make", "-C", HERE, "ref"]
        if force:
            args.insert(1, "-B")
        subprocess.run(args, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)


def _ptr(a, typ):
    return a.ctypes.data_as(typ)


class Oracle:
    """The C restatement (libchll_oracle.so)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        self.lib = ct.CDLL(path)
        L = self.lib
        L.oracle_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.oracle_philox4x32_10.restype = None
        L.oracle_stream_uniform.argtypes = [ct.c_uint64, ct.c_uint64]
        L.oracle_stream_uniform.restype = ct.c_double
        L.oracle_stream_fill.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_int64, _f64p]
        L.oracle_stream_fill.restype = None
        L.oracle_init.argtypes = [_i32p, _i32p, ct.c_int, ct.c_int, ct.c_int]
        L.oracle_init.restype = None
        L.oracle_site_energy.argtypes = [_f64p, _f64p, _f64p, _i32p, ct.c_int, ct.c_int, ct.c_int,
                                         ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                         ct.c_double, ct.c_double, ct.c_double, ct.c_int,
                                         ct.c_double, ct.c_double]
        L.oracle_site_energy.restype = ct.c_double
        L.oracle_trial_director.argtypes = [ct.c_double] * 5 + [_f64p]
        L.oracle_trial_director.restype = None
        L.oracle_etot.argtypes = [_f64p, _f64p, _f64p, _i32p, ct.c_int, ct.c_int, ct.c_int,
                                  ct.c_double, ct.c_double]
        L.oracle_etot.restype = ct.c_double
        L.oracle_energy_true.argtypes = [_f64p, _f64p, _f64p, _i32p, ct.c_int, ct.c_int, ct.c_int,
                                         ct.c_int, ct.c_double, ct.c_double]
        L.oracle_energy_true.restype = ct.c_double
        L.oracle_run.argtypes = [ct.c_int, ct.c_uint64, _u64p, _f64p, ct.c_int64,
                                 ct.c_uint32, ct.c_uint32, ct.c_uint32, ct.c_int, ct.c_int,
                                 ct.c_int, ct.c_int, ct.c_int,
                                 _f64p, _f64p, _f64p, _i32p,
                                 ct.c_double, ct.c_double, _f64p, ct.c_double,
                                 ct.c_int, ct.c_int, ct.c_int, _i32p, _f64p, _i64p]
        L.oracle_run.restype = ct.c_int
        L.oracle_run_mt.argtypes = L.oracle_run.argtypes + [ct.c_int, ct.c_int]
        L.oracle_run_mt.restype = ct.c_int
        L.oracle_rank.argtypes = [ct.c_int] * 5
        L.oracle_rank.restype = ct.c_int64
        L.oracle_init_state.argtypes = [_f64p, _f64p, _f64p, _i32p, ct.c_int, ct.c_int, ct.c_int, ct.c_int,
                                        ct.c_uint64, ct.c_uint32, ct.c_double]
        L.oracle_init_state.restype = None

    # -- RNG --------------------------------------------------------------
    def philox(self, ctr, key):
        c = np.ascontiguousarray(ctr, dtype=np.uint32)
        k = np.ascontiguousarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.oracle_philox4x32_10(_ptr(c, _u32p), _ptr(k, _u32p), _ptr(out, _u32p))
        return out

    def stream(self, seed: int, n0: int, count: int) -> np.ndarray:
        out = np.empty(count, dtype=np.float64)
        self.lib.oracle_stream_fill(seed, n0, count, _ptr(out, _f64p))
        return out

    # -- reference functions ------------------------------------------------
    def init(self, ni, nj, nk):
        """Sublattice lists (chll/chll_mod.f90:9-38)."""
        nsub = (ni * nj * nk) // 2
        n1 = (ni * nj * nk + 1) // 2
        sl1 = np.zeros(n1, dtype=np.int32)
        sl2 = np.zeros(max(nsub, 1), dtype=np.int32)
        self.lib.oracle_init(_ptr(sl1, _i32p), _ptr(sl2, _i32p), ni, nj, nk)
        return sl1, sl2

    def site_energy(self, st, i, j, k, n3, snew, g, KK, bc=BC_OPEN, arith=ARITH_COMPAT):
        """Site energy (chll/chll_mod.f90:363-412); st = dict of F-order arrays."""
        ni, nj, nk = st["nx"].shape
        return self.lib.oracle_site_energy(
            _ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p), _ptr(st["s"], _i32p),
            ni, nj, nk, bc, arith, i, j, k, float(n3[0]), float(n3[1]), float(n3[2]), int(snew),
            float(g), float(KK))

    def trial_director(self, n3, d, u_phi):
        out = np.zeros(3)
        self.lib.oracle_trial_director(float(n3[0]), float(n3[1]), float(n3[2]), float(d), float(u_phi),
                                       _ptr(out, _f64p))
        return out

    def etot(self, st, g, KK):
        """Reported total energy per site (chll/chll_mod.f90:414-456)."""
        ni, nj, nk = st["nx"].shape
        return self.lib.oracle_etot(_ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                                    _ptr(st["s"], _i32p), ni, nj, nk, float(g), float(KK))

    def energy_true(self, st, g, KK, bc=BC_OPEN):
        ni, nj, nk = st["nx"].shape
        return self.lib.oracle_energy_true(_ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                                           _ptr(st["s"], _i32p), ni, nj, nk, bc, float(g), float(KK))

    def init_state(self, ni, nj, nk, kind=0, seed=0, replica=0, p_flip=0.5):
        """Initial state keyed by (seed, replica, site): kind 0 isotropic,
        1 aligned x, 2 aligned xz; spins -1 with probability p_flip."""
        st = {f: np.zeros((ni, nj, nk), dtype=np.float64, order="F") for f in ("nx", "ny", "nz")}
        st["s"] = np.zeros((ni, nj, nk), dtype=np.int32, order="F")
        self.lib.oracle_init_state(_ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                                   _ptr(st["s"], _i32p), ni, nj, nk, int(kind), int(seed), int(replica),
                                   float(p_flip))
        st["dope"] = np.zeros((ni, nj, nk), dtype=np.int32, order="F")
        return st

    def rank(self, ni, nj, i, j, k):
        """Rank of site (i, j, k) (0-based) in its sublattice list."""
        return int(self.lib.oracle_rank(ni, nj, i, j, k))

    def run(self, st, nstart, nstop, nout, g, KK, d, kbt, *, rng=RNG_STREAM, seed=0, stream_pos=0,
            buf=None, run_index=0, replica=0, bc=BC_OPEN, arith=ARITH_COMPAT, want_counts=False,
            site_offset=0, threads=None, hostfill=False):
        """Sweep driver (chll/chll_mod.f90:40-115).  Mutates st in place.

        site_offset (Philox draws only) is added to the sublattice rank of
        every draw's counter, so st can be the slab of planes k0.. of a larger
        lattice with the same ni, nj (site_offset = ni nj k0 / 2, k0 even).
        threads: run the half-sweeps on that many OpenMP threads (oracle_run_mt;
        same results).  hostfill (stream draws): fill each sweep's 10 nsub
        uniforms into a host array first, as the reference's random_number
        calls do (same results).

        Returns (msteps, m, d_out, stream_pos_out[, counts])."""
        ni, nj, nk = st["nx"].shape
        rows = max((1 + nstop - nstart) // nout, 0)
        msteps = np.zeros(max(rows, 1), dtype=np.int32)
        m = np.zeros((max(rows, 1), 5), dtype=np.float64, order="F")
        dd = ct.c_double(d)
        pos = ct.c_uint64(stream_pos)
        nsw = nstop - nstart + 1
        counts = np.zeros(2 * max(nsw, 1), dtype=np.int64)
        if buf is not None:
            buf = np.ascontiguousarray(buf, dtype=np.float64)
            bptr, blen = _ptr(buf, _f64p), buf.size
        else:
            bptr, blen = None, 0
        args = (rng, seed, ct.byref(pos), bptr, blen, run_index, replica, site_offset, bc, arith,
                nstart, nstop, nout,
                _ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                _ptr(st["s"], _i32p), float(g), float(KK), ct.byref(dd), float(kbt),
                ni, nj, nk, _ptr(msteps, _i32p), _ptr(m, _f64p), _ptr(counts, _i64p))
        if threads is None and not hostfill:
            rc = self.lib.oracle_run(*args)
        else:
            rc = self.lib.oracle_run_mt(*args, int(threads or 1), 1 if hostfill else 0)
        if rc == -2:
            raise MemoryError("oracle_run_mt: host uniform array")
        if rc != 0:
            raise RuntimeError("oracle_run: uniform buffer exhausted")
        out = (msteps[:rows], m[:rows], dd.value, pos.value)
        if want_counts:
            out = out + (counts[:2 * nsw].reshape(nsw, 2),)
        return out


class Reference:
    """The reference's chll/chll_mod.f90 compiled here from source by
    oracle/Makefile (flang, serial).  ``stream=True`` loads the build whose
    random_number is replaced by the documented uniform stream."""

    def __init__(self, stream: bool = True):
        path = REF_STREAM_SO if stream else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.stream = stream
        self.lib = ct.CDLL(path)
        L = self.lib
        L.init.argtypes = [_f64p, _f64p, _f64p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.init.restype = None
        L.run.argtypes = [_i32p, _i32p, _i32p, _f64p, _f64p, _f64p, _i32p, _i32p, _i32p, _i32p,
                          _f64p, _f64p, _f64p, _f64p, _i32p, _i32p, _i32p, _i32p, _i32p, _f64p]
        L.run.restype = None
        if stream:
            L.chll_ref_set_stream.argtypes = [ct.c_uint64, ct.c_uint64, ct.c_int64]
            L.chll_ref_set_buffer.argtypes = [_f64p, ct.c_int64, ct.c_uint64, ct.c_int64]
            L.chll_ref_get_pos.restype = ct.c_uint64
            L.chll_ref_overflow.restype = ct.c_int

    def init(self, st):
        ni, nj, nk = st["nx"].shape
        nsub = (ni * nj * nk) // 2
        sl1 = np.zeros(nsub, dtype=np.int32)
        sl2 = np.zeros(nsub, dtype=np.int32)
        a = [ct.c_int32(v) for v in (ni, nj, nk, nsub)]
        self.lib.init(_ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                      _ptr(st["s"], _i32p), _ptr(st["dope"], _i32p), _ptr(sl1, _i32p), _ptr(sl2, _i32p),
                      *[ct.byref(x) for x in a])
        return sl1, sl2

    def run(self, st, nstart, nstop, nout, g, KK, d, kbt, *, seed=0, stream_pos=0, buf=None, quiet=True):
        """Calls the reference run(); returns (msteps, m, d_out, stream_pos_out)."""
        ni, nj, nk = st["nx"].shape
        nsub = (ni * nj * nk) // 2
        sl1, sl2 = self.init(st)
        rows = max((1 + nstop - nstart) // nout, 0)
        msteps = np.zeros(max(rows, 1), dtype=np.int32)
        m = np.zeros((max(rows, 1), 5), dtype=np.float64, order="F")
        if self.stream:
            if buf is not None:
                self._buf = np.ascontiguousarray(buf, dtype=np.float64)
                self.lib.chll_ref_set_buffer(_ptr(self._buf, _f64p), self._buf.size, stream_pos, nsub)
            else:
                self.lib.chll_ref_set_stream(seed, stream_pos, nsub)
        sc = [ct.c_int32(v) for v in (nstart, nstop, nout)]
        dims = [ct.c_int32(v) for v in (ni, nj, nk, nsub)]
        gg, kk, dd, tt = ct.c_double(g), ct.c_double(KK), ct.c_double(d), ct.c_double(kbt)
        # rows of m are only addressed when m has the (rows, 5) shape the reference assumes
        mm = m if rows > 0 else np.zeros((1, 5), order="F")
        with _silenced_stdout(quiet):
            self.lib.run(ct.byref(sc[0]), ct.byref(sc[1]), ct.byref(sc[2]),
                         _ptr(st["nx"], _f64p), _ptr(st["ny"], _f64p), _ptr(st["nz"], _f64p),
                         _ptr(st["s"], _i32p), _ptr(st["dope"], _i32p), _ptr(sl1, _i32p), _ptr(sl2, _i32p),
                         ct.byref(gg), ct.byref(kk), ct.byref(dd), ct.byref(tt),
                         *[ct.byref(x) for x in dims], _ptr(msteps, _i32p), _ptr(mm, _f64p))
        pos = self.lib.chll_ref_get_pos() if self.stream else 0
        if self.stream and self.lib.chll_ref_overflow():
            raise RuntimeError("reference: uniform buffer exhausted")
        return msteps[:rows], mm[:rows], dd.value, pos


class _silenced_stdout:
    """Redirect fd 1 to /dev/null around a native call that prints."""

    def __init__(self, on=True):
        self.on = on

    def __enter__(self):
        if self.on:
            import sys
            sys.stdout.flush()
            self.saved = os.dup(1)
            dn = os.open(os.devnull, os.O_WRONLY)
            os.dup2(dn, 1)
            os.close(dn)
        return self

    def __exit__(self, *exc):
        if self.on:
            os.dup2(self.saved, 1)
            os.close(self.saved)
        return False


def isotropic_state(ni, nj, nk, seed, p_flip=0.5):
    """Seeded isotropic unit directors and +-1 spins, Fortran-ordered.

    Directors follow the reference's initialiser formula (scripts/pdanneal.py:5-16:
    costh = 2(u1-0.5), sinth = sqrt(1-costh^2), phi = 2 pi u2) and spins
    chll/chll.py:84-90 (s = +1, then -1 when u < 0.5), vectorised with numpy's
    legacy MT19937 in the reference's i-major draw order."""
    rs = np.random.RandomState(seed)
    u = rs.random_sample((ni, nj, nk, 2))
    costh = 2.0 * (u[..., 0] - 0.5)
    sinth = np.sqrt(1.0 - costh * costh)
    phi = u[..., 1] * 2.0 * np.pi
    st = {
        "nx": np.asfortranarray(sinth * np.cos(phi)),
        "ny": np.asfortranarray(sinth * np.sin(phi)),
        "nz": np.asfortranarray(costh),
    }
    us = rs.random_sample((ni, nj, nk))
    st["s"] = np.asfortranarray(np.where(us < p_flip, -1, 1).astype(np.int32))
    st["dope"] = np.zeros((ni, nj, nk), dtype=np.int32, order="F")
    return st


def copy_state(st):
    return {k: np.asfortranarray(v.copy()) for k, v in st.items()}
