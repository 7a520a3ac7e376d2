"""``ChLLSim``: the host-side simulation class of mdeutsc2/ChLLModel
(chll/chll.py:38-304) on top of the MI355X library.

Same constructor, methods, array layouts (Fortran-ordered float64 directors,
int32 spins), output folders and file formats as the reference; the two native
calls go through the same ctypes signatures (``ftn_init``, ``ftn_run``,
chll/chll.py:9-36) into libchll_mod.so built from this package's HIP sources.

Keyword-only extensions (defaults reproduce the reference): ``bc`` ('open' |
'periodic'), ``arith`` ('compat' | 'fp64'), ``rng`` ('philox' | 'stream'),
``seed`` (Philox key / stream seed) and ``verbose``.

Deliberate differences from the reference, each a crash or a silent loss
there: ``save_params`` is skipped when ``name`` is None (the reference raises
TypeError, chll/chll.py:141); the initialisers are vectorised with numpy in
the reference's draw order instead of Python triple loops.
"""
from __future__ import annotations

import ctypes as ct
import os
import time

import numpy as np

from . import _lib
from ._lib import ARITH, BC, RNG, ROW_NAMES

libchll_mod = _lib.lib()
ftn_init = libchll_mod.init
ftn_run = libchll_mod.run


class ChLLSim:
    def __init__(self, ni, nj, nk, kbt, d, KK, g, rho, name=None, overwrite=False, *,
                 bc="open", arith="compat", rng="philox", seed=None, verbose=True):
        # parameters (chll/chll.py:39-49)
        self.name = name
        self.ni = ni
        self.nj = nj
        self.nk = nk
        self.kbt = kbt
        self.d = d
        self.KK = KK
        self.g = g
        self.rho = rho
        # internal parameters (:51-54)
        self.overwrite = overwrite
        self.simpath = None
        self.nsub = int((self.ni * self.nj * self.nk) / 2)
        # options of this implementation
        if bc not in BC or arith not in ARITH or rng not in ("philox", "stream"):
            raise ValueError(f"bad option bc={bc!r} arith={arith!r} rng={rng!r}")
        self.bc, self.arith, self.rng, self.seed, self.verbose = bc, arith, rng, seed, verbose
        self.d_out = None  # adapted step returned by the last run (the reference drops it, chll.py:170)
        self.ext = None    # full measurement rows of the last run (see ROW_NAMES)
        # arrays, Fortran-ordered (:56-82)
        self.nx = np.zeros((self.ni, self.nj, self.nk), dtype=np.float64, order="F")
        self.ny = np.zeros((self.ni, self.nj, self.nk), dtype=np.float64, order="F")
        self.nz = np.zeros((self.ni, self.nj, self.nk), dtype=np.float64, order="F")
        self.s = np.zeros((self.ni, self.nj, self.nk), dtype=np.int32, order="F")
        self.dope = np.zeros((self.ni, self.nj, self.nk), dtype=np.int32, order="F")
        self.sl1 = np.zeros(self.nsub, dtype=np.int32, order="F")
        self.sl2 = np.zeros(self.nsub, dtype=np.int32, order="F")
        self.m = None
        self.msteps = None
        self.setup_folders()

    # -- initial states (chll/chll.py:84-115) ---------------------------------
    def _default_directors(self):
        if np.mean(np.sqrt(self.nx * self.nx + self.ny * self.ny + self.nz * self.nz)) == 0.0:
            print("> Nx,Ny,Nz not initialized, setting nx[:,:,:] = 1.0")
            self.nx[:, :, :] = 1.0
            self.ny[:, :, :] = 0.0
            self.nz[:, :, :] = 0.0

    def _ftn_init(self):
        ftn_init(self.nx, self.ny, self.nz, self.s, self.dope, self.sl1, self.sl2,
                 ct.c_int(self.ni), ct.c_int(self.nj), ct.c_int(self.nk), ct.c_int(self.nsub))
        err = _lib.last_error()
        if err:
            raise _lib.ChllError(err)
        print("> Init Done!")

    def init_rand_s(self):
        """Spins +1, then -1 where np.random.rand() < 0.5, one global-stream draw
        per site in i-major order (chll/chll.py:85-90)."""
        u = np.random.rand(self.ni, self.nj, self.nk)
        self.s[:, :, :] = np.where(u < 0.5, -1, 1).astype(np.int32)
        self._default_directors()
        self._ftn_init()

    def init_with_s(self, init_s):
        self.s[:, :, :] = init_s
        self._default_directors()
        self._ftn_init()

    # -- output folders and files (chll/chll.py:117-262) ----------------------
    def setup_folders(self):
        """Claim the output folder ./data/<name> (contract of chll/chll.py:117-138):
        scripts run from the project root, ./data must already exist, an
        existing folder is reused only with overwrite=True, and either failure
        ends the program with exit(); name None saves nothing."""
        self.simpath = None
        if self.name is None:
            print("> Not saving sim")
            return
        root = "./data"
        if not os.path.isdir(root):
            print(f"ERROR: '{root}' does not exist.")
            exit()
        target = os.path.join(root, self.name)
        existed = os.path.exists(target)
        if existed and not self.overwrite:
            print(f"> ERROR: '{target}' exists and overwrite=False")
            exit()
        os.makedirs(target, exist_ok=True)
        print(f"> {'reusing' if existed else 'created'} {target}")
        self.simpath = target

    def _output_path(self, filename):
        """Where an output file goes, or None when the simulation is not saved.

        The rules are the reference's (chll/chll.py:200-262): the simulation
        folder must still exist and a file already there is replaced only with
        overwrite=True; a breach of either ends the program with exit()."""
        if self.name is None:
            return None
        if not os.path.exists(self.simpath):
            print(f"> ERROR: '{self.simpath}' does not exist")
            exit()
        path = os.path.join(self.simpath, filename)
        if os.path.isfile(path) and not self.overwrite:
            print(f"> ERROR: '{path}' exists and overwrite=False")
            exit()
        return path

    def save_params(self, nsteps, nout):
        """<name><name>_params.txt, always rewritten (chll/chll.py:139-150)."""
        if self.name is None:
            return
        fields = [("nsteps", nsteps), ("nout", nout)]
        lines = [self.name, " ".join(f"{k} = {v}" for k, v in fields),
                 f"ni,nj,nk = {self.ni} {self.nj} {self.nk}"]
        lines += [f"{k} = {getattr(self, k)}" for k in ("kbt", "d", "KK", "g", "rho")]
        with open(os.path.join(self.simpath, f"{self.name}{self.name}_params.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")

    def _configure(self):
        L = _lib.lib()
        rng = RNG[self.rng]
        _lib.check(L.chll_configure(rng, ARITH[self.arith], BC[self.bc], 1 if self.verbose else 0))
        if self.seed is not None:
            if rng == RNG["philox"]:
                _lib.check(L.chll_set_seed(int(self.seed), 0))
                self.seed = None  # later runs continue the run counter
            else:
                _lib.check(L.chll_set_stream(int(self.seed), 0))
                self.seed = None

    def run(self, nsteps, nout, save=False):
        """chll/chll.py:154-178."""
        self.save_params(nsteps, nout)
        self.msteps = np.zeros((int(nsteps / nout)), dtype=np.int32, order="F")
        self.m = np.zeros((int(nsteps / nout), 5), dtype=np.float64, order="F")
        self._configure()
        print("> RUNNING")
        st = time.perf_counter()
        d = ct.c_double(self.d)
        ftn_run(ct.c_int(1), ct.c_int(nsteps), ct.c_int(nout),
                self.nx, self.ny, self.nz, self.s, self.dope, self.sl1, self.sl2,
                ct.c_double(self.g), ct.c_double(self.KK), d, ct.c_double(self.kbt),
                ct.c_int(self.ni), ct.c_int(self.nj), ct.c_int(self.nk), ct.c_int(self.nsub),
                self.msteps, self.m)
        et = time.perf_counter()
        err = _lib.last_error()
        if err:
            raise _lib.ChllError(err)
        self.d_out = d.value
        rows = max(int(nsteps / nout), 0)
        ext = np.zeros((max(rows, 1), _lib.ROW_COLS))
        got = _lib.lib().chll_last_rows(ext.ctypes.data_as(ct.POINTER(ct.c_double)), rows)
        self.ext = {k: ext[:got, i].copy() for i, k in enumerate(ROW_NAMES)}
        if save:
            self.save_macros()
            self.save_config()
        print("> DONE! ", et - st)

    def save_config(self):
        """<name>_data.npz: the final state, compressed (chll/chll.py:200-225)."""
        path = self._output_path(f"{self.name}_data.npz")
        if path is None:
            return
        np.savez_compressed(path, **{k: getattr(self, k) for k in ("nx", "ny", "nz", "s", "dope")})
        print(f"> Wrote '{path}'")

    def save_macros(self):
        """<name>.csv: one line per measurement row (chll/chll.py:227-262)."""
        path = self._output_path(f"{self.name}.csv")
        if path is None:
            return
        body = "".join(",".join([str(int(step))] + [str(v) for v in self.m[r, :5]]) + "\n"
                       for r, step in enumerate(self.msteps))
        with open(path, "w") as f:
            f.write("Step,TotalEnergy,Eexcess,Paccept,Faccept,D\n" + body)
        print(f"> Wrote '{path}'")

    def plot_config(self, path=None):
        """Picture of the mid plane j = nj // 2 (the slice chll/chll.py:264-282
        draws), saved as <simpath>/<name>.png unless `path` is given: every
        site is a headless stick along its director's (n_x, n_z) projection,
        centred on (i, k) and coloured by its chirality.  Returns the file."""
        from matplotlib.backends.backend_agg import FigureCanvasAgg
        from matplotlib.collections import LineCollection
        from matplotlib.figure import Figure

        j = self.nj // 2
        ii, kk = np.meshgrid(np.arange(self.ni, dtype=float), np.arange(self.nk, dtype=float), indexing="ij")
        centre = np.stack([ii, kk], axis=-1)
        arm = 0.4 * np.stack([self.nx[:, j, :], self.nz[:, j, :]], axis=-1)
        sticks = np.stack([centre - arm, centre + arm], axis=-2).reshape(-1, 2, 2)
        fig = Figure(figsize=(10, 10))
        FigureCanvasAgg(fig)
        ax = fig.add_subplot()
        lines = LineCollection(sticks, array=self.s[:, j, :].astype(float).ravel(), cmap="coolwarm", linewidths=1.5)
        lines.set_clim(-1.0, 1.0)
        ax.add_collection(lines)
        ax.set_xlim(-1, self.ni)
        ax.set_ylim(-1, self.nk)
        ax.set_aspect("equal")
        ax.set_xlabel("i (x)")
        ax.set_ylabel("k (z)")
        ax.set_title(f"{self.name or ''}  y = {j}")
        fig.colorbar(lines, ax=ax, label="chirality s")
        if path is None:
            if self.simpath is None:
                raise ValueError("plot_config: no output folder (name is None) and no path given")
            path = os.path.join(self.simpath, f"{self.name}.png")
        fig.savefig(path)
        return path
