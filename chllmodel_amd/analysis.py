"""Lattice analysis on the GPU: correlation functions, the chirality structure
factor, and VTK export.

Drop-in replacements, under the reference's names and return values, for the
analysis helpers of mdeutsc2/ChLLModel's notebooks and scripts:

* ``calc_n_corr(nx, ny, nz, nshells) -> (g, c)`` and
  ``calc_s_corr(s, nshells) -> (g, c)`` (notebooks/paper_plots2.ipynb, cell 0):
  shell-averaged nematic correlation <(n_r . n_{r+d})^2> and chiral correlation
  <s_r s_{r+d}> - <s>^2 over |d| < nshells, shell floor(|d|), open edges.
  ``fit_corr_exp`` / ``fit_corr_exp2`` are the notebook's exponential fits.
* ``structure_factor(data, bins, k_min, k_max, plane) -> (sf, kvs)``
  (notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:10-29):
  |sum_r s_r exp(-i kv plane . (r + 1))|^2 / N^2 on np.linspace(k_min, k_max, bins).
* ``convert_npz(fn)`` (viz/np2vti.py:8-32): the ChLLSim ``*_data.npz`` state as
  a VTK XML image (``.vti``) with point arrays ``s`` (Int32) and ``n``
  (Float32 x 3), written directly (no vtk module needed).

The sums run in libchll_mod.so (``chll_shell_corr``, ``chll_structure_factor``);
arrays are (ni, nj, nk) indexed [i, j, k] like the reference's, as numpy arrays
or torch tensors.  There is no CPU fallback: without the library or a GPU the
functions raise.
"""
from __future__ import annotations

import ctypes as ct
import os
import sys

import numpy as np

from ._lib import lib

_vp = ct.c_void_p


def _declare(L):
    if getattr(L, "_analysis_declared", False):
        return L
    i32 = ct.c_int32
    L.chll_analysis_last_error.restype = ct.c_char_p
    L.chll_corr_workspace_bytes.argtypes = [i32] * 4
    L.chll_corr_workspace_bytes.restype = ct.c_size_t
    L.chll_corr_counts.argtypes = [i32] * 4 + [ct.POINTER(ct.c_int64)]
    L.chll_corr_counts.restype = ct.c_int
    L.chll_shell_corr.argtypes = [_vp] * 4 + [i32] * 4 + [_vp, _vp, _vp, ct.c_size_t, _vp]
    L.chll_shell_corr.restype = ct.c_int
    L.chll_sf_workspace_bytes.argtypes = [i32, i32, i32, ct.POINTER(ct.c_int32), i32, i32]
    L.chll_sf_workspace_bytes.restype = ct.c_size_t
    L.chll_structure_factor.argtypes = [_vp, i32, i32, i32, ct.POINTER(ct.c_int32), i32,
                                        ct.POINTER(ct.c_double), i32, _vp, _vp, ct.c_size_t, _vp]
    L.chll_structure_factor.restype = ct.c_int
    L._analysis_declared = True
    return L


def _fail(L, what):
    raise RuntimeError(f"{what}: {L.chll_analysis_last_error().decode()}")


def _shape3(a):
    if a.ndim != 3:
        raise ValueError("lattice arrays must have shape (ni, nj, nk)")
    return tuple(int(x) for x in a.shape)


def _device(a, dtype):
    """(ni, nj, nk) array -> flat Fortran-order torch tensor on the GPU."""
    import torch

    if isinstance(a, torch.Tensor):
        t = a.to(device="cuda", dtype=dtype)
        return t.permute(2, 1, 0).contiguous().view(-1)
    arr = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(arr.transpose(2, 1, 0)).astype(
        {torch.float64: np.float64, torch.int8: np.int8}[dtype], copy=False))
    return t.reshape(-1).to("cuda")


def _spins_i8(s):
    import torch

    if isinstance(s, torch.Tensor):
        lo, hi = int(s.min()), int(s.max())
    else:
        s = np.asarray(s)
        lo, hi = int(s.min()), int(s.max())
    if lo < -127 or hi > 127:
        raise ValueError("spin values must lie in [-127, 127]")
    return _device(s, torch.int8)


def corr_counts(shape, nshells: int) -> np.ndarray:
    """Pair counts per shell (the reference's second return value), exact."""
    L = _declare(lib())
    ni, nj, nk = shape
    out = np.zeros(nshells, dtype=np.int64)
    if L.chll_corr_counts(ni, nj, nk, nshells, out.ctypes.data_as(ct.POINTER(ct.c_int64))) != 0:
        _fail(L, "chll_corr_counts")
    return out


def prepare_directors(nx, ny, nz):
    """(ni, nj, nk) director components -> three flat Fortran-order float64
    tensors on the GPU (the layout the kernels read)."""
    import torch

    shape = _shape3(nx)
    if tuple(ny.shape) != shape or tuple(nz.shape) != shape:
        raise ValueError("arrays must share one shape")
    return [_device(a, torch.float64) for a in (nx, ny, nz)]


def prepare_spins(s):
    """(ni, nj, nk) spins -> flat Fortran-order int8 tensor on the GPU."""
    _shape3(s)
    return _spins_i8(s)


def shell_correlations_device(shape, nshells, dn=None, ds=None, workspace=None):
    """Device-resident form: dn = prepare_directors(...) or None, ds =
    prepare_spins(...) or None.  Returns (g_nem, g_chi) as GPU tensors (None
    where not asked), enqueued on torch's current stream."""
    import torch

    L = _declare(lib())
    if dn is None and ds is None:
        raise ValueError("pass the directors, the spins, or both")
    ni, nj, nk = shape
    N = ni * nj * nk
    for t in (dn or []) + ([ds] if ds is not None else []):
        if t.numel() != N or not t.is_cuda:
            raise ValueError("device arrays must hold ni*nj*nk elements on the GPU")
    gn = torch.empty(nshells, dtype=torch.float64, device="cuda") if dn is not None else None
    gs = torch.empty(nshells, dtype=torch.float64, device="cuda") if ds is not None else None
    wsb = L.chll_corr_workspace_bytes(ni, nj, nk, nshells)
    if wsb == 0:
        raise ValueError("nshells must be in 1..127 and the shape positive")
    ws = workspace if workspace is not None and workspace.numel() >= wsb else \
        torch.empty(wsb, dtype=torch.uint8, device="cuda")
    ptr = lambda t: _vp(t.data_ptr()) if t is not None else None  # noqa: E731
    dx = dn if dn is not None else [None] * 3
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.chll_shell_corr(ptr(dx[0]), ptr(dx[1]), ptr(dx[2]), ptr(ds), ni, nj, nk, nshells, ptr(gn), ptr(gs),
                           ptr(ws), wsb, _vp(stream))
    if rc != 0:
        _fail(L, "chll_shell_corr")
    return gn, gs


def shell_correlations(nx=None, ny=None, nz=None, s=None, *, nshells: int):
    """(g_nem or None, g_chi or None, counts as float64) in one kernel pass."""
    nem = nx is not None
    chi = s is not None
    if not nem and not chi:
        raise ValueError("pass the directors, the spins, or both")
    shape = _shape3(nx if nem else s)
    if chi and tuple(s.shape) != shape:
        raise ValueError("arrays must share one shape")
    dn = prepare_directors(nx, ny, nz) if nem else None
    ds = prepare_spins(s) if chi else None
    gn, gs = shell_correlations_device(shape, nshells, dn, ds)
    c = corr_counts(shape, nshells).astype(np.float64)
    return (gn.cpu().numpy() if nem else None), (gs.cpu().numpy() if chi else None), c


def calc_n_corr(nx, ny, nz, nshells):
    """Nematic shell correlation (notebooks/paper_plots2.ipynb, cell 0)."""
    g, _, c = shell_correlations(nx, ny, nz, nshells=int(nshells))
    return g, c


def calc_s_corr(s, nshells):
    """Chiral shell correlation with <s>^2 subtracted (same cell)."""
    _, g, c = shell_correlations(s=s, nshells=int(nshells))
    return g, c


def _corr_model(x, a, l):
    return a * np.exp(-x / l)


def _corr_model2(x, a, l, c):
    return a * np.exp(-x / l) + c


def fit_corr_exp(g):
    """a exp(-r / l) fitted to the shells; returns (a, l)."""
    import scipy.optimize

    popt, _ = scipy.optimize.curve_fit(_corr_model, np.arange(len(g)), g)
    return popt[0], popt[1]


def fit_corr_exp2(g):
    """a exp(-r / l) + c fitted to the shells; returns (a, l, c)."""
    import scipy.optimize

    popt, _ = scipy.optimize.curve_fit(_corr_model2, np.arange(len(g)), g)
    return popt[0], popt[1], popt[2]


def structure_factors_device(ds, shape, kvs, planes):
    """Device-resident form: ds = prepare_spins(...).  Returns an array
    (P, bins) on the GPU; planes are grouped so that each call's projection
    histograms fit."""
    import torch

    L = _declare(lib())
    ni, nj, nk = shape
    if ds.numel() != ni * nj * nk or ds.dtype != torch.int8 or not ds.is_cuda:
        raise ValueError("ds must be the flat int8 spins on the GPU")
    kv = np.ascontiguousarray(np.asarray(kvs, dtype=np.float64))
    pl = np.ascontiguousarray(np.asarray(planes, dtype=np.int32).reshape(-1, 3))
    out = torch.empty((len(pl), len(kv)), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    p0 = 0
    while p0 < len(pl):
        p1 = p0 + 1
        while p1 < len(pl) and p1 - p0 < 16 and _sf_ws(L, shape, pl[p0:p1 + 1], len(kv)) > 0:
            p1 += 1
        group = np.ascontiguousarray(pl[p0:p1])
        wsb = _sf_ws(L, shape, group, len(kv))
        if wsb == 0:
            _fail(L, f"plane {pl[p0].tolist()}")
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        rc = L.chll_structure_factor(_vp(ds.data_ptr()), ni, nj, nk, group.ctypes.data_as(ct.POINTER(ct.c_int32)),
                                     len(group), kv.ctypes.data_as(ct.POINTER(ct.c_double)), len(kv),
                                     _vp(out[p0:p1].data_ptr()), _vp(ws.data_ptr()), wsb, _vp(stream))
        if rc != 0:
            _fail(L, "chll_structure_factor")
        p0 = p1
    return out


def structure_factors(data, kvs, planes) -> np.ndarray:
    """S(kv) for every plane vector in `planes` and every kv: array (P, bins)."""
    shape = _shape3(data)
    return structure_factors_device(prepare_spins(data), shape, kvs, planes).cpu().numpy()


def _sf_ws(L, shape, planes, bins):
    planes = np.ascontiguousarray(planes, dtype=np.int32)
    return L.chll_sf_workspace_bytes(*shape, planes.ctypes.data_as(ct.POINTER(ct.c_int32)), len(planes), bins)


def structure_factor(data, bins, k_min, k_max, plane):
    """(sf, kvs) with kvs = np.linspace(k_min, k_max, bins)
    (notebooks/.ipynb_checkpoints/hero_stripe-checkpoint.py:19-29)."""
    kvs = np.linspace(k_min, k_max, bins)
    return structure_factors(data, kvs, [plane])[0], kvs


# VTK export ---------------------------------------------------------------

def _vti_arrays(nx, ny, nz, s, layout):
    if layout == "reference":
        # viz/np2vti.py:13-22: s.flatten() (C order) and np.stack(n).reshape(-1, 3)
        sv = np.asarray(s).flatten().astype(np.int32)
        nv = np.stack((nx, ny, nz)).reshape(-1, 3).astype(np.float32)
    elif layout == "points":
        # VTK point order (x = i fastest), one director per point
        sv = np.asarray(s).ravel(order="F").astype(np.int32)
        nv = np.stack([np.asarray(a).ravel(order="F") for a in (nx, ny, nz)], axis=1).astype(np.float32)
    else:
        raise ValueError("layout must be 'reference' or 'points'")
    return np.ascontiguousarray(sv), np.ascontiguousarray(nv)


def write_vti(path, nx, ny, nz, s, layout: str = "reference") -> None:
    """VTK XML ImageData, dimensions nx.shape, point arrays "s" (Int32) and
    "n" (Float32, 3 components), raw appended little-endian data with UInt64
    block headers.  layout="reference" stores the arrays in the order
    viz/np2vti.py hands them to VTK (C-order flatten of s; the stacked
    (3, ni, nj, nk) directors cut into consecutive triples); layout="points"
    stores one director per point in VTK's x-fastest point order."""
    ni, nj, nk = _shape3(np.asarray(nx))
    sv, nv = _vti_arrays(nx, ny, nz, s, layout)
    if sv.size != ni * nj * nk or nv.shape[0] != ni * nj * nk:
        raise ValueError("arrays must share one shape")
    ext = f"0 {ni - 1} 0 {nj - 1} 0 {nk - 1}"
    off_n = 8 + sv.nbytes
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'  <ImageData WholeExtent="{ext}" Origin="0 0 0" Spacing="1 1 1">\n'
        f'    <Piece Extent="{ext}">\n'
        '      <PointData>\n'
        '        <DataArray type="Int32" Name="s" format="appended" offset="0"/>\n'
        f'        <DataArray type="Float32" Name="n" NumberOfComponents="3" format="appended" offset="{off_n}"/>\n'
        '      </PointData>\n'
        '      <CellData>\n'
        '      </CellData>\n'
        '    </Piece>\n'
        '  </ImageData>\n'
        '  <AppendedData encoding="raw">\n   _'
    )
    with open(path, "wb") as f:
        f.write(head.encode())
        f.write(np.uint64(sv.nbytes).tobytes())
        f.write(sv.astype("<i4").tobytes())
        f.write(np.uint64(nv.nbytes).tobytes())
        f.write(nv.astype("<f4").tobytes())
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")


def read_vti(path) -> dict:
    """Reads a file written by ``write_vti``: {"dims": (ni, nj, nk), "s": int32
    array, "n": float32 (N, 3) array}."""
    import re

    raw = open(path, "rb").read()
    mark = raw.index(b'<AppendedData encoding="raw">')
    head = raw[:mark].decode()
    ext = [int(x) for x in re.search(r'WholeExtent="([^"]+)"', head).group(1).split()]
    dims = (ext[1] + 1, ext[3] + 1, ext[5] + 1)
    data = raw[raw.index(b"_", mark) + 1:]
    out = {"dims": dims}
    pos = 0
    for name, dt, comps in (("s", "<i4", 1), ("n", "<f4", 3)):
        nbytes = int(np.frombuffer(data[pos:pos + 8], dtype="<u8")[0])
        arr = np.frombuffer(data[pos + 8:pos + 8 + nbytes], dtype=dt)
        out[name] = arr.reshape(-1, 3) if comps == 3 else arr
        pos += 8 + nbytes
    return out


def convert_npz(fn, layout: str = "reference") -> str:
    """<name>.npz (keys nx, ny, nz, s) -> <name>.vti beside it (viz/np2vti.py:8-32)."""
    outfile = ".".join(fn.split(".")[:-1]) + ".vti"
    data = np.load(fn)
    write_vti(outfile, data["nx"], data["ny"], data["nz"], data["s"], layout=layout)
    print(fn, " => ", outfile)
    return outfile


def main(argv=None) -> int:
    """python -m chllmodel_amd.analysis <file.npz | directory>: converts one
    state file, or every .npz below a directory, to .vti (viz/np2vti.py:35-47)."""
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m chllmodel_amd.analysis <file.npz | directory>")
        return 2
    infile = argv[0]
    if os.path.isdir(infile):
        for root, _dirs, files in os.walk(infile):
            for file in files:
                if file.endswith(".npz"):
                    convert_npz(os.path.join(root, file))
    elif infile.endswith(".npz"):
        convert_npz(infile)
    else:
        print("not npz file or directory!")
        return 1
    print("Done!")
    return 0


if __name__ == "__main__":
    sys.exit(main())
