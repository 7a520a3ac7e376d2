"""chllmodel_amd: MI355X-native chiral Lebwohl-Lasher Monte Carlo.

A drop-in replacement for mdeutsc2/ChLLModel's ``libchll_mod.so`` (same C-ABI
``init``/``run``) built from hand-written HIP for gfx950, with the reference's
``ChLLSim`` host class (``chllmodel_amd.chll``) and a device-resident
``Lattice`` API (``chllmodel_amd.lattice``).
"""
from ._lib import ARITH, BC, RNG, ROW_NAMES, ChllError, LIB_PATH  # noqa: F401

__all__ = ["ARITH", "BC", "RNG", "ROW_NAMES", "ChllError", "LIB_PATH"]
