"""afid_amd: MI355X engine for AFiD's per-substep time advance.

The product is libafid_hip.so (HIP kernels for gfx950 behind the C ABI of
include/afid.h); this package holds its build step, a ctypes host mirror of
the reference routine interface (engine.Cell), the Fortran binding
(fortran/afid_hip.F90), the continuation/statistics I/O library (io/), and
the batched-sweep and serving loops over many resident cells (sweep, serve).
"""
from .engine import PREC_FP64, PREC_MIXED, AfidError, Cell, CellParams, load_library  # noqa: F401

__all__ = ["AfidError", "Cell", "CellParams", "load_library", "PREC_FP64", "PREC_MIXED"]
