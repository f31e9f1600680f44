"""afid_amd: MI355X engine for AFiD's per-substep time advance.

The product is libafid_hip.so (HIP kernels for gfx950 behind the C ABI of
include/afid.h); this package holds its build step, a ctypes host mirror of
the reference routine interface (engine.Cell), and the Fortran binding
(fortran/afid_hip.F90).
"""
from .engine import AfidError, Cell, CellParams, load_library  # noqa: F401

__all__ = ["AfidError", "Cell", "CellParams", "load_library"]
