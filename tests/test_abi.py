"""The C-ABI library loads and exports every entry point include/afid.h declares.
No compute calls (runs without a GPU)."""
import ctypes
import os
import re

import pytest

from afid_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_declares_reference_entry_points():
    names = engine.exported_symbols()
    for n in ("afid_create", "afid_destroy", "afid_explicit_terms_vx", "afid_implicit_and_update_temp",
              "afid_solve_pressure_correction", "afid_correct_velocity", "afid_correct_pressure",
              "afid_time_marcher", "afid_step", "afid_calc_max_cfl", "afid_check_divergence",
              "afid_calc_plate_nu", "afid_global_quantities", "afid_calc_dissipation_nu"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libafid_hip.so not built")
    lib = ctypes.CDLL(engine.LIB_PATH)
    missing = [n for n in engine.exported_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_io_library_exports_every_declared_symbol():
    if not os.path.exists(engine.IO_LIB_PATH):
        pytest.skip("libafid_io.so not built")
    lib = ctypes.CDLL(engine.IO_LIB_PATH)
    names = engine.exported_symbols("afid_io.h")
    assert {"afid_write_continuation", "afid_read_continuation"} <= set(names)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_without_device():
    if not os.path.exists(engine.LIB_PATH):
        pytest.skip("libafid_hip.so not built")
    L = engine.load_library()
    assert L.afid_strerror(0) == b"ok"
    assert b"2, 3 and 5" in L.afid_strerror(-4)


def test_no_oracle_in_product():
    """The product package and library never reference the oracle."""
    pkg = os.path.join(ROOT, "afid_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".F90")):
                src = open(os.path.join(dirpath, f), errors="ignore").read()
                assert not re.search(r"\boracle\b|afo_|liboracle", src), f


def test_fortran_binding_covers_the_abi():
    """Every entry point of include/afid.h and include/afid_io.h has a bind(C)
    interface in the Fortran module, except the profiling and test hooks."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(engine.LIB_PATH), "..", "fortran", "afid_hip.F90")).read()
    bound = set(re.findall(r"name='(afid_\w+)'", src))
    hooks = {"afid_profile_enable", "afid_profile_read", "afid_profile_reset", "afid_pressure_spectrum",
             "afid_h5_write_field", "afid_h5_read_field"}
    declared = set(engine.exported_symbols()) | set(engine.exported_symbols("afid_io.h"))
    assert sorted(declared - bound - hooks) == []
    assert bound <= declared
