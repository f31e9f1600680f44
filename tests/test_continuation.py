"""Continuation files (include/afid_io.h): the reference's HDF5 layout
(WriteFlowField.F90:17-45, HdfRoutines.F90:388-477) and restarts.

CPU: a field written by afid_h5_write_field has the reference's dataset name,
Fortran shape (nx, nym, nzm) and element order, checked with h5dump.  It reads
back bit for bit, and shape and missing-file errors come back as codes.
GPU: a restart from the files continues bit for bit.  With the state file
removed, the restart follows the reference: pr, histories and the stale row
start at zero, checked against the oracle at rel-L2 <= 1e-11.  The Fortran
host writes the files at a normal exit and restarts from them (nread = 1).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import demean, rel_l2

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def io_or_skip():
    from afid_amd import engine
    if not os.path.exists(engine.IO_LIB_PATH):
        pytest.skip("libafid_io.so not built")
    return engine


def test_field_file_layout(tmp_path):
    engine = io_or_skip()
    nzm, nym, nx = 3, 2, 4
    a = np.arange(nzm * nym * nx, dtype=np.float64).reshape(nzm, nym, nx) * 0.5 + 1.0
    path = str(tmp_path / "continua_temp.h5")
    engine.h5_write_field(path, "var", a)
    b = engine.h5_read_field(path, "var", a.shape)
    np.testing.assert_array_equal(a, b)
    if not os.path.exists(H5DUMP):
        pytest.skip("h5dump not available")
    out = subprocess.run([H5DUMP, path], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "var"' in out and "H5T_IEEE_F64LE" in out
    # HDF5 lists dims slowest first: Fortran (nx, nym, nzm) appears as (nzm, nym, nx)
    assert f"SIMPLE {{ ( {nzm}, {nym}, {nx} ) / ( {nzm}, {nym}, {nx} ) }}" in out
    data = re.sub(r"\(\d+(,\d+)*\):", " ", out[out.index("DATA {") + 6:])
    vals = [float(v) for v in data.replace("}", " ").replace(",", " ").split()]
    np.testing.assert_array_equal(vals, a.ravel())   # x fastest: element (k, j, i) = a[i, j, k]


def test_field_file_errors(tmp_path):
    engine = io_or_skip()
    a = np.ones((2, 3, 4))
    path = str(tmp_path / "f.h5")
    engine.h5_write_field(path, "var", a)
    with pytest.raises(engine.AfidError, match="another shape"):
        engine.h5_read_field(path, "var", (2, 3, 5))
    with pytest.raises(engine.AfidError, match="could not be read"):
        engine.h5_read_field(str(tmp_path / "missing.h5"), "var", a.shape)


CASE = dict(nxm=24, nym=16, nzm=20, istr3=4, str3=3.0, ylen=2.0, zlen=1.5, ray=1e6, pra=1.0)


@pytest.mark.gpu
def test_restart_continues_bit_for_bit(tmp_path):
    engine = io_or_skip()
    a = engine.Cell(engine.CellParams(**CASE))
    a.init_linear_noise(6, 1e-3)
    for n in range(3):
        a.step(n, 1, 2e-3, 1e-4, 5e-3, 1.2)
    a.WriteFlowField(str(tmp_path))
    for f in ("continua_master.h5", "continua_temp.h5", "continua_vx.h5", "continua_vy.h5", "continua_vz.h5",
              "continua_state.h5"):
        assert os.path.exists(tmp_path / f), f
    b = engine.Cell(engine.CellParams(**CASE))
    b.ReadFlowField(str(tmp_path))
    assert b.time == a.time
    for n in range(3, 6):
        ea, da = a.step(n, 1, 2e-3, 1e-4, 5e-3, 1.2)
        eb, db = b.step(n, 1, 2e-3, 1e-4, 5e-3, 1.2)
        assert (ea, da) == (eb, db)
    for nm in ("vx", "vy", "vz", "temp", "pr"):
        np.testing.assert_array_equal(a.download(nm), b.download(nm), err_msg=nm)
    a.close()
    b.close()


@pytest.mark.gpu
def test_reference_style_restart_matches_oracle(tmp_path):
    """Without continua_state.h5 the restart starts from pr = 0, zero histories
    and a zero stale row, as the reference's ReadFlowField leaves them."""
    engine = io_or_skip()
    from oracle.oracle import Oracle
    a = engine.Cell(engine.CellParams(**CASE))
    a.init_linear_noise(8, 1e-3)
    for n in range(2):
        a.step(n, 0, 2e-3)
    a.WriteFlowField(str(tmp_path))
    os.remove(tmp_path / "continua_state.h5")
    b = engine.Cell(engine.CellParams(**CASE))
    b.ReadFlowField(str(tmp_path))
    o = Oracle(nsst=3, alx3=1.0, **CASE)
    for nm in ("vx", "vy", "vz", "temp"):
        o.set_interior(nm, engine.h5_read_field(str(tmp_path / f"continua_{nm}.h5"), "var",
                                                (CASE["nzm"], CASE["nym"], CASE["nxm"] + 1)))
    np.testing.assert_array_equal(b.download("pr"), 0.0)
    for n in range(2):
        o.step(n, 0, 2e-3, 1e-8, 1e-2, 1.2)
        b.step(n, 0, 2e-3)
    for nm in ("vx", "vy", "vz", "temp"):
        assert rel_l2(b.download(nm), o.interior(nm)) < 1e-11, nm
    nxm = CASE["nxm"]
    assert rel_l2(demean(b.download("pr")[:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm])) < 1e-11
    a.close()
    b.close()


@pytest.mark.gpu
def test_continuation_grid_that_cannot_be_interpolated_is_refused(tmp_path):
    """gridnew (interp.F90:254-299) handles the stretchings 0 and 6 only, and
    ReadFlowField.F90:72-82 stops at more than twice the old resolution."""
    engine = io_or_skip()
    a = engine.Cell(engine.CellParams(**CASE))        # istr3 = 4
    a.init_linear_noise(1, 1e-3)
    a.WriteFlowField(str(tmp_path))
    b = engine.Cell(engine.CellParams(**dict(CASE, nym=18)))
    with pytest.raises(engine.AfidError, match="cannot be interpolated"):
        b.ReadFlowField(str(tmp_path))
    a.close()
    b.close()
    a = engine.Cell(engine.CellParams(**dict(CASE, istr3=6, str3=4.0)))
    a.init_linear_noise(1, 1e-3)
    a.WriteFlowField(str(tmp_path))
    b = engine.Cell(engine.CellParams(**dict(CASE, istr3=6, str3=4.0, nzm=2 * CASE["nzm"] + 10)))
    with pytest.raises(engine.AfidError, match="cannot be interpolated"):
        b.ReadFlowField(str(tmp_path))
    a.close()
    b.close()


@pytest.mark.gpu
def test_fortran_host_restart(tmp_path):
    """afid_run writes the continuation at its normal exit (errorcode 555) and
    restarts from it with nread = 1; fixed dt, so 2 + 2 steps equal 4 steps."""
    from test_fortran_host import RUN, write_bou
    if not os.path.exists(RUN):
        pytest.skip("afid_run not built")
    one, two = tmp_path / "one", tmp_path / "two"
    one.mkdir()
    two.mkdir()
    write_bou(one / "bou.in", ntst=3, idtv=0, dt=2e-3)
    r = subprocess.run([RUN, "bou.in", "--seed", "5", "--out", "."], cwd=one, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and "errorcode 555" in r.stdout, r.stdout + r.stderr
    write_bou(two / "bou.in", ntst=1, idtv=0, dt=2e-3)
    r = subprocess.run([RUN, "bou.in", "--seed", "5"], cwd=two, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    text = (two / "bou.in").read_text().splitlines()
    text[1] = text[1].rsplit(None, 1)[0] + " 1"   # NREAD = 1
    (two / "bou.in").write_text("\n".join(text) + "\n")
    r = subprocess.run([RUN, "bou.in", "--out", "."], cwd=two, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "errorcode 555" in r.stdout, r.stdout + r.stderr
    for nm in ("vx", "vy", "vz", "temp", "pr"):
        np.testing.assert_array_equal(np.fromfile(one / f"{nm}.bin"), np.fromfile(two / f"{nm}.bin"), err_msg=nm)
