"""main.F90's time-step rule (main.F90:171-189) against its closed form.

With vx = vy = 0 and a uniform vz = U, CalcMaxCFL (CalcMaxCFL.F90:28-33)
is exactly U*dz (the two-point average of equal values is exact), so the dt
and error code of a step are fixed in closed form:
  idtv = 1: dt = dtmin for ntime <= 1; otherwise limitcfl/cfl, clipped at dtmax
            (a still field gives cfl = 1e-8, the floor, hence dtmax);
            errorcode 166 when that dt is below dtmin.
  idtv = 0: dt as given; errorcode 165 when cfl*dt > limitcfl.
The oracle is checked on the CPU; the engine's host-driven step (afid_step)
and its device step control (afid_steps) on the GPU.
"""
import numpy as np
import pytest

GRID = dict(nxm=16, nym=8, nzm=8, istr3=6, str3=6.0, ylen=1.0, zlen=2.0, ray=1e6, pra=1.0)
DZ = GRID["nzm"] / GRID["zlen"]
DTMIN, DTMAX, LIMIT = 1e-6, 1e-2, 1.2

# (ntime, idtv, dt given, U) -> expected (errorcode, dt)
CASES = [
    (0, 1, 1e-3, 0.3, 0, DTMIN),                              # first two steps: dtmin
    (1, 1, 1e-3, 0.3, 0, DTMIN),
    (2, 1, 1e-3, 50.0, 0, LIMIT / (50.0 * DZ)),                # CFL-limited, below dtmax
    (2, 1, 1e-3, 0.3, 0, DTMAX),                               # clipped at dtmax
    (5, 1, 1e-3, 0.0, 0, DTMAX),                               # still field: cfl floor 1e-8, dtmax
    (2, 1, 1e-3, 1e6, 166, LIMIT / (1e6 * DZ)),                # below dtmin
    (3, 0, 1e-3, 100.0, 0, 1e-3),                              # fixed dt, cfl*dt = 0.4
    (3, 0, 1e-3, 500.0, 165, 1e-3),                            # fixed dt, cfl*dt = 2.0 > limit
]
IDS = ["dtmin-0", "dtmin-1", "cfl", "dtmax", "still", "166", "fixed", "165"]


def _uniform_vz(shape, U):
    vz = np.full(shape, U)
    vz[:, :, -1] = 0.0     # row nx of vz is outside the field (kept 0, as the reference leaves it)
    return vz


def test_case_table_is_consistent():
    assert 500.0 * DZ * 1e-3 > LIMIT and 100.0 * DZ * 1e-3 < LIMIT
    assert DTMIN < LIMIT / (50.0 * DZ) < DTMAX and LIMIT / (0.3 * DZ) > DTMAX and LIMIT / (1e6 * DZ) < DTMIN


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_dt_rule(case):
    from oracle.oracle import Oracle
    ntime, idtv, dtin, U, err_x, dt_x = case
    o = Oracle(nsst=3, **GRID)
    try:
        shape = (o.nzm, o.nym, o.nx)
        o.set_interior("vz", _uniform_vz(shape, U))
        assert o.cfl() == max(1e-8, U * DZ)
        err, dt = o.step(ntime, idtv, dtin, DTMIN, DTMAX, LIMIT)
        assert err == err_x
        assert dt == dt_x
    finally:
        o.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=IDS)
@pytest.mark.parametrize("path", ["host", "device"])
def test_engine_dt_rule(case, path):
    from afid_amd.engine import Cell, CellParams
    ntime, idtv, dtin, U, err_x, dt_x = case
    cell = Cell(CellParams(nsst=3, **GRID))
    try:
        cell.upload("vz", _uniform_vz((cell.nzm, cell.nym, cell.nx), U))
        assert cell.CalcMaxCFL() == max(1e-8, U * DZ)
        if path == "host":
            err, dt = cell.step(ntime, idtv, dtin, DTMIN, DTMAX, LIMIT)
        else:
            err, dt, done = cell.steps(ntime, 1, idtv, dtin, DTMIN, DTMAX, LIMIT)
            assert done == 1
        assert err == err_x
        assert dt == pytest.approx(dt_x, rel=1e-15)
    finally:
        cell.close()
