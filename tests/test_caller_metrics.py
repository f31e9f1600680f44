"""Caller-supplied grid arrays (afid_metrics) and the refusal of a
factorisation that would need a row interchange, on the GPU through the C ABI.

- A cell created with the oracle's CreateGrid / InitPressureSolver arrays steps
  like the cell built from its own grid, and both match the oracle
  (rel-L2 <= 1e-11 after three RK3 steps, the bar of the path).
- The singular (0,0) pressure mode (zgttrf info > 0, where the reference aborts:
  SolvePressureCorrection.F90:128-134) is decided on the coefficients the solve
  uses: a stretched cell given a uniform grid's pressure coefficients is
  refused, a uniform cell given a stretched grid's is not.  scipy's zgttrf on
  the same (0,0) operator fixes which grid is singular.
- A wall-normal system made non-dominant through caller metrics needs a row
  interchange in dgttrf.  The engine's pivot-free LU refuses it
  (AFID_NUM_PIVOT) on the host-driven and on the device step path, in both
  cases before any field changes.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SMALL = dict(nxm=16, nym=8, nzm=8)
METRIC_ARRAYS = ("xc", "xm", "g3rc", "g3rm", "udx3c", "udx3m", "ap3ck", "ac3ck", "am3ck", "ap3sk", "ac3sk",
                 "am3sk", "ap3ssk", "ac3ssk", "am3ssk", "amphk", "acphk", "apphk", "ak1", "ak2")


def oracle_metrics(o, names=METRIC_ARRAYS):
    """The oracle's 1-based grid arrays, from index 1, as afid_metrics takes them."""
    n = {"amphk": o.nxm, "acphk": o.nxm, "apphk": o.nxm, "ak1": o.nzm, "ak2": o.nym}
    return {nm: np.array(o.arr(nm)[1:1 + n.get(nm, o.nx)]) for nm in names}


def zero_mode_info(m, n):
    """zgttrf info of the (ky,kz)=(0,0) pressure operator built from metrics m
    as SolvePressureCorrection.F90:117-128 builds it (diagonal normalised to 1)."""
    b = 1.0 / m["acphk"][:n]
    amph, apph = m["amphk"][:n] * b, m["apphk"][:n] * b
    return lapack.zgttrf(amph[1:].astype(complex), np.ones(n, complex), apph[:-1].astype(complex))[-1]


def test_oracle_metrics_step_like_own_grid():
    from afid_amd.engine import Cell, CellParams
    from oracle.oracle import Oracle
    p = dict(nxm=24, nym=16, nzm=20, istr3=6, str3=6.0, ylen=2.0, zlen=1.5, ray=1e6, pra=0.7)
    o = Oracle(nsst=3, **p)
    own = Cell(CellParams(nsst=3, **p))
    given = Cell(CellParams(nsst=3, **p), metrics=oracle_metrics(o))
    try:
        for c in (o, own, given):
            c.init_linear_noise(4, 1e-3)
        for n in range(3):
            eo, dto = o.step(n, 1, 1e-3, 1e-8, 1e-2, 1.2)
            for c in (own, given):
                eg, dtg = c.step(n, 1, 1e-3, 1e-8, 1e-2, 1.2)
                assert eg == eo and dtg == pytest.approx(dto, rel=1e-12)
        for nm in ("vx", "vy", "vz", "temp"):
            assert rel_l2(given.download(nm), o.interior(nm)) < 1e-11, nm
            assert rel_l2(given.download(nm), own.download(nm)) < 1e-13, nm
    finally:
        own.close()
        given.close()
        o.close()


def test_singular_mode_decided_on_caller_metrics():
    from afid_amd.engine import AfidError, Cell, CellParams
    from oracle.oracle import Oracle
    uni = oracle_metrics(Oracle(istr3=0, **SMALL), ("amphk", "acphk", "apphk"))
    cheb = oracle_metrics(Oracle(istr3=6, str3=6.0, **SMALL), ("amphk", "acphk", "apphk"))
    n = SMALL["nxm"]
    assert zero_mode_info(uni, n) > 0 and zero_mode_info(cheb, n) == 0   # what scipy's zgttrf says

    def pressure_solve(params, metrics):
        cell = Cell(CellParams(**params), metrics=metrics)
        try:
            cell.init_linear_noise(1)
            cell.set_substep(1, 1e-3)
            cell.CalcLocalDivergence()
            cell.SolvePressureCorrection()
        finally:
            cell.close()

    pressure_solve(dict(istr3=6, str3=6.0, **SMALL), None)        # its own grid: regular
    with pytest.raises(AfidError, match="singular"):
        pressure_solve(dict(istr3=6, str3=6.0, **SMALL), uni)      # given a singular operator
    with pytest.raises(AfidError, match="singular"):
        pressure_solve(dict(istr3=0, **SMALL), None)                # its own grid: singular
    pressure_solve(dict(istr3=0, **SMALL), cheb)                    # given a regular operator


def test_row_interchange_refused_before_fields_change():
    from afid_amd.engine import AfidError, Cell, CellParams
    p = CellParams(nsst=3, nxm=24, nym=16, nzm=20, istr3=6, str3=6.0, ylen=2.0, zlen=1.5, ray=1e5, pra=1.0)
    dt = 1e-3
    base = Cell(p)
    try:
        ac3ck = base.metric("ac3ck", p.nxm + 1)
        ren = base.metric("ren", 1)[0]
    finally:
        base.close()
    # Row kc = 6 of the vx system: 1 - ac3ck*betadx = 1e-5 in substep 1 (al = 8/15),
    # so its normalised sub-diagonal far exceeds the previous pivot.
    betadx = dt / ren * 0.5 * (8.0 / 15.0)
    bad = ac3ck.copy()
    bad[5] = (1.0 - 1e-5) / betadx
    for path in ("host", "device"):
        cell = Cell(p, metrics={"ac3ck": bad})
        try:
            cell.init_linear_noise(3, 1e-3)
            before = {nm: cell.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}
            with pytest.raises(AfidError) as ei:
                if path == "host":
                    cell.step(0, 0, dt)
                else:
                    cell.steps(0, 3, 0, dt)
            assert ei.value.code == 1, path      # AFID_NUM_PIVOT
            for nm, a in before.items():
                np.testing.assert_array_equal(cell.download(nm), a, err_msg=f"{path}: {nm} changed")
            assert cell.time == 0.0, path
        finally:
            cell.close()
