"""Mixed-precision mode (AFID_PREC_MIXED): the explicit terms computed in FP32
from the FP64 fields and their histories stored in FP32, an FP32 FFT spectrum
in the pressure solve, FP64 elsewhere (right-hand sides, the wall-normal
factorisations and solves, the reductions), and one FP64-residual refinement
of the projection per step.

Stated error against the FP64 path on the same seeded input: after 12 steps of
a small cell started from rest the per-field rel-L2 difference is below 2e-5
for vx, vy, vz (measured 7.8e-6: the vx history carries the hydrostatic
buoyancy, whose FP32 rounding is large next to the still tiny velocities),
below 1e-9 for temp (1.1e-10) and below 1e-6 for pr modulo its mean (7.3e-9),
and CheckDivergence stays below 1e-8.  The FP64 path is itself checked
against the oracle elsewhere, so these bounds are the error against the oracle
up to that path's 1e-11.  History fields round-trip through upload/download
as FP32 values; the per-routine ImplicitAndUpdate entry points are refused in
this mode.
"""

import numpy as np
import pytest

from conftest import demean, rel_l2

CASE = dict(nxm=24, nym=16, nzm=20, istr3=6, str3=6.0, ylen=2.0, zlen=2.0, ray=1e6, pra=0.7)


def _cell(precision, **kw):
    from afid_amd.engine import Cell, CellParams
    return Cell(CellParams(**dict(CASE, precision=precision, **kw)))


@pytest.mark.gpu
@pytest.mark.parametrize("device_steps", [False, True], ids=["host", "graph"])
def test_mixed_tracks_fp64(device_steps):
    from afid_amd.engine import PREC_FP64, PREC_MIXED
    out = {}
    for prec in (PREC_FP64, PREC_MIXED):
        c = _cell(prec)
        c.init_linear_noise(9, 1e-3)
        if device_steps:
            err, _, done = c.steps(0, 12, 1, 1e-3, 1e-8, 5e-3, 1.2)
            assert err == 0 and done == 12
        else:
            for n in range(12):
                c.step(n, 1, 1e-3, 1e-8, 5e-3, 1.2)
        out[prec] = {nm: c.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}
        out[prec]["div"] = c.CheckDivergence()
        out[prec]["time"] = c.time
        c.close()
    a, b = out[PREC_FP64], out[PREC_MIXED]
    errs = {nm: rel_l2(b[nm], a[nm]) for nm in ("vx", "vy", "vz", "temp")}
    nxm = CASE["nxm"]
    errs["pr"] = rel_l2(demean(b["pr"][:, :, :nxm]), demean(a["pr"][:, :, :nxm]))
    print("mixed vs fp64 rel-L2", errs, "div", b["div"], a["div"], flush=True)
    for nm in ("vx", "vy", "vz"):
        assert errs[nm] < 2e-5, nm
    assert errs["temp"] < 1e-9
    assert errs["pr"] < 1e-6
    assert b["div"] < 1e-8 and a["div"] < 1e-11
    assert abs(b["time"] - a["time"]) <= 1e-6 * a["time"]


@pytest.mark.gpu
def test_mixed_history_roundtrip():
    from afid_amd.engine import PREC_MIXED
    c = _cell(PREC_MIXED)
    h = np.random.default_rng(2).standard_normal((CASE["nzm"], CASE["nym"], CASE["nxm"] + 1))
    c.upload("rux", h)
    np.testing.assert_array_equal(c.download("rux"), h.astype(np.float32).astype(np.float64))
    c.close()


@pytest.mark.gpu
def test_mixed_refusals():
    from afid_amd.engine import PREC_MIXED, AfidError
    c = _cell(PREC_MIXED)
    c.init_linear_noise(1, 1e-3)
    c.set_substep(1, 1e-3)
    c.ExplicitTermsVX()
    with pytest.raises(AfidError, match="code -5"):
        c.ImplicitAndUpdateVX()
    with pytest.raises(AfidError, match="code -5"):
        c.steps_cfl()
    c.close()


def test_params_precision_field():
    from afid_amd.engine import CellParams, PREC_MIXED, _Params
    import ctypes
    p = CellParams(nxm=8, nym=8, nzm=8, precision=PREC_MIXED).to_c()
    assert p.precision == PREC_MIXED and ctypes.sizeof(_Params) % 8 == 0


@pytest.mark.gpu
@pytest.mark.parametrize("c", [dict(nxm=12, nym=768, nzm=6), dict(nxm=10, nym=6, nzm=750)], ids=["y768", "z750"])
def test_mixed_long_fft_lines(c):
    """The FP32 FFT instances on a y line of 768 and a z line of 750: mixed
    against FP64 after 8 steps of dt 5e-4 (velocities < 1e-4, measured 1.4e-5;
    temp < 1e-7 rel-L2, measured 1.3e-8; max|div| < 1e-8)."""
    from afid_amd.engine import Cell, CellParams, PREC_FP64, PREC_MIXED
    out = {}
    for prec in (PREC_FP64, PREC_MIXED):
        cell = Cell(CellParams(istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0, precision=prec, **c))
        cell.init_linear_noise(5, 1e-3)
        err, _, done = cell.steps(0, 8, 0, 5e-4)
        assert err == 0 and done == 8
        out[prec] = {nm: cell.download(nm) for nm in ("vx", "vy", "vz", "temp")}
        out[prec]["div"] = cell.CheckDivergence()
        cell.close()
    errs = {nm: rel_l2(out[PREC_MIXED][nm], out[PREC_FP64][nm]) for nm in ("vx", "vy", "vz", "temp")}
    print("long-line mixed vs fp64", c, errs, out[PREC_MIXED]["div"], flush=True)
    for nm in ("vx", "vy", "vz"):
        assert errs[nm] < 1e-4, nm
    assert errs["temp"] < 1e-7 and out[PREC_MIXED]["div"] < 1e-8
