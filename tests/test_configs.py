"""The measurement configurations on the GPU, each against the oracle or against
what fixes its answer.

cfg1  64x128x128, tanh stretching istr3=4 str3=4, ylen=zlen=2, Ra=1e6, Pr=1,
      RK3, fixed dt=1e-3, 100 steps.  As written, the (0,0) pressure mode of this
      grid meets an exact zero pivot in netlib's zgttrf (scipy's zgttrf on the
      oracle's operator returns info = 64), so the reference aborts
      (SolvePressureCorrection.F90:128-134): the engine must refuse it with
      AFID_NUM_SINGULAR, on the per-routine, host-step and device-step paths.
      The same cell with clipped-Chebyshev stretching (istr3=6, str3=6), the
      substitution used for measurements, runs the 100 steps against the oracle:
      rel-L2 <= 1e-11 per field (pr modulo its mean), max|div| <= 1e-10, and the
      plate, volume and dissipation Nusselt numbers to rel 1e-10.
cfg3  256x512x512, istr3=6 str3=6, ylen=zlen=2, Ra=1e7, Pr=0.7, CFL dt.  FP64
      against the oracle for 3 steps (rel-L2 <= 1e-11); then the mixed-precision
      mode against FP64 after 200 steps: rel-L2 below 3e-4 for the velocities
      (measured 3.9e-5), 1e-7 for temp (2.2e-8, the explicit terms being FP32),
      1e-8 for pr modulo its mean (4.0e-10), max|div| <= 1e-8.
cfg4  The 64-cell batched sweep (sweep_specs(64, 500)): every cell completes its
      500 steps with error code 0 and max|div| <= 1e-8; the four smallest cells
      equal their runs alone bit for bit; the sweep's two smallest cells,
      batched for 5 steps, match the oracle (rel-L2 <= 1e-11).
cfg5  The 24-request serving stream (serving_requests(24), lambda = 4/s): the
      refused requests are exactly those whose (0,0) pressure mode scipy's
      zgttrf finds singular; every served cell ends with error code 0 and
      max|div| <= 1e-8; the three smallest served cells equal their runs alone
      bit for bit.
"""
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import demean, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp")
CFG1 = dict(nxm=64, nym=128, nzm=128, istr3=4, str3=4.0, ylen=2.0, zlen=2.0, ray=1e6, pra=1.0)
CFG3 = dict(nxm=256, nym=512, nzm=512, istr3=6, str3=6.0, ylen=2.0, zlen=2.0, ray=1e7, pra=0.7)
CFL_CTL = (1, 1e-3, 1e-8, 1e-2, 1.2)    # idtv, dt, dtmin, dtmax, cflmax (bou.in)


def _threads():
    from oracle import oracle as O
    O.lib().afo_set_threads(min(16, os.cpu_count() or 1))


def zero_mode_zgttrf_info(nxm, istr3, str3, alx3=1.0):
    """scipy zgttrf info of the (0,0) pressure operator of this x grid, built from
    the oracle's InitPressureSolver arrays as SolvePressureCorrection.F90:117-128
    builds it.  The (0,0) mode does not depend on the lateral sizes."""
    from oracle.oracle import Oracle
    o = Oracle(nxm=nxm, nym=4, nzm=4, istr3=istr3, str3=str3, alx3=alx3)
    try:
        n = nxm
        am, ac, ap = (np.array(o.arr(nm)[1:n + 1]) for nm in ("amphk", "acphk", "apphk"))
    finally:
        o.close()
    b = 1.0 / ac
    amph, apph = am * b, ap * b
    return lapack.zgttrf(amph[1:].astype(complex), np.ones(n, complex), apph[:-1].astype(complex))[-1]


def compare_fields(cell, o, tol, tag=""):
    nxm = o.nxm
    for nm in FIELDS:
        e = rel_l2(cell.download(nm), o.interior(nm))
        print(f"{tag}{nm}: rel-L2 {e:.2e}", flush=True)
        assert e <= tol, nm
    e = rel_l2(demean(cell.download("pr")[:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm]))
    print(f"{tag}pr (mod mean): rel-L2 {e:.2e}", flush=True)
    assert e <= tol


def compare_nusselt(cell, o, rel):
    nb, nt = cell.CalcPlateNu()
    ob, ot = o.plate_nu()
    assert nb == pytest.approx(ob, rel=rel) and nt == pytest.approx(ot, rel=rel)
    gq, ogq = cell.GlobalQuantities(), o.global_quantities()
    np.testing.assert_allclose(gq, ogq, rtol=rel)
    mu, th, _, _ = cell.CalcDissipationNu()
    omu, oth, _, _ = o.dissipation()
    assert mu == pytest.approx(omu, rel=rel) and th == pytest.approx(oth, rel=rel)
    print(f"Nu plate {nb:.6f}/{nt:.6f}, Nu vol {gq[0]:.6f}, Nu diss {mu:.6f}/{th:.6f}", flush=True)


# ---------------------------------------------------------------- cfg1
def test_cfg1_as_written_is_refused_like_the_reference():
    from afid_amd.engine import AfidError, Cell, CellParams
    assert zero_mode_zgttrf_info(CFG1["nxm"], CFG1["istr3"], CFG1["str3"]) > 0
    assert zero_mode_zgttrf_info(CFG1["nxm"], 6, 6.0) == 0            # the substitution is regular
    for path in ("routine", "host", "device"):
        cell = Cell(CellParams(nsst=3, **CFG1))
        try:
            cell.init_linear_noise(1, 1e-3)
            before = cell.download("temp")
            with pytest.raises(AfidError, match="singular") as ei:
                if path == "routine":
                    cell.set_substep(1, 1e-3)
                    cell.CalcLocalDivergence()
                    cell.SolvePressureCorrection()
                elif path == "host":
                    cell.step(0, 0, 1e-3)
                else:
                    cell.steps(0, 100, 0, 1e-3)
            assert ei.value.code == 2, path                                # AFID_NUM_SINGULAR
            if path == "device":
                np.testing.assert_array_equal(cell.download("temp"), before)
        finally:
            cell.close()


def test_cfg1_substituted_hundred_steps_match_oracle():
    from afid_amd.engine import Cell, CellParams
    from oracle.oracle import Oracle
    _threads()
    p = dict(CFG1, istr3=6, str3=6.0)
    print("cfg1 (istr3=6): oracle and GPU, 100 steps at fixed dt 1e-3", flush=True)
    o = Oracle(nsst=3, **p)
    cell = Cell(CellParams(nsst=3, **p))
    try:
        o.init_linear_noise(1, 1e-3)
        cell.init_linear_noise(1, 1e-3)
        for n in range(100):
            eo, _ = o.step(n, 0, 1e-3, 1e-8, 1e-2, 1.2)
            assert eo == 0, n
        err, dt, done = cell.steps(0, 100, 0, 1e-3, 1e-8, 1e-2, 1.2)
        assert (err, done) == (0, 100) and dt == 1e-3
        assert cell.time == pytest.approx(0.1, rel=1e-13)
        compare_fields(cell, o, 1e-11, "cfg1 ")
        assert cell.CheckDivergence() <= 1e-10 and o.divmax() <= 1e-10
        compare_nusselt(cell, o, 1e-10)
    finally:
        cell.close()
        o.close()


# ---------------------------------------------------------------- cfg3
def test_cfg3_fp64_three_steps_match_oracle():
    from afid_amd.engine import Cell, CellParams
    from oracle.oracle import Oracle
    _threads()
    print("cfg3 256x512x512 FP64: oracle and GPU, 3 steps", flush=True)
    o = Oracle(nsst=3, **CFG3)
    cell = Cell(CellParams(nsst=3, **CFG3))
    try:
        o.init_linear_noise(3, 1e-3)
        cell.init_linear_noise(3, 1e-3)
        for n in range(3):
            eo, dto = o.step(n, *CFL_CTL)
            assert eo == 0
        err, dt, done = cell.steps(0, 3, *CFL_CTL)
        assert (err, done) == (0, 3) and dt == pytest.approx(dto, rel=1e-12)
        compare_fields(cell, o, 1e-11, "cfg3 ")
        compare_nusselt(cell, o, 1e-10)
    finally:
        cell.close()
        o.close()


def test_cfg3_mixed_two_hundred_steps_track_fp64():
    from afid_amd.engine import PREC_FP64, PREC_MIXED, Cell, CellParams
    out = {}
    for prec in (PREC_FP64, PREC_MIXED):
        print(f"cfg3 precision {prec}: 200 device steps", flush=True)
        cell = Cell(CellParams(nsst=3, precision=prec, **CFG3))
        try:
            cell.init_linear_noise(3, 1e-3)
            err, _, done = cell.steps(0, 200, *CFL_CTL)
            assert (err, done) == (0, 200)
            out[prec] = {nm: cell.download(nm) for nm in FIELDS + ("pr",)}
            out[prec]["div"] = cell.max_divergence()
        finally:
            cell.close()
    a, b = out[PREC_FP64], out[PREC_MIXED]
    nxm = CFG3["nxm"]
    errs = {nm: rel_l2(b[nm], a[nm]) for nm in FIELDS}
    errs["pr"] = rel_l2(demean(b["pr"][:, :, :nxm]), demean(a["pr"][:, :, :nxm]))
    print("cfg3 mixed vs FP64 after 200 steps:", {k: f"{v:.2e}" for k, v in errs.items()},
          f"max|div| mixed {b['div']:.2e} fp64 {a['div']:.2e}", flush=True)
    for nm in ("vx", "vy", "vz"):
        assert errs[nm] < 3e-4, nm
    assert errs["temp"] < 1e-7 and errs["pr"] < 1e-8
    assert b["div"] <= 1e-8 and a["div"] <= 1e-10


# ---------------------------------------------------------------- cfg4
def test_cfg4_sweep_batched():
    from afid_amd.engine import Cell
    from afid_amd.sweep import STEP_CTL, run_batched, sweep_specs
    from oracle.oracle import Oracle
    specs = sweep_specs(64, 500, seed=0)
    order = sorted(range(len(specs)), key=lambda q: (specs[q].params.cells, q))
    alone_ids, oracle_ids = order[:4], order[:2]
    kept, divs = {}, {}

    def on_done(q, cell):
        divs[q] = cell.max_divergence()   # inf for a cell whose fields are not finite
        if q in alone_ids:
            kept[q] = {nm: cell.download(nm) for nm in FIELDS}

    print("cfg4: 64 cells x 500 steps, batched", flush=True)
    res = run_batched(specs, on_done=on_done)
    print(f"cfg4 batched: {res.seconds:.1f} s, {res.cells_per_hour:.0f} cells/hour, "
          f"max|div| {res.max_divergence:.2e}", flush=True)
    assert res.errorcodes == [0] * 64 and res.steps_done == [500] * 64
    assert max(divs.values()) <= 1e-8 and len(divs) == 64
    ctl = STEP_CTL
    for q in alone_ids:
        s = specs[q]
        c = Cell(s.params)
        try:
            c.init_linear_noise(s.seed, 1e-3)
            err, _, done = c.steps(0, s.steps, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
            assert (err, done) == (0, 500)
            for nm in FIELDS:
                np.testing.assert_array_equal(c.download(nm), kept[q][nm], err_msg=f"cell {q} {nm}")
        finally:
            c.close()
    # the same cells batched for 5 steps against the oracle
    _threads()
    short = [type(specs[q])(specs[q].params, specs[q].seed, 5) for q in oracle_ids]
    got = {}
    run_batched(short, on_done=lambda q, cell: got.__setitem__(q, {nm: cell.download(nm) for nm in FIELDS}))
    for q, s in enumerate(short):
        p = s.params
        o = Oracle(nsst=3, nxm=p.nxm, nym=p.nym, nzm=p.nzm, istr3=p.istr3, str3=p.str3, alx3=p.alx3,
                   ylen=p.ylen, zlen=p.zlen, ray=p.ray, pra=p.pra)
        try:
            o.init_linear_noise(s.seed, 1e-3)
            for n in range(5):
                o.step(n, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
            for nm in FIELDS:
                e = rel_l2(got[q][nm], o.interior(nm))
                assert e <= 1e-11, (q, nm, e)
        finally:
            o.close()


# ---------------------------------------------------------------- cfg5
def test_cfg5_serving_stream():
    from afid_amd.engine import Cell
    from afid_amd.serve import Server, serving_requests, summarize
    from afid_amd.sweep import STEP_CTL
    reqs = serving_requests(24, lam=4.0, seed=0)
    by_rid = {r.rid: r for r in reqs}
    singular = {r.rid for r in reqs
                if zero_mode_zgttrf_info(r.spec.params.nxm, r.spec.params.istr3, r.spec.params.str3) > 0}
    ran = sorted((r for r in reqs if r.rid not in singular), key=lambda r: (r.spec.cell_updates, r.rid))
    sample = {r.rid for r in ran[:3]}
    kept = {}

    def on_finish(req, cell):
        if req.rid in sample:
            kept[req.rid] = {nm: cell.download(nm) for nm in FIELDS}

    print(f"cfg5: 24 requests, {len(singular)} with a singular (0,0) mode", flush=True)
    srv = Server(budget_bytes=200e9, on_finish=on_finish)
    done = srv.run(reqs)
    s = summarize(done)
    print(f"cfg5: {s}", flush=True)
    assert sorted(c.rid for c in done) == list(range(24))
    assert {c.rid for c in done if c.refused} == singular
    for c in done:
        if not c.refused:
            assert c.errorcode == 0 and c.steps_done == by_rid[c.rid].spec.steps, c.rid
            assert c.max_divergence <= 1e-8, c.rid
        else:
            assert c.errorcode == 2, c.rid                                 # AFID_NUM_SINGULAR
    assert srv.peak_bytes <= 200e9 and srv.peak_active >= 2
    ctl = STEP_CTL
    for rid in sample:
        r = by_rid[rid]
        c = Cell(r.spec.params)
        try:
            c.init_linear_noise(r.spec.seed, 1e-3)
            err, _, n = c.steps(0, r.spec.steps, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
            assert (err, n) == (0, r.spec.steps)
            for nm in FIELDS:
                np.testing.assert_array_equal(c.download(nm), kept[rid][nm], err_msg=f"request {rid} {nm}")
        finally:
            c.close()
