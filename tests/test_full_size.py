"""Parity at the production size: 384 x 768 x 768 (clipped-Chebyshev stretching
istr3 = 6, str3 = 6), ylen = zlen = 1, Ra = 1e8, Pr = 1, RK3, seeded conduction
profile plus +-1e-3 noise.  Ten main.F90 steps with the CFL-limited dt rule
(the first two at dtmin, as main.F90 starts; then the dt the CFL number
allows, clipped at dtmax), the GPU path (host-driven steps, and steps from
captured graphs, the path bench.py times) against the oracle on the same
input: the same error code and dt per step, per-field rel-L2 <= 1e-11 (pr
modulo its mean), the maximum local divergence below 1e-10 in both, and the
plate, volume and dissipation Nusselt numbers (CalcPlateNu, GlobalQuantities,
CalcDissipationNu) to rel 1e-10.  The oracle holds about 40 GB of host memory
and runs the ten steps in about three minutes on 16 threads.
"""
import os

import numpy as np
import pytest

from conftest import demean, rel_l2

pytestmark = pytest.mark.gpu

CELL = dict(nxm=384, nym=768, nzm=768, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e8, pra=1.0)
CTL = (1, 1e-3, 1e-8, 1e-2, 1.2)   # idtv, dt, dtmin, dtmax, cflmax as in bou.in
NSTEPS = 10


def _check(cell, o, tag):
    nxm = CELL["nxm"]
    for nm in ("vx", "vy", "vz", "temp"):
        e = rel_l2(cell.download(nm), o.interior(nm))
        print(f"{tag} {nm}: rel-L2 {e:.2e}", flush=True)
        assert e < 1e-11, nm
    e = rel_l2(demean(cell.download("pr")[:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm]))
    print(f"{tag} pr (mod mean): rel-L2 {e:.2e}", flush=True)
    assert e < 1e-11
    assert cell.CheckDivergence() < 1e-10
    nb, nt = cell.CalcPlateNu()
    ob, ot = o.plate_nu()
    assert nb == pytest.approx(ob, rel=1e-10) and nt == pytest.approx(ot, rel=1e-10)
    gq, ogq = cell.GlobalQuantities(), o.global_quantities()
    np.testing.assert_allclose(gq, ogq, rtol=1e-10)
    mu, th, _, _ = cell.CalcDissipationNu()
    omu, oth, _, _ = o.dissipation()
    assert mu == pytest.approx(omu, rel=1e-10) and th == pytest.approx(oth, rel=1e-10)
    print(f"{tag} Nu plate {nb:.8f}/{nt:.8f} (oracle {ob:.8f}/{ot:.8f}), Nu vol {gq[0]:.8f} "
          f"(oracle {ogq[0]:.8f}), Nu diss {mu:.8f}/{th:.8f} (oracle {omu:.8f}/{oth:.8f})", flush=True)


@pytest.mark.timeout(900)
def test_production_size_steps_match_oracle():
    from afid_amd.engine import Cell, CellParams
    from oracle import oracle as O
    O.lib().afo_set_threads(min(16, os.cpu_count() or 1))
    print(f"full size 384x768x768: oracle and GPU, {NSTEPS} steps", flush=True)
    o = O.Oracle(nsst=3, alx3=1.0, **CELL)
    cell = Cell(CellParams(nsst=3, alx3=1.0, **CELL))
    try:
        o.init_linear_noise(2, 1e-3)
        cell.init_linear_noise(2, 1e-3)
        dts = []
        for n in range(NSTEPS):
            eo, dto = o.step(n, *CTL)
            eg, dtg = cell.step(n, *CTL)
            assert eg == eo == 0, n
            assert dtg == pytest.approx(dto, rel=1e-12), n
            dts.append(dtg)
            print(f"step {n}: errorcode {eg}, dt {dtg:.6e}", flush=True)
        assert dts[0] == dts[1] == CTL[2] and dts[-1] > CTL[2]
        assert o.divmax() < 1e-10
        _check(cell, o, "host steps")
        # the same steps from captured step graphs (afid_steps), the timed path of bench.py
        dev = Cell(CellParams(nsst=3, alx3=1.0, **CELL))
        try:
            dev.init_linear_noise(2, 1e-3)
            err, dt, done = dev.steps(0, NSTEPS, *CTL)
            assert (err, done) == (0, NSTEPS)
            assert dt == pytest.approx(dts[-1], rel=1e-12)
            _check(dev, o, "graph steps")
            # dt and the factorisations come from the device here, from the host
            # there: the two GPU paths agree to rounding, not bit for bit
            for nm in ("vx", "vy", "vz", "temp"):
                assert rel_l2(dev.download(nm), cell.download(nm)) < 1e-12, nm
        finally:
            dev.close()
    finally:
        cell.close()
        o.close()
