"""The two forms of the substep update against each other and the oracle.

The one-pass form (k_substep_update: stencil and wall-normal solves in one
kernel, b' kept on chip) and the two-pass form (k_stencil_rhs, then
k_column_solve through HBM) compute the same explicit terms in the same
operation order and solve the same factored systems; the one-pass solve
applies the chunk carries in a second pass instead of through precomputed
G/H rows, so the two agree to rounding.  AFID_UPDATE, read when a cell is
created, selects the one-pass form ("fused") where the column pitch allows it
(at most 448 stored positions), else the two-pass form.

Tolerances: one-pass against two-pass rel-L2 <= 1e-12 per field after
host-driven and device (graph) steps; both against the oracle <= 1e-11, the
bar of the path; the one-pass cell must have run k_substep_update and the
two-pass cell k_stencil_rhs + k_column_solve (HIP-event profile names).
"""
import os

import numpy as np
import pytest

from conftest import SMALL_CASES, case_id, demean, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp")
CASES = SMALL_CASES + [
    # three y columns, a pitch of exactly 448 (7 stencil waves, the one-pass
    # limit) and a chunk length of 25 (the production pitch of 400)
    dict(nxm=40, nym=3, nzm=8, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0),
    dict(nxm=447, nym=4, nzm=5, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0),
    dict(nxm=384, nym=6, nzm=8, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e7, pra=0.7),
]


def _cell(c, form):
    from afid_amd.engine import Cell, CellParams
    old = os.environ.get("AFID_UPDATE")
    os.environ["AFID_UPDATE"] = form
    try:
        return Cell(CellParams(nsst=3, **dict(c, alx3=c.get("alx3", 1.0))))
    finally:
        if old is None:
            del os.environ["AFID_UPDATE"]
        else:
            os.environ["AFID_UPDATE"] = old


@pytest.mark.parametrize("c", CASES, ids=case_id)
def test_one_pass_update_matches_two_pass_and_oracle(c):
    from oracle.oracle import Oracle
    print("update forms", case_id(c), flush=True)
    o = Oracle(nsst=3, **dict(c, alx3=c.get("alx3", 1.0)))
    one, two = _cell(c, "fused"), _cell(c, "split")
    try:
        for x in (o, one, two):
            x.init_linear_noise(6, 1e-3)
        one.profile(True)
        two.profile(True)
        for n in range(4):
            o.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
            one.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
            two.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
        k1, k2 = one.profile_read(), two.profile_read()
        assert k1.get("substep_update", (0, 0))[1] == 12 and "stencil_rhs" not in k1
        assert k2.get("stencil_rhs", (0, 0))[1] == 12 and k2.get("column_solve", (0, 0))[1] == 12
        assert "substep_update" not in k2
        one.profile(False)
        nxm = c["nxm"]
        for nm in FIELDS:
            a, b = one.download(nm), two.download(nm)
            assert rel_l2(a, b) <= 1e-12, nm
            assert rel_l2(a, o.interior(nm)) <= 1e-11, nm
        assert rel_l2(demean(one.download("pr")[:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm])) <= 1e-11
        # the stale row is the difference of two near-equal values (vz at k = 1
        # after and before the solve): compare it on the scale of vz itself
        np.testing.assert_allclose(one.get_stale(), two.get_stale(), rtol=0,
                                   atol=1e-12 * float(np.abs(two.download("vz")).max()))
        # device steps from captured graphs, continuing from the same state
        e1, dt1, d1 = one.steps(4, 6, 1, 5e-4, 1e-4, 5e-3, 1.2)
        e2, dt2, d2 = two.steps(4, 6, 1, 5e-4, 1e-4, 5e-3, 1.2)
        assert (e1, d1) == (e2, d2) == (0, 6)
        assert dt1 == pytest.approx(dt2, rel=1e-12)
        for n in range(4, 10):
            o.step(n, 1, 5e-4, 1e-4, 5e-3, 1.2)
        for nm in FIELDS:
            a = one.download(nm)
            assert rel_l2(a, two.download(nm)) <= 1e-12, nm
            assert rel_l2(a, o.interior(nm)) <= 1e-11, nm
        assert one.CheckDivergence() <= 1e-10
    finally:
        one.close()
        two.close()
        o.close()


def test_one_pass_form_refused_beyond_its_pitch():
    """nxm = 448 gives a pitch of 464 positions: the cell falls back to the
    two-pass form even when the one-pass form is asked for."""
    c = dict(nxm=448, nym=4, nzm=4, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0)
    cell = _cell(c, "fused")
    try:
        cell.init_linear_noise(1, 1e-3)
        cell.profile(True)
        cell.step(0, 0, 5e-4, 1e-4, 5e-3, 1.2)
        k = cell.profile_read()
        assert "substep_update" not in k and k["stencil_rhs"][1] == 3
    finally:
        cell.close()
