"""Oracle vs the reference AFiD routines, through the golden fixtures that
tests/golden/make_golden.py wrote from the reference Fortran (compiled
unchanged from the reference tree into oracle/_ref).

What the fixtures pin: CreateGrid; ExplicitTerms{VX,VY,VZ,Temp};
ImplicitAndUpdate{VX,VY,VZ,Temp} with SolveImpEqnUpdate_{X,YZ,Temp} (dgttrf/
dgttrs from MKL) including the stale rhs(1) row; CalcLocalDivergence;
CorrectVelocity; CorrectPressure; CalcMaxCFL; CheckDivergence;
GlobalQuantities; CalcDissipationNu; CalcPlateNu (6 digits); and 10 composed TimeMarcher steps with
main.F90's dt selection.  The reference pressure solve needs FFTW3, which is
absent, so the composed march runs the oracle's Poisson solve inside the
reference TimeMarcher; the Poisson solve itself is pinned in
test_oracle_math.py.
"""
import os

import numpy as np
import pytest

from conftest import demean, rel_l2
from oracle.oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PKEYS = ("nxm", "nym", "nzm", "nsst", "alx3", "istr3", "str3", "ylen", "zlen", "ray", "pra", "inslws", "inslwn")


def load(name):
    d = np.load(os.path.join(GOLD, name + ".npz"))
    p = {k: d[k].item() for k in PKEYS}
    return d, p


@pytest.mark.parametrize("case", ["tanh", "cheb", "unif_slip"])
def test_grid_matches_reference(case):
    d, p = load("grid_" + case)
    o = Oracle(**p)
    nx, nxm = p["nxm"] + 1, p["nxm"]
    g = d["raw"]
    off = 0
    for name in ("xc", "xm", "g3rc", "g3rm", "udx3c", "udx3m", "ap3ck", "ac3ck", "am3ck", "ap3sk", "ac3sk",
                 "am3sk", "ap3ssk", "ac3ssk", "am3ssk"):
        r = g[off:off + nx]
        off += nx
        a = o.arr(name)[1:nx + 1]
        n = nxm if name in ("xm", "g3rm", "udx3m") else nx
        np.testing.assert_array_equal(a[:n], r[:n], err_msg=name)
    ints = g[off:off + 4 * nxm].reshape(4, nxm)
    off += 4 * nxm
    for name, r in zip(("kmv", "kpv", "kmc", "kpc"), ints):
        np.testing.assert_array_equal(o.arr(name)[1:nxm + 1], r)
    sc = g[off:off + 14]
    assert (sc[0], sc[1], sc[2]) == (o.scalar("dx"), o.scalar("dy"), o.scalar("dz"))
    assert (sc[3], sc[4]) == (o.scalar("ren"), o.scalar("pec"))
    np.testing.assert_array_equal(sc[5:8], [o.scalar(f"gam{q}") for q in (1, 2, 3)])
    np.testing.assert_array_equal(sc[8:11], [o.scalar(f"rom{q}") for q in (1, 2, 3)])


def load_stage_oracle(d, p):
    o = Oracle(**p)
    for n in ("vx", "vy", "vz", "temp", "pr"):
        o.arr(n)[...] = d["in_" + n]
        o.L.afo_update_halo(o.h, n.encode())
    for n in ("rux", "ruy", "ruz", "rutemp", "rhs"):
        o.arr(n)[...] = d["in_" + n]
    o.set_substep(int(d["ns"]), float(d["dt"]))
    return o


@pytest.mark.parametrize("case", ["rk_sub1", "rk_sub2", "ab2"])
def test_stages_match_reference(case):
    d, p = load("stages_" + case)
    o = load_stage_oracle(d, p)
    nxm = p["nxm"]
    for r in ("explicit_terms_vx", "explicit_terms_vy", "explicit_terms_vz", "explicit_terms_temp"):
        o.call(r)
    # qcap, hro defined for k = 2..nxm; dph, dq for k = 1..nxm
    assert rel_l2(o.arr("qcap")[:, :, 1:nxm], d["qcap"][:, :, 1:nxm]) < 1e-15
    assert rel_l2(o.arr("dph"), d["dph_expl"]) < 1e-15
    assert rel_l2(o.arr("dq")[:, :, :nxm], d["dq"][:, :, :nxm]) < 1e-15
    assert rel_l2(o.arr("hro")[:, :, 1:nxm], d["hro"][:, :, 1:nxm]) < 1e-15
    for r in ("implicit_and_update_vx", "implicit_and_update_vy", "implicit_and_update_vz",
              "implicit_and_update_temp"):
        o.call(r)
    o.L.afo_update_halo(o.h, b"vy")
    o.L.afo_update_halo(o.h, b"vz")
    for n in ("vx", "vy", "vz", "temp"):
        assert rel_l2(o.arr(n), d[n + "_impl"]) < 1e-15, n
    for n in ("rhs", "rux", "ruy", "ruz", "rutemp"):
        assert rel_l2(o.arr(n), d[n + "_impl"]) < 1e-15, n
    # stale row: rhs(1,:,:) holds the vz increment at k=1 after the substep's solves
    np.testing.assert_allclose(o.arr("rhs")[:, :, 0], d["rhs_impl"][:, :, 0], rtol=1e-14, atol=0)
    o.call("calc_local_divergence")
    assert rel_l2(o.arr("dph"), d["dph_div"]) < 1e-15
    o.arr("dph")[...] = d["dph_in"]
    o.call("copy_dphhalo")
    o.call("correct_velocity")
    o.call("correct_pressure")
    for n in ("vx", "vy", "vz", "pr"):
        assert rel_l2(o.arr(n)[1:-1, 1:-1], d[n + "_corr"][1:-1, 1:-1]) < 1e-15, n
    o.L.afo_update_all_halos(o.h)
    E = d["diag"]
    assert o.cfl() == pytest.approx(E[0], rel=1e-15)
    assert o.divmax() == pytest.approx(E[1], rel=1e-15)
    gq = o.global_quantities()
    np.testing.assert_allclose(gq[4:7], E[2:5], rtol=1e-15)            # vmax(1..3)
    np.testing.assert_allclose(gq[1:4], E[5:8], rtol=1e-13)            # tempm, tempmax, tempmin
    assert gq[0] == pytest.approx(float(d["nu_vol"]), rel=1e-13)
    # nu_plate.out holds 6 significant digits (format e14.6, CalcPlateNu.F90:51-52)
    np.testing.assert_allclose(o.plate_nu(), d["nu_plate"], rtol=5e-6)
    mu, th, de, dth = o.dissipation()
    np.testing.assert_allclose([mu, th], d["nu_diss"], rtol=1e-13)
    np.testing.assert_allclose(de, E[8:8 + nxm], rtol=1e-13)
    np.testing.assert_allclose(dth, E[8 + nxm:8 + 2 * nxm], rtol=1e-13)


@pytest.mark.parametrize("case", ["fixed_dt", "cfl_dt"])
def test_march_matches_reference(case):
    d, p = load("march_" + case)
    o = Oracle(**p)
    o.init_linear_noise(int(d["seed"]), 1e-3)
    diag = d["diag"]
    for n in range(int(d["nsteps"])):
        err, dt = o.step(n, int(d["idtv"]), float(d["dt"]), float(d["dtmin"]), float(d["dtmax"]),
                         float(d["limitcfl"]))
        assert dt == pytest.approx(diag[n, 0], rel=1e-12)
    for n in ("vx", "vy", "vz", "temp"):
        assert rel_l2(o.arr(n), d[n]) < 1e-12, n
    nxm = p["nxm"]
    assert rel_l2(demean(o.arr("pr")[1:-1, 1:-1, :nxm]), demean(d["pr"][1:-1, 1:-1, :nxm])) < 1e-12
    assert o.divmax() < 1e-10
