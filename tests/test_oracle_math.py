"""Pins of the C oracle that do not need the reference build: library
routines (numpy.fft, scipy LAPACK), closed forms and invariants.

Each check is chosen so that a plausible slip in the oracle (wrong sign,
index, normalisation, transposed operand, dropped term) makes it fail.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from conftest import demean, rel_l2
from oracle import oracle as O


# ---------------------------------------------------------------- FFT (FFTW conventions)
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 15, 16, 30, 45, 60, 64, 96, 100, 128, 240, 384, 768])
def test_fft_c2c_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(O.fft_c2c(x, -1), np.fft.fft(x), rtol=0, atol=1e-13 * n)
    # backward is unnormalised: n * ifft
    assert np.allclose(O.fft_c2c(x, +1), n * np.fft.ifft(x), rtol=0, atol=1e-13 * n)


@pytest.mark.parametrize("n", [4, 6, 15, 16, 30, 128, 768])
def test_fft_r2c_c2r(n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    X = O.fft_r2c(x)
    assert np.allclose(X, np.fft.rfft(x), atol=1e-13 * n)
    # c2r is unnormalised and ignores the imaginary parts of DC / Nyquist
    Xb = X.copy()
    Xb[0] += 0.37j
    if n % 2 == 0:
        Xb[-1] += 0.91j
    y = O.fft_c2r(Xb, n)
    assert np.allclose(y, n * x, atol=1e-12 * n)


# ---------------------------------------------------------------- LAPACK restatement
@pytest.mark.parametrize("n,pivot", [(5, False), (40, False), (40, True), (3, True)])
def test_dgttrf_dgttrs_match_scipy(n, pivot):
    rng = np.random.default_rng(n)
    dl = rng.standard_normal(n - 1)
    du = rng.standard_normal(n - 1)
    d = 4.0 + rng.random(n) if not pivot else 0.1 * rng.standard_normal(n)  # weak diagonal forces interchanges
    b = rng.standard_normal((n, 3))
    sdl, sd, sdu, sdu2, sipiv, info = lapack.dgttrf(dl, d, du)
    assert info == 0
    sx, info = lapack.dgttrs(sdl, sd, sdu, sdu2, sipiv, b)[:2]
    import ctypes
    L = O.lib()
    odl, od, odu = dl.copy(), d.copy(), du.copy()
    odu2 = np.zeros(max(n - 2, 1))
    ipiv = np.zeros(n, dtype=np.int32)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    info = L.afo_dgttrf(n, dp(odl), dp(od), dp(odu), dp(odu2), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert info == 0
    assert np.array_equal(ipiv + 1, sipiv)               # same interchanges
    np.testing.assert_allclose(od, sd, rtol=1e-14, atol=0)
    ob = np.asfortranarray(b.copy())
    L.afo_dgttrs(n, 3, dp(odl), dp(od), dp(odu), dp(odu2), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                 dp(ob), n)
    np.testing.assert_allclose(ob, sx, rtol=1e-12, atol=1e-13)
    A = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
    np.testing.assert_allclose(A @ ob, b, atol=1e-11)


# ---------------------------------------------------------------- grid closed forms
def test_grid_closed_forms():
    o = O.Oracle(20, 12, 10, istr3=4, str3=3.0, alx3=1.0, ylen=2.0, zlen=1.0)
    nx, nxm = 21, 20
    xc = o.arr("xc")[1:nx + 1]
    k = np.arange(1, nx + 1)
    ref = (1 + np.tanh(3.0 * (2 * k - nx - 1) / nxm) / np.tanh(3.0)) * 0.5
    ref[0] = 0.0
    np.testing.assert_allclose(xc, ref, atol=1e-15)
    # the wall-normal second-derivative stencils are exact on quadratics: d2(x^2)/dx2 = 2
    am, ac, ap = o.arr("am3ck"), o.arr("ac3ck"), o.arr("ap3ck")
    for kc in range(2, nxm + 1):
        # centred variable lives on xc (xc[kc-1] is xc(kc))
        lap = ap[kc] * xc[kc] ** 2 + ac[kc] * xc[kc - 1] ** 2 + am[kc] * xc[kc - 2] ** 2
        assert abs(lap - 2.0) < 1e-9
    xm = o.arr("xm")[1:nx]
    am, ac, ap = o.arr("am3sk"), o.arr("ac3sk"), o.arr("ap3sk")
    for kc in range(2, nxm):
        # staggered variable lives on xm (xm[kc-1] is xm(kc)); the finite-volume stencil
        # dx^2/g3rm(k) * [(q(k+1)-q(k))/g3rc(k+1) - (q(k)-q(k-1))/g3rc(k)] is exact on linears and
        # maps x^2 to (xm(k+1)-xm(k-1))/(xc(k+1)-xc(k))
        lin = ap[kc] * xm[kc] + ac[kc] * xm[kc - 1] + am[kc] * xm[kc - 2]
        assert abs(lin) < 1e-9 * abs(ap[kc])
        lap = ap[kc] * xm[kc] ** 2 + ac[kc] * xm[kc - 1] ** 2 + am[kc] * xm[kc - 2] ** 2
        assert abs(lap - (xm[kc] - xm[kc - 2]) / (xc[kc] - xc[kc - 1])) < 1e-9
    # modified wavenumbers: ak2(j) = 2(1-cos(2 pi (j-1)/nym)) (nym/ylen)^2, even in the index
    ak2 = o.arr("ak2")[1:13]
    jj = np.arange(12)
    np.testing.assert_allclose(ak2, 2 * (1 - np.cos(2 * np.pi * jj / 12)) * (12 / 2.0) ** 2, rtol=1e-13, atol=1e-12)
    # pressure x operator rows sum to zero (Neumann), amphk(1)=apphk(nxm)=0
    amp, acp, app = o.arr("amphk"), o.arr("acphk"), o.arr("apphk")
    assert amp[1] == 0.0 and app[nxm] == 0.0
    np.testing.assert_allclose(amp[1:nxm + 1] + acp[1:nxm + 1] + app[1:nxm + 1], 0.0, atol=1e-9)
    # RK3 coefficients: sum(alpha) = 1, gamma/rho exact rationals
    assert abs(o.scalar("alm1") + o.scalar("alm2") + o.scalar("alm3") - 1.0) < 1e-15
    assert o.scalar("gam1") == 8 / 15 and o.scalar("rom2") == -17 / 60


# ---------------------------------------------------------------- Poisson solve
def _lap(o, phi):
    """Discrete operator applied by CorrectPressure / the pressure solve: dzz*dzq + dyy*dyq + x-operator."""
    nxm = o.nxm
    amp, acp, app = o.arr("amphk")[1:nxm + 1], o.arr("acphk")[1:nxm + 1], o.arr("apphk")[1:nxm + 1]
    dyq, dzq = o.scalar("dyq"), o.scalar("dzq")
    out = (np.roll(phi, -1, 0) - 2 * phi + np.roll(phi, 1, 0)) * dzq
    out += (np.roll(phi, -1, 1) - 2 * phi + np.roll(phi, 1, 1)) * dyq
    km = np.concatenate([[0], np.arange(nxm - 1)])
    kp = np.concatenate([np.arange(1, nxm), [nxm - 1]])
    out += phi[:, :, kp] * app + phi * acp + phi[:, :, km] * amp
    return out


@pytest.mark.parametrize("nxm,nym,nzm,istr3,str3", [(16, 12, 10, 4, 3.0), (24, 15, 16, 6, 6.0), (9, 8, 25, 4, 2.0)])
def test_pressure_solve_inverts_operator(nxm, nym, nzm, istr3, str3):
    o = O.Oracle(nxm, nym, nzm, istr3=istr3, str3=str3, ylen=1.3, zlen=0.7)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((nzm, nym, nxm))
    # make it compatible with the Neumann x-operator: remove the g3rm-weighted mean
    g3rm = o.arr("g3rm")[1:nxm + 1]
    rhs -= (rhs * g3rm).sum() / (g3rm.sum() * nym * nzm)
    o.arr("dph")[...] = rhs
    assert o.L.afo_solve_pressure_correction(o.h) == 0
    phi = o.arr("dph").copy()
    res = _lap(o, phi) - rhs
    assert np.abs(res).max() < 1e-9 * np.abs(rhs).max() * max(o.scalar("dyq"), o.scalar("dzq"))
    # brute-force: dense solve on the same operator with the mean fixed
    n = nxm * nym * nzm
    if n <= 2500:
        A = np.zeros((n, n))
        for c in range(n):
            e = np.zeros(n)
            e[c] = 1
            A[:, c] = _lap(o, e.reshape(nzm, nym, nxm)).ravel()
        A2 = np.vstack([A, np.ones(n)])
        sol = np.linalg.lstsq(A2, np.concatenate([rhs.ravel(), [0.0]]), rcond=None)[0]
        assert rel_l2(demean(phi), demean(sol.reshape(phi.shape))) < 1e-9


def test_projection_makes_velocity_solenoidal():
    o = O.Oracle(16, 12, 10, istr3=4, str3=3.0)
    rng = np.random.default_rng(9)
    for nm in ("vx", "vy", "vz"):
        v = rng.standard_normal((10, 12, 17)) * 1e-2
        if nm == "vx":
            v[:, :, 0] = 0.0
            v[:, :, -1] = 0.0
        if nm != "vx":
            v[:, :, -1] = 0.0
        o.set_interior(nm, v)
    o.set_substep(2, 1e-3)
    before = o.divmax()
    o.call("calc_local_divergence")
    assert o.L.afo_solve_pressure_correction(o.h) == 0
    o.call("copy_dphhalo")
    o.call("correct_velocity")
    o.call("update_all_halos")
    assert before > 1.0
    assert o.divmax() < 1e-12 * before


def test_uniform_grid_zero_mode_is_singular():
    # rows of the (0,0) pressure mode sum to zero; on a uniform grid the netlib
    # elimination leaves an exactly zero pivot, and the reference aborts
    o = O.Oracle(16, 8, 8, istr3=0)
    o.arr("dph")[...] = 0.0
    assert o.L.afo_solve_pressure_correction(o.h) == 16


# ---------------------------------------------------------------- invariants of the substep
def test_conduction_state_is_a_fixed_point():
    o = O.Oracle(20, 12, 10, istr3=4, str3=3.0)
    o.init_linear_noise(1, amp=0.0)   # T = 1 - xc exactly, u = 0
    T0 = o.interior("temp")
    for ntime in range(3):
        o.step(ntime, 0, 1e-3, 1e-8, 1e-2, 1.2)
    np.testing.assert_allclose(o.interior("temp"), T0, atol=1e-14)
    assert np.abs(o.interior("vx")).max() < 1e-14
    nb, nt = o.plate_nu()
    assert abs(nb - 1.0) < 1e-12 and abs(nt - 1.0) < 1e-12
    assert abs(o.global_quantities()[0] - 1.0) < 1e-12


def _reflect_y(a, face):
    """y -> -y on the staggered grid: cell centres ym(j) map to ym(1-j), faces yc(j) to yc(2-j)."""
    nym = a.shape[1]
    idx = (-np.arange(nym)) % nym if face else (-1 - np.arange(nym)) % nym
    return a[:, idx, :]


def test_y_reflection_symmetry():
    # Reflecting y (vy odd on faces, everything else even on centres) commutes with the
    # substep.  A wrong stencil index or sign in any y-term breaks it.  (y<->z exchange is
    # not a symmetry of the reference: the stale rhs(1) row feeds the vz increment, never
    # the vy one, into the vx and temperature solves.)
    o1 = O.Oracle(13, 10, 12, istr3=4, str3=2.0, ylen=1.3, zlen=0.9)
    o2 = O.Oracle(13, 10, 12, istr3=4, str3=2.0, ylen=1.3, zlen=0.9)
    rng = np.random.default_rng(4)
    sh = (12, 10, 14)
    T = 0.5 + 0.1 * rng.standard_normal(sh)
    T[:, :, 0], T[:, :, -1] = 1.0, 0.0
    vx, vy, vz = (1e-2 * rng.standard_normal(sh) for _ in range(3))
    vx[:, :, 0] = vx[:, :, -1] = 0
    vy[:, :, -1] = vz[:, :, -1] = 0
    for nm, a in (("temp", T), ("vx", vx), ("vy", vy), ("vz", vz)):
        o1.set_interior(nm, a)
        o2.set_interior(nm, -_reflect_y(a, True) if nm == "vy" else _reflect_y(a, False))
    for ns in (1, 2, 3, 1, 2, 3):
        o1.substep(ns, 1e-3)
        o2.substep(ns, 1e-3)
    for nm in ("temp", "vx", "vz"):
        assert rel_l2(o1.interior(nm), _reflect_y(o2.interior(nm), False)) < 1e-13, nm
    assert rel_l2(o1.interior("vy"), -_reflect_y(o2.interior("vy"), True)) < 1e-13
