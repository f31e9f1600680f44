"""Restart regridding (interp.F90, ReadFlowField.F90:56-158).

The oracle's afo_interp is pinned three ways on the CPU:
- against the reference's own interp, compiled from its source into
  oracle/_ref/ref_harness (mode 4), bit for bit, on refinement, coarsening, a
  change of stretching and uniform -> stretched, for all four fields;
- against closed forms that follow from interp.F90 as written: on its own
  grid a field comes back shifted down one x level, with the extrapolated
  2 a(1) - a(2) at k = 1 (the upper-x corners are read at the lower level);
  fields linear in y and z and independent of x are reproduced exactly
  (trilinear weights on the uniform lateral grids, linear halo
  extrapolation); an x-only field g(k) gives
  g(lo) + (g(hi) - g(lo)) wx' wy wz, the one upper-x corner's weight;
- gridnew against CreateGrid's stretched faces and the uniform formula.
The refusals (more than twice the resolution, stretching 4, a point outside
the old window) come back as codes.
GPU: afid_interp_field equals the oracle bit for bit; a continuation written
on one grid and read on another gives the oracle's interpolated fields with
pr = 0, and the next steps match the oracle from there.
"""
import os
import tempfile

import numpy as np
import pytest

from conftest import rel_l2

LAT = dict(ylen=1.0, zlen=1.5)
# (old grid, new grid): (nxm, nym, nzm, istr3, str3)
CASES = [
    ((16, 12, 10, 6, 6.0), (24, 20, 16, 6, 6.0)),     # refine
    ((16, 12, 10, 6, 6.0), (16, 12, 10, 6, 4.0)),     # stretching only
    ((24, 20, 16, 6, 6.0), (16, 12, 10, 6, 6.0)),     # coarsen by 1.5
    ((16, 12, 10, 0, 0.0), (20, 16, 12, 6, 3.0)),     # uniform -> stretched
    ((16, 12, 10, 6, 6.0), (32, 24, 20, 6, 8.0)),     # refine by 2 (the limit)
]


def ids(c):
    o, n = c
    return f"{o[0]}x{o[1]}x{o[2]}s{o[4]:g}-{n[0]}x{n[1]}x{n[2]}s{n[4]:g}"


def new_grid(n):
    return dict(nxm=n[0], nym=n[1], nzm=n[2], istr3=n[3], str3=n[4], **LAT)


def old_dims(o):
    return o[0] + 1, o[1] + 1, o[2] + 1          # nxo, nyo, nzo as continua_master.h5 holds them


def harness_params(n):
    return dict(new_grid(n), nsst=3, alx3=1.0, ray=1e6, pra=1.0, inslws=1, inslwn=1)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_oracle_interp_matches_reference_routine(case):
    from oracle import oracle as O
    if not O.ref_available():
        pytest.skip("reference harness not built (oracle/_ref)")
    o, n = case
    nxo, nyo, nzo = old_dims(o)
    old = np.random.default_rng(3).standard_normal((nzo - 1, nyo - 1, nxo))
    for iv in (1, 2, 3, 4):
        code, got = O.interp(old, nxo, nyo, nzo, o[3], o[4], new_grid(n), iv)
        assert code == 0
        with tempfile.TemporaryDirectory() as d:
            ref = O.run_ref_interp(d, harness_params(n), old, nxo, nyo, nzo, o[3], o[4], iv)
        np.testing.assert_array_equal(got, ref, err_msg=f"intvar {iv}")


def test_gridnew_faces():
    from oracle import oracle as O
    from oracle.oracle import Oracle
    rc, rm = O.gridnew(9, 1.5, 0, 1.0)
    np.testing.assert_allclose(rc, 1.5 * np.arange(9) / 8, rtol=1e-15, atol=0)
    np.testing.assert_allclose(rm[:-1], 1.5 * (np.arange(8) + 0.5) / 8, rtol=1e-15)
    assert rm[-1] == 2 * rc[-1] - rm[-2]
    for nxm, s in ((24, 6.0), (31, 3.0)):
        rc, rm = O.gridnew(nxm + 1, 1.0, 6, s)
        o = Oracle(nxm=nxm, nym=8, nzm=8, istr3=6, str3=s, alx3=1.0)
        xc = o.arr("xc")[1:nxm + 2].copy()      # a view into the oracle's memory: copy before closing
        o.close()
        np.testing.assert_allclose(rc, xc, rtol=1e-14, atol=1e-16)
        assert np.all(np.diff(rc) > 0)


@pytest.mark.parametrize("intvar", [1, 2, 3, 4])
def test_interp_on_its_own_grid_shifts_one_level(intvar):
    from oracle import oracle as O
    o = (20, 12, 10, 6, 5.0)
    nxo, nyo, nzo = old_dims(o)
    old = np.random.default_rng(intvar).standard_normal((nzo - 1, nyo - 1, nxo))
    code, got = O.interp(old, nxo, nyo, nzo, o[3], o[4], new_grid(o), intvar)
    assert code == 0
    want = np.empty_like(old)
    want[:, :, 1:] = old[:, :, :-1]
    want[:, :, 0] = 2.0 * old[:, :, 0] - old[:, :, 1]
    np.testing.assert_allclose(got, want, rtol=4e-15, atol=4e-15 * np.abs(old).max())


def axes(intvar, nzo_or_nz, nyo_or_ny, nxo_or_nx, istr, s, old):
    """Coordinates interp.F90:73-101 assigns to (i, j, k) for one field."""
    from oracle import oracle as O
    tc, tm = O.gridnew(nzo_or_nz, LAT["zlen"], 0, 1.0)
    rc, rm = O.gridnew(nyo_or_ny, LAT["ylen"], 0, 1.0)
    zz, zm = O.gridnew(nxo_or_nx, 1.0, istr, s)
    xa, ya, za = (tc if intvar == 1 else tm), (rc if intvar == 2 else rm), (zz if intvar == 3 else zm)
    if old:    # 0..n+1 with the linear end extrapolation
        ext = lambda v: np.concatenate([[2 * v[0] - v[1]], v, [2 * v[-1] - v[-2]]])
        return ext(xa), ext(ya), ext(za)
    return xa, ya, za


@pytest.mark.parametrize("intvar", [1, 2, 3, 4])
def test_interp_reproduces_lateral_linear_fields(intvar):
    from oracle import oracle as O
    o, n = (16, 12, 10, 6, 6.0), (24, 20, 16, 6, 6.0)
    nxo, nyo, nzo = old_dims(o)
    xo, yo, zo = axes(intvar, nzo, nyo, nxo, o[3], o[4], True)
    xn, yn, zn = axes(intvar, n[2] + 1, n[1] + 1, n[0] + 1, n[3], n[4], False)
    a, b, c = 0.3, -1.7, 0.9
    # old(i, j) at the file's points i = 1..nzom, j = 1..nyom
    old = (a + b * xo[1:nzo, None, None] + c * yo[None, 1:nyo, None]) * np.ones((1, 1, nxo))
    code, got = O.interp(old, nxo, nyo, nzo, o[3], o[4], new_grid(n), intvar)
    assert code == 0
    want = (a + b * xn[:n[2], None, None] + c * yn[None, :n[1], None]) * np.ones((1, 1, n[0] + 1))
    # points whose brackets touch the halo in two axes read the halo edges
    # the reference leaves unset; compare the rest
    sel = ~edge_points(xo, yo, zo, xn, yn, zn, nzo, nyo, nxo)
    assert sel.mean() > 0.8
    np.testing.assert_allclose(got[sel], want[sel], rtol=1e-13, atol=1e-13)


def bracket(oldc, nn, x):
    """First l in 1..nn with oldc[l] >= x gives (l-1, l); (nn-1, nn) when none."""
    for l in range(1, nn + 1):
        if oldc[l] >= x:
            return l - 1, l
    return nn - 1, nn


def edge_points(xo, yo, zo, xn, yn, zn, nzo, nyo, nxo):
    """Mask (i, j, k) of the new points that read a halo edge or corner."""
    hi = np.array([bracket(xo, nzo, x)[0] == 0 or bracket(xo, nzo, x)[1] == nzo for x in xn[:len(xn) - 1]])
    hj = np.array([bracket(yo, nyo, y)[0] == 0 or bracket(yo, nyo, y)[1] == nyo for y in yn[:len(yn) - 1]])
    hk = np.array([bracket(zo, nxo, z)[0] == 0 for z in zn])
    cnt = hi[:, None, None].astype(int) + hj[None, :, None] + hk[None, None, :]
    return cnt >= 2


@pytest.mark.parametrize("intvar", [1, 2, 3, 4])
def test_interp_x_only_field_uses_one_upper_corner(intvar):
    from oracle import oracle as O
    o, n = (16, 12, 10, 6, 6.0), (24, 20, 16, 6, 6.0)
    nxo, nyo, nzo = old_dims(o)
    xo, yo, zo = axes(intvar, nzo, nyo, nxo, o[3], o[4], True)
    xn, yn, zn = axes(intvar, n[2] + 1, n[1] + 1, n[0] + 1, n[3], n[4], False)
    g = np.cos(3.0 * np.arange(1, nxo + 1)) + 2.0
    old = np.broadcast_to(g, (nzo - 1, nyo - 1, nxo)).copy()
    code, got = O.interp(old, nxo, nyo, nzo, o[3], o[4], new_grid(n), intvar)
    assert code == 0
    gpad = np.concatenate([[2 * g[0] - g[1]], g, [2 * g[-1] - g[-2]]])
    edge = edge_points(xo, yo, zo, xn, yn, zn, nzo, nyo, nxo)
    ok = 0
    for i in range(n[2]):
        fi, ci = bracket(xo, nzo, xn[i])
        for j in range(n[1]):
            fj, cj = bracket(yo, nyo, yn[j])
            for k in range(n[0] + 1):
                if edge[i, j, k]:
                    continue
                fk, ck = bracket(zo, nxo, zn[k])
                w = ((xo[ci] - xn[i]) / (xo[ci] - xo[fi]) * (yn[j] - yo[fj]) / (yo[cj] - yo[fj])
                     * (zn[k] - zo[fk]) / (zo[ck] - zo[fk]))
                want = gpad[fk] + (gpad[ck] - gpad[fk]) * w
                assert abs(got[i, j, k] - want) <= 3e-13, (i, j, k)
                ok += 1
    assert ok > 0.9 * got.size


def test_interp_refusals():
    from oracle import oracle as O
    o = (8, 8, 8, 6, 6.0)
    nxo, nyo, nzo = old_dims(o)
    old = np.zeros((nzo - 1, nyo - 1, nxo))
    assert O.interp(old, nxo, nyo, nzo, 6, 6.0, new_grid((8, 8, 20, 6, 6.0)), 4)[0] == 1   # > 2x in z
    assert O.interp(old, nxo, nyo, nzo, 4, 3.0, new_grid((8, 8, 8, 6, 6.0)), 4)[0] == 2    # gridnew: no istr 4
    o2 = (8, 16, 8, 6, 6.0)                     # y coarsened by 2: vy's lower bracket leaves the window
    nxo, nyo, nzo = old_dims(o2)
    old = np.zeros((nzo - 1, nyo - 1, nxo))
    assert O.interp(old, nxo, nyo, nzo, 6, 6.0, new_grid((8, 8, 8, 6, 6.0)), 2)[0] == 3


# ---- GPU -------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=ids)
def test_interp_field_matches_oracle(case):
    from afid_amd import engine
    from oracle import oracle as O
    o, n = case
    nxo, nyo, nzo = old_dims(o)
    cell = engine.Cell(engine.CellParams(ray=1e6, pra=1.0, **new_grid(n)))
    rng = np.random.default_rng(11)
    for iv, nm in ((1, "vx"), (2, "vy"), (3, "vz"), (4, "temp")):
        old = rng.standard_normal((nzo - 1, nyo - 1, nxo))
        print("interp", ids(case), nm, flush=True)
        cell.interp(nm, old, nxo, nyo, nzo, o[3], o[4])
        code, want = O.interp(old, nxo, nyo, nzo, o[3], o[4], new_grid(n), iv)
        assert code == 0
        np.testing.assert_array_equal(cell.download(nm), want, err_msg=nm)
    cell.close()


@pytest.mark.gpu
def test_interp_field_refusals():
    from afid_amd import engine
    cell = engine.Cell(engine.CellParams(ray=1e6, pra=1.0, **new_grid((8, 8, 8, 6, 6.0))))
    for (o, nm) in (((8, 8, 3, 6, 6.0), "temp"), ((8, 8, 8, 4, 3.0), "temp"), ((8, 16, 8, 6, 6.0), "vy")):
        nxo, nyo, nzo = old_dims(o)
        with pytest.raises(engine.AfidError, match="cannot be interpolated"):
            cell.interp(nm, np.zeros((nzo - 1, nyo - 1, nxo)), nxo, nyo, nzo, o[3], o[4])
    with pytest.raises(engine.AfidError):
        cell.interp("pr", np.zeros((8, 8, 9)), 9, 9, 9, 6, 6.0)
    cell.close()


@pytest.mark.gpu
def test_restart_on_another_grid(tmp_path):
    from afid_amd import engine
    from oracle.oracle import Oracle
    from oracle import oracle as O
    if not os.path.exists(engine.IO_LIB_PATH):
        pytest.skip("libafid_io.so not built")
    o, n = (16, 12, 10, 6, 6.0), (24, 20, 16, 6, 5.0)
    a = engine.Cell(engine.CellParams(ray=1e6, pra=1.0, **new_grid(o)))
    a.init_linear_noise(6, 1e-3)
    for s in range(2):
        a.step(s, 0, 2e-3)
    a.WriteFlowField(str(tmp_path))
    old = {nm: a.download(nm) for nm in ("vx", "vy", "vz", "temp")}
    t_a = a.time
    b = engine.Cell(engine.CellParams(ray=1e6, pra=1.0, **new_grid(n)))
    b.init_linear_noise(1, 1e-3)
    b.ReadFlowField(str(tmp_path))
    assert b.time == t_a
    nxo, nyo, nzo = old_dims(o)
    orc = Oracle(nsst=3, alx3=1.0, ray=1e6, pra=1.0, **new_grid(n))
    for iv, nm in ((1, "vx"), (2, "vy"), (3, "vz"), (4, "temp")):
        code, want = O.interp(old[nm], nxo, nyo, nzo, o[3], o[4], new_grid(n), iv)
        assert code == 0
        np.testing.assert_array_equal(b.download(nm), want, err_msg=nm)
        orc.set_interior(nm, want)
    np.testing.assert_array_equal(b.download("pr"), 0.0)
    for s in range(2):
        orc.step(s, 0, 1e-3, 1e-8, 1e-2, 1.2)
        b.step(s, 0, 1e-3)
    for nm in ("vx", "vy", "vz", "temp"):
        assert rel_l2(b.download(nm), orc.interior(nm)) < 1e-11, nm
    a.close()
    b.close()
