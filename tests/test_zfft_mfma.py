"""The z transforms with their 16-point stage on the matrix cores
(AFID_ZFFT=mfma, read when a cell is created; FP64, nzm = 16 Q with Q in
16, 32, 48, 64) against numpy.fft and against the radix-stage kernels.

Tolerances: the forward and backward z stages against numpy.fft (FFTW
conventions: unnormalised, forward sign -1) at rel-L2 <= 1e-14, as the radix
kernels; steps of a cell with the matrix-core transforms against the same
cell with the radix kernels at rel-L2 <= 1e-13 per field, and against the
oracle at <= 1e-11, the bar of the path.
"""
import os

import numpy as np
import pytest

from conftest import demean, rel_l2

pytestmark = pytest.mark.gpu


def _cell(p, zfft):
    from afid_amd.engine import Cell, CellParams
    old = os.environ.get("AFID_ZFFT")
    os.environ["AFID_ZFFT"] = zfft
    try:
        return Cell(CellParams(nsst=3, alx3=1.0, **p))
    finally:
        if old is None:
            del os.environ["AFID_ZFFT"]
        else:
            os.environ["AFID_ZFFT"] = old


@pytest.mark.parametrize("nzm", [256, 512, 768, 1024])
def test_mfma_z_stages_match_numpy(nzm):
    p = dict(nxm=6, nym=10, nzm=nzm, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0)
    cell = _cell(p, "mfma")
    try:
        rng = np.random.default_rng(nzm)
        d = rng.standard_normal((nzm, p["nym"], p["nxm"]))
        cell.upload("dph", d)
        s1 = np.fft.rfft(d, axis=1)
        s2 = np.fft.fft(s1, axis=0)
        assert rel_l2(cell.pressure_spectrum(2), s2) < 1e-14
        g3 = cell.pressure_spectrum(3)
        assert rel_l2(cell.pressure_spectrum(4), np.fft.ifft(g3, axis=0) * nzm) < 1e-14
    finally:
        cell.close()


def test_mfma_z_steps_match_radix_and_oracle():
    from oracle.oracle import Oracle
    p = dict(nxm=16, nym=12, nzm=256, istr3=6, str3=6.0, ylen=1.0, zlen=2.0, ray=1e6, pra=1.0)
    o = Oracle(nsst=3, alx3=1.0, **p)
    mf, rx = _cell(p, "mfma"), _cell(p, "radix")
    try:
        for x in (o, mf, rx):
            x.init_linear_noise(4, 1e-3)
        for n in range(4):
            o.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
            mf.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
            rx.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
        e1, _, d1 = mf.steps(4, 4, 0, 5e-4, 1e-4, 5e-3, 1.2)
        e2, _, d2 = rx.steps(4, 4, 0, 5e-4, 1e-4, 5e-3, 1.2)
        assert (e1, d1) == (e2, d2) == (0, 4)
        for n in range(4, 8):
            o.step(n, 0, 5e-4, 1e-4, 5e-3, 1.2)
        for nm in ("vx", "vy", "vz", "temp"):
            a = mf.download(nm)
            assert rel_l2(a, rx.download(nm)) <= 1e-13, nm
            assert rel_l2(a, o.interior(nm)) <= 1e-11, nm
        nxm = p["nxm"]
        assert rel_l2(demean(mf.download("pr")[:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm])) <= 1e-11
    finally:
        mf.close()
        rx.close()
        o.close()
