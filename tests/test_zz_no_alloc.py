"""afid_calc_stats_noalloc against afid_calc_stats, through the C ABI.

The no-allocation entry point adds one statistics sample into a caller device
array, with its block partials in the library's static per-GPU scratch.  Two
calls in a row, the second on a smaller cell, must each equal what
afid_calc_stats accumulates for the same fields (the same kernels in the same
order, so bit for bit).  A cell whose partials exceed the scratch is refused
before any work.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIG = dict(nxm=40, nym=48, nzm=36, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e6, pra=1.0)
SMALL = dict(nxm=17, nym=12, nzm=10, istr3=4, str3=3.0, ylen=2.0, zlen=1.5, ray=1e5, pra=0.7)


def _stepped_cell(p, seed):
    from afid_amd.engine import Cell, CellParams
    cell = Cell(CellParams(nsst=3, **p))
    cell.init_linear_noise(seed, 1e-3)
    for n in range(2):
        cell.step(n, 1, 1e-3, 1e-8, 1e-2, 1.2)
    return cell


class DeviceArray:
    """n doubles of device memory from the HIP runtime the library itself uses
    (the test's own buffer: the entry point under test allocates nothing)."""

    def __init__(self, n):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8 * n)) == 0
        assert self.hip.hipMemset(self.ptr, 0, ctypes.c_size_t(8 * n)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(8 * self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _noalloc_sample(cell):
    acc = DeviceArray(9 * cell.nxm)
    try:
        cell.CalcStatsNoAlloc(acc.ptr.value)
        cell.synchronize()
        return acc.get()
    finally:
        acc.free()


def _reference_sample(cell):
    cell.reset_stats()
    cell.CalcStats()
    st, n = cell.stats()
    assert n == 1
    from afid_amd.engine import STAT_NAMES
    return np.concatenate([st[nm] for nm in STAT_NAMES[:9]])


def test_noalloc_stats_match_calc_stats_two_calls():
    cells = [_stepped_cell(BIG, 3), _stepped_cell(SMALL, 5)]
    try:
        for cell in cells:           # the second call is on the smaller cell
            print(f"noalloc stats: {cell.nxm}x{cell.nym}x{cell.nzm}", flush=True)
            got = _noalloc_sample(cell)
            ref = _reference_sample(cell)
            assert got.shape == ref.shape
            assert np.all(np.isfinite(got)) and np.any(got != 0.0)
            np.testing.assert_array_equal(got, ref)
            _, n = cell.stats()
            assert n == 1            # the context's own sample count is untouched
    finally:
        for cell in cells:
            cell.close()


def test_noalloc_stats_accumulate_and_refuse_oversize():
    from afid_amd.engine import AfidError, Cell, CellParams
    cell = _stepped_cell(SMALL, 7)
    acc = DeviceArray(9 * cell.nxm)
    try:
        cell.CalcStatsNoAlloc(acc.ptr.value)
        cell.CalcStatsNoAlloc(acc.ptr.value)
        cell.synchronize()
        two = acc.get()
        cell.reset_stats()
        cell.CalcStats()
        cell.CalcStats()
        st, n = cell.stats()
        from afid_amd.engine import STAT_NAMES
        np.testing.assert_array_equal(two, np.concatenate([st[nm] for nm in STAT_NAMES[:9]]))
    finally:
        acc.free()
        cell.close()
    # 256*9*ldx doubles of partials with ldx = 912 exceed the 16 MiB scratch
    big = Cell(CellParams(nsst=3, nxm=896, nym=4, nzm=4, istr3=6, str3=6.0, ylen=1.0, zlen=1.0,
                          ray=1e6, pra=1.0))
    acc = DeviceArray(9 * 896)
    try:
        with pytest.raises(AfidError) as ei:
            big.CalcStatsNoAlloc(acc.ptr.value)
        assert ei.value.code == -1
    finally:
        acc.free()
        big.close()
