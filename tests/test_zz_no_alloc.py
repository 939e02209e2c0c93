"""The allocation-free SpAdd (dv_range_spadd_noalloc) against dv_range_spadd:
device-resident operands, caller-allocated outputs, two calls in a row (the
second smaller, so the shared scratch array holds the first call's
contents), the capacity error and the size limit."""
import ctypes as C

import numpy as np
import pytest

import diverg_amd as dv
from diverg_amd import _native as N
from tests.rcrs_util import assert_identical

pytestmark = pytest.mark.gpu

HIP = C.CDLL("libamdhip64.so")


class DevBuf:
    """Device memory the test allocates (the entry point allocates none)."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        assert HIP.hipMalloc(C.byref(self.ptr), C.c_size_t(max(8, nbytes))) == 0

    def get(self, dtype, n):
        out = np.zeros(n, dtype)
        if n:
            assert HIP.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr,
                                 C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        HIP.hipFree(self.ptr)


def spadd_noalloc(ctx, a, b, capacity):
    """Call the entry point on device copies of a and b; returns the sum as
    a RangeCRS (or raises DiVerGError) and the entry count it reported."""
    da, db = ctx.to_device(a), ctx.to_device(b)
    keep = []
    n = a.nrows
    rm_dt = np.uint32 if a.size_bytes == 4 else np.uint64
    rm, en = DevBuf(a.size_bytes * (n + 1)), DevBuf(4 * capacity)
    total, nnz = C.c_uint64(), C.c_uint64()
    try:
        rc = N.lib().dv_range_spadd_noalloc(
            ctx._h, C.byref(da._desc(keep)), C.byref(db._desc(keep)), rm.ptr,
            C.cast(en.ptr, N.U32P), capacity, C.byref(total), C.byref(nnz), None)
        if rc:
            raise dv.DiVerGError(rc, "dv_range_spadd_noalloc")
        out = dv.RangeCRS(a.ncols, en.get(np.uint32, total.value), rm.get(rm_dt, n + 1),
                          nnz.value)
        return out, total.value
    finally:
        rm.free()
        en.free()
        da.free()
        db.free()


def graphs(bp, seed, size_bytes):
    a = dv.generate_graph(2, backbone_bp=bp, seed=seed, size_bytes=size_bytes)
    b = dv.generate_graph(2, backbone_bp=bp, seed=seed + 1, size_bytes=size_bytes)
    n = min(a.nrows, b.nrows)
    # the same shape: keep the first n rows of each, columns below n
    def cut(m):
        rm = m.rowmap[:n + 1].astype(np.int64)
        en = m.entries[:rm[-1]].copy()
        en = np.minimum(en, n - 1)
        return dv.RangeCRS(n, en, rm.astype(m.rowmap.dtype))
    return cut(a), cut(b)


@pytest.mark.parametrize("size_bytes", [4, 8])
def test_noalloc_spadd_matches_spadd_twice(ctx, size_bytes):
    for bp, seed in [(400_000, 3), (60_000, 9)]:  # the second call is smaller
        a, b = graphs(bp, seed, size_bytes)
        want = ctx.range_spadd(a, b)
        print(f"spadd_noalloc {a.nrows} rows, S={size_bytes}", flush=True)
        got, total = spadd_noalloc(ctx, a, b, len(a.entries) + len(b.entries))
        assert total == len(want.entries)
        assert_identical(got, want, f"noalloc spadd {bp} bp")
        i = ctx.identity(a.nrows, size_bytes)
        got, _ = spadd_noalloc(ctx, i, a, len(i.entries) + len(a.entries))
        assert_identical(got, ctx.range_spadd(i, a), "noalloc I + A")


def test_noalloc_spadd_capacity_and_limits(ctx):
    a, b = graphs(30_000, 5, 8)
    want = ctx.range_spadd(a, b)
    with pytest.raises(dv.DiVerGError) as e:
        spadd_noalloc(ctx, a, b, 2)
    assert e.value.status == 2  # DV_ERR_OVERFLOW: C does not fit c_capacity
    got, _ = spadd_noalloc(ctx, a, b, len(want.entries))  # exactly enough
    assert_identical(got, want, "exact capacity")
    cap = N.lib().dv_range_spadd_noalloc_max_rows()
    assert 1_000_000 < cap < (32 << 20) // 4
    big = cap + 1
    rm = np.full(big + 1, 2, np.uint64)
    rm[0] = 0
    m = dv.RangeCRS(big, np.zeros(2, np.uint32), rm)
    with pytest.raises(dv.DiVerGError) as e:
        spadd_noalloc(ctx, m, m, 0)
    assert e.value.status == 5  # DV_ERR_BUDGET
