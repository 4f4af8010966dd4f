"""sgipc_convert_srbk_noalloc against sgipc_convert_srbk through the C ABI:
the same unique upper BCOO, bit for bit (rows, cols and summed values), on two
calls in a row, the second with a smaller input (the shared scratch keeps the
first call's contents, which must not leak into the second); the capacity and
size limits return their status codes.  Tolerance: bit-exact (index and
summation-order work, both entry points sum each block in arena order)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_TRIPLETS = 1 << 18


def _convert(L, ctx, fn, dv, dr, dc, n, cap):
    ov = ctx.empty((cap, 9), np.float64)
    orr = ctx.empty(cap, np.int32)
    oc = ctx.empty(cap, np.int32)
    out = ctypes.c_int(-1)
    st = getattr(L, fn)(ctx.h, dv.vp, dr.vp, dc.vp, n, ov.vp, orr.vp, oc.vp, cap, ctypes.byref(out))
    return st, out.value, ov, orr, oc


def _check_same(pkg, ctx, orc, rv, rr, rc):
    L = pkg._lib.lib()
    dv, dr, dc = ctx.to_device(rv, np.float64), ctx.to_device(rr, np.int32), ctx.to_device(rc, np.int32)
    cap = rr.size
    st_a, n_a, av, ar, ac = _convert(L, ctx, "sgipc_convert_srbk", dv, dr, dc, rr.size, cap)
    st_b, n_b, bv, br, bc = _convert(L, ctx, "sgipc_convert_srbk_noalloc", dv, dr, dc, rr.size, cap)
    assert st_a == st_b == pkg._lib.SGIPC_OK, (st_a, st_b)
    assert n_a == n_b
    np.testing.assert_array_equal(br.download()[:n_b], ar.download()[:n_a])
    np.testing.assert_array_equal(bc.download()[:n_b], ac.download()[:n_a])
    np.testing.assert_array_equal(bv.download()[:n_b], av.download()[:n_a])
    # and both equal the oracle's restatement
    ov, orr, oc = orc.convert(rv, rr, rc)
    assert n_b == orr.size
    np.testing.assert_array_equal(bv.download()[:n_b], ov)
    return n_b


def test_noalloc_convert_matches_convert_twice(pkg, orc, gpu_ctx, scene_c0, scene_tiny):
    big = orc.triplets(scene_c0)[:3]
    small = orc.triplets(scene_tiny)[:3]
    assert big[1].size <= MAX_TRIPLETS and small[1].size < big[1].size
    n1 = _check_same(pkg, gpu_ctx, orc, *big)
    n2 = _check_same(pkg, gpu_ctx, orc, *small)
    print(f"noalloc convert: {big[1].size} -> {n1} blocks, then {small[1].size} -> {n2}")


def test_noalloc_convert_limits(pkg, orc, gpu_ctx, scene_tiny):
    L = pkg._lib.lib()
    rv, rr, rc = orc.triplets(scene_tiny)[:3]
    dv, dr, dc = gpu_ctx.to_device(rv, np.float64), gpu_ctx.to_device(rr, np.int32), gpu_ctx.to_device(rc, np.int32)
    need = orc.convert(rv, rr, rc)[1].size
    # undersized output: CAPACITY with the count the caller must provide
    st, n, *_ = _convert(L, gpu_ctx, "sgipc_convert_srbk_noalloc", dv, dr, dc, rr.size, need - 1)
    assert st == pkg._lib.SGIPC_CAPACITY and n == need
    # beyond the scratch's maximum: CAPACITY before any device work
    st, n, *_ = _convert(L, gpu_ctx, "sgipc_convert_srbk_noalloc", dv, dr, dc, MAX_TRIPLETS + 1, 4)
    assert st == pkg._lib.SGIPC_CAPACITY and n == 0
    assert b"allocation-free maximum" in L.sgipc_last_error(gpu_ctx.h)
    # and a full call afterwards still matches
    _check_same(pkg, gpu_ctx, orc, rv, rr, rc)
