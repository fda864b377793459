"""The allocation-free entry point seg4dL2normNoAlloc against _l2norm_
(segmentation4D.c:194-213), through the C ABI: the same value bit for bit,
on two calls in a row with the second on a smaller input (the fixed device
scratch keeps the first call's partial sums), and a refusal beyond the
scratch's limit.

Device buffers come from hipMalloc of the HIP runtime the library itself is
linked against, resolved through the library's own handle."""
import ctypes
from ctypes import POINTER, c_double, c_int, c_long, c_size_t, c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip(seglib):
    # dlsym through the product library's handle searches its dependency
    # tree: these are the HIP runtime functions the library itself calls
    # (another copy of the runtime, e.g. one bundled with torch, may share
    # the process)
    H = seglib.lib
    H.hipMalloc.restype = c_int
    H.hipMalloc.argtypes = [POINTER(c_void_p), c_size_t]
    H.hipFree.restype = c_int
    H.hipFree.argtypes = [c_void_p]
    H.hipMemcpy.restype = c_int
    H.hipMemcpy.argtypes = [c_void_p, c_void_p, c_size_t, c_int]
    H.hipDeviceSynchronize.restype = c_int
    return H


def _dev(H, nbytes):
    p = c_void_p()
    assert H.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p


def _noalloc(lib, H, a, b, hh):
    """seg4dL2normNoAlloc on device copies of a, b (caller-owned buffers)."""
    f = lib.lib.seg4dL2normNoAlloc
    f.restype = c_int
    f.argtypes = [c_void_p, c_void_p, c_long, c_double, c_void_p, c_void_p]
    n = a.size
    da, db, dr = _dev(H, 8 * n), _dev(H, 8 * n), _dev(H, 8)
    try:
        assert H.hipMemcpy(da, a.ctypes.data, 8 * n, H2D) == 0
        assert H.hipMemcpy(db, b.ctypes.data, 8 * n, H2D) == 0
        rc = f(da, db, n, hh, dr, None)
        assert rc == 0, lib.last_error()
        assert H.hipDeviceSynchronize() == 0
        out = np.zeros(1)
        assert H.hipMemcpy(out.ctypes.data, dr, 8, D2H) == 0
        return float(out[0])
    finally:
        for p in (da, db, dr):
            H.hipFree(p)


@pytest.mark.parametrize("order", ["large_then_small"])
def test_no_alloc_l2norm_matches_l2norm(seglib, hip, order):
    print("[gpu] no-alloc _l2norm_ variant", flush=True)
    rng = np.random.default_rng(7)
    for dims, hh in (((40, 36, 20, 6), 1.0), ((12, 10, 8, 4), 0.01)):
        nx, ny, nz, nt = dims
        seglib.allocate(nx, ny, nz, nt)
        try:
            seglib.set(h=hh)
            a = rng.standard_normal(seglib.shape)
            b = rng.standard_normal(seglib.shape)
            seglib.arr("u")[...] = a
            seglib.arr("prev_u")[...] = b
            want = seglib.l2norm("u", "prev_u")
        finally:
            seglib.deallocate()
        got = _noalloc(seglib, hip, a, b, hh)
        print(f"  {dims} h={hh}: {got!r} vs _l2norm_ {want!r}", flush=True)
        assert got == want
        assert got == pytest.approx(np.sqrt(np.sum((a - b) ** 2) * hh ** 4), rel=1e-12)


def test_no_alloc_l2norm_refuses_oversize(seglib):
    f = seglib.lib.seg4dL2normNoAlloc
    f.restype = c_int
    f.argtypes = [c_void_p, c_void_p, c_long, c_double, c_void_p, c_void_p]
    # refused before any device access: the pointers are never touched
    dummy = c_void_p(256)
    assert f(dummy, dummy, 65536 * 4096 + 1, 1.0, dummy, None) == -1
    assert "outside 1..268435456" in seglib.last_error()
    assert f(dummy, dummy, 0, 1.0, dummy, None) == -1
    assert f(None, dummy, 100, 1.0, dummy, None) == -1
