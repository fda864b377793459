"""The oracle's threaded branches (nthreads > 1), which the benchmark times as
the CPU baseline and the large-volume parity tests use as the checker.

Pins:
  - every threaded branch gives bitwise the serial oracle's u after a fixed
    number of sweeps: colour passes, presmoothing passes and the coefficient
    and edge-indicator loops are order-independent, so splitting the outer t
    loop over threads (mcf_4D_openmp.c:127-128, :155, :184, :218) changes no
    bit;
  - par_reduce = 2 (terms in parallel, serial sum) gives bitwise the serial
    residuals, hence the serial sweep counts under a tolerance;
  - par_reduce = 1 (OpenMP reduction, the reference OpenMP program's order,
    mcf_4D_openmp.c:184-214) changes only the residual's last bits;
  - the threaded MCF step gives bitwise the u of the reference's own OpenMP
    program (mcf_4D_openmp.c:104-348, compiled into oracle/_ref).
"""
from ctypes import c_double, c_int

import numpy as np
import pytest

from harness import DP, RefMcf, params, ref_available, zeros

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0)


def _volume(shape, seed):
    rng = np.random.default_rng(seed)
    L, Kd, J, I = shape
    l, k, j, i = np.meshgrid(np.arange(L), np.arange(Kd), np.arange(J), np.arange(I),
                             indexing="ij")
    img = np.round(255 * rng.uniform(0, 1, shape) * 0.3 +
                   255 * 0.6 * (((i - I / 2) ** 2 / 30 + (j - J / 2) ** 2 / 20 +
                                 (k - Kd / 2) ** 2 / 10) < 1 + 0.2 * np.sin(l))) / 255
    u0 = 1.0 / (np.sqrt((i - I / 3) ** 2 + (j - J / 2) ** 2 + (k - Kd / 2) ** 2) + 1.0)
    for a in (img, u0):
        a[0] = a[-1] = 0
        a[:, 0] = a[:, -1] = 0
        a[:, :, 0] = a[:, :, -1] = 0
        a[..., 0] = a[..., -1] = 0
    return img, u0


def _edge(oracle, p, img):
    a, b, prev = img.copy(), np.zeros(img.shape), np.zeros(img.shape)
    g = zeros(img.shape, 8)
    it, res = oracle.gf(p, a, b, prev, g)
    return a, g, it, res


@pytest.mark.parametrize("par_reduce", [0, 2])
def test_threaded_segmentation_equals_serial(oracle, par_reduce):
    """Edge indicator and three tolerance-driven time steps: u, g, every
    residual and every sweep count bitwise those of the serial oracle."""
    shape = (10, 14, 22, 26)  # 24 x 20 x 12 x 8 interior
    img, u0 = _volume(shape, 5)
    ser = params(shape, gsTol=1e-9, maxNoGSIteration=5000, maxNoGSIteration_sigma=500.0, **KW)
    thr = params(shape, gsTol=1e-9, maxNoGSIteration=5000, maxNoGSIteration_sigma=500.0,
                 nthreads=8, par_reduce=par_reduce, **KW)
    a_s, g_s, it_s, res_s = _edge(oracle, ser, img)
    a_t, g_t, it_t, res_t = _edge(oracle, thr, img)
    assert it_s == it_t and res_s == res_t
    assert np.array_equal(a_s, a_t)
    for x, y in zip(g_s, g_t):
        assert np.array_equal(x, y)
    us, ut = u0.copy(), u0.copy()
    prev_s, prev_t = np.zeros(shape), np.zeros(shape)
    cs, ct = zeros(shape, 8), zeros(shape, 8)
    mask = np.ones(shape)
    for st in range(3):
        zs, Rs = oracle.seg_time_step(ser, us, prev_s, g_s, cs, mask)
        zt, Rt = oracle.seg_time_step(thr, ut, prev_t, g_t, ct, mask)
        assert zs == zt and Rs == Rt, f"step {st + 1}: {zs} / {zt} sweeps, R {Rs} / {Rt}"
        assert np.array_equal(us, ut), f"step {st + 1}"
        for x, y in zip(cs, ct):
            assert np.array_equal(x, y)


def test_timed_branch_equals_serial(oracle):
    """The CPU baseline's branch (nthreads 8, OpenMP reduction): bitwise the
    serial u after a fixed sweep count (the residual feeds only the stopping
    test), and the same sweep count under a tolerance."""
    shape = (10, 14, 22, 26)
    img, u0 = _volume(shape, 9)
    base = params(shape, **KW)
    _, g, _, _ = _edge(oracle, base, img)
    for gs, cap in ((-1.0, 7), (1e-9, 5000)):
        ser = params(shape, gsTol=gs, maxNoGSIteration=cap, **KW)
        thr = params(shape, gsTol=gs, maxNoGSIteration=cap, nthreads=8, par_reduce=1, **KW)
        us, ut = u0.copy(), u0.copy()
        cs, ct = zeros(shape, 8), zeros(shape, 8)
        oracle.coefficients(ser, us, g, cs)
        oracle.coefficients(thr, ut, g, ct)
        for x, y in zip(cs, ct):
            assert np.array_equal(x, y)
        zs, Rs = oracle.sor(ser, us, u0, cs)
        zt, Rt = oracle.sor(thr, ut, u0, ct)
        assert zs == zt
        if gs < 0:
            assert zs == cap
            assert np.array_equal(us, ut)
        assert Rt == pytest.approx(Rs, rel=1e-12)


@pytest.mark.skipif(not ref_available("libref_mcf_omp.so"),
                    reason="reference OpenMP MCF program not compiled (make -C oracle ref)")
@pytest.mark.parametrize("gstol,cap", [(-1.0, 6), (1e-12, 1000)])
def test_threaded_mcf_equals_reference_openmp(oracle, gstol, cap):
    """The oracle's threaded MCF step (OpenMP reduction) against the
    reference's OpenMP program driven through its globals: u bitwise after
    every step, with a fixed sweep count (gsTol < 0) and under the program's
    own tolerance."""
    m = 14  # 12^4 interior
    hh = 2.5 / (m - 2)
    ref = RefMcf("libref_mcf_omp.so")
    ref.seti(iMax=m, jMax=m, kMax=m, lMax=m, maxNoGSIteration=cap)
    ref.setd(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=gstol, omega_c=1.8, mcfTol=1e-12)
    assert ref.fn("_allocateMem_mcf")()
    try:
        exact = ref.fn("_exactSolution_mcf", restype=None, argtypes=[c_double])
        copy = ref.fn("_copy4DdataFromSrcToDest_mcf", restype=None, argtypes=[DP, DP])
        step = ref.fn("_mcfTimeStep")
        exact(0.0)
        copy(ref.ptr("exact_u"), ref.ptr("u"))
        shape = (m, m, m, m)
        p = params(shape, h=hh, tau=hh * hh, eps2=hh * hh, gsTol=gstol, omega_c=1.8,
                   maxNoGSIteration=cap, nthreads=8, par_reduce=1)
        u = np.zeros(shape)
        oracle.mcf_exact(p, u, 0.0)
        assert np.array_equal(u, ref.arr("u"))
        prev, coef = np.zeros(shape), zeros(shape, 8)
        for st in range(1, 4):
            c_int.in_dll(ref.L, "numberOfTimeStep").value = st
            assert step()
            z, _ = oracle.mcf_time_step(p, u, prev, coef, st)
            if gstol < 0:
                assert z == cap
            assert np.array_equal(u, ref.arr("u")), f"step {st}"
    finally:
        ref.fn("_deallocateMem_mcf")()
