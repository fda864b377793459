"""FP32-refined form (FP64 defect correction around FP32 SOR) against the FP64
oracle.  Not bitwise by design: every step must pass the reference's
stopping test on the FP64 residual, and the trajectory must stay within the
stated error of the FP64 oracle with a mask IoU >= 0.999.

Tolerance: both solutions satisfy ||b - A u|| <= 1e-6 ||b|| per step, so
they differ by O(1e-6) relative per step; after 6 steps the test allows
max|u - u_FP64| <= 1e-4 * max|u_FP64|."""
import numpy as np
import pytest

from harness import params, zeros

pytestmark = pytest.mark.gpu

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0, maxNoGSIteration=100000, maxNoGSIteration_sigma=100000.0)
REL = 1e-6


def _volume(lib, dims, seed):
    spec = lib.gen_spec(0, seed)
    spec.nx, spec.ny, spec.nz, spec.nt = dims
    spec.axes[0][0], spec.axes[0][1], spec.axes[0][2] = dims[0] / 4, dims[1] / 5, dims[2] / 4
    spec.axes[1][0], spec.axes[1][1], spec.axes[1][2] = dims[0] / 8, dims[1] / 8, dims[2] / 6
    shape = (dims[3] + 2, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    img, u0 = np.zeros(shape), np.zeros(shape)
    lib.generate(spec, img, u0)
    return shape, img, u0


@pytest.mark.parametrize("dims,inner_omega", [((40, 36, 20, 6), -1.0), ((33, 17, 12, 5), -1.0),
                                              ((40, 36, 20, 6), 0.0), ((33, 17, 12, 5), 1.5)])
def test_fp32_refined_vs_oracle(seglib, oracle, dims, inner_omega):
    """inner_omega: SOR factor of the correction solves (-1 keeps the
    default 1.2, 0 takes omega_c = 1.8)."""
    print(f"[gpu] fp32-refined {dims} inner omega {inner_omega}", flush=True)
    lib = seglib
    shape, img, u0 = _volume(lib, dims, 3)
    lib.allocate(*dims)
    om_keep = lib.lib.seg4dSetInnerOmega(inner_omega)
    assert om_keep == 1.2
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(REL)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        lib.set_precision(1)
        lib.arr("u")[...] = u0
        u = u0.copy()
        prev = np.zeros(shape)
        coef = zeros(shape, 8)
        for st in range(6):
            # oracle FP64 step with the same relative bound (sequential sum)
            p = params(shape, gsTol=REL * REL * oracle.sumsq_interior(params(shape, **KW), u), **KW)
            oracle.seg_time_step(p, u, prev, g, coef, np.ones(shape))
            lib.set(numberOfTimeStep=st + 1)
            z, R = lib.seg_time_step()
            rounds = lib.last_rounds()
            tol = lib.lib.seg4dLastTolerance()
            assert R <= tol, f"step {st + 1}: FP64 residual {R} above bound {tol}"
            assert 1 <= rounds <= 8
            ug = lib.arr("u")
            err = np.abs(ug - u).max()
            scale = np.abs(u).max()
            assert err <= 1e-4 * scale, f"step {st + 1}: max|du| {err} vs {scale}"
            iou = oracle.iou(params(shape), oracle.mask(params(shape), u),
                             oracle.mask(params(shape), np.ascontiguousarray(ug)))
            assert iou >= 0.999, f"step {st + 1}: IoU {iou}"
            print(f"  step {st + 1}: fp32 sweeps {z}, rounds {rounds}, max|du| {err:.2e}, IoU {iou}",
                  flush=True)
    finally:
        lib.lib.seg4dSetInnerOmega(om_keep)
        lib.set_precision(0)
        lib.set_relative_tolerance(0.0)
        lib.deallocate()


@pytest.mark.parametrize("precision", [0, 1])
def test_warm_mode_switches_keep_history(seglib, oracle, precision):
    """The history levels and u live in buffers that rotate by pointer after
    each set-up (the oldest level's buffer receives u^(n+1)); which buffers
    rotate depends on the warm-start mode.  Switching modes between steps in
    resident mode: after every step prev_u holds u^n bitwise (ghosts
    included), and an FP64 step with mode 0 (start u^n) equals the oracle's
    step from the device's u^n bitwise."""
    lib = seglib
    dims = (24, 20, 12, 6)
    shape, img, u0 = _volume(lib, dims, 11)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        tol = REL * REL * oracle.sumsq_interior(params(shape, **KW), u0)
        lib.set(gsTol=tol, **KW)
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.upload("Mmask")
        lib.set_precision(precision)
        modes = [4, 4, 4, 0, 1, 1, 0, 3, 4, 4, 0, 1, 0]
        exact = 0
        for st, warm in enumerate(modes):
            lib.set_warm_start(warm)
            lib.download("u")
            before = lib.arr("u").copy()
            z, R = lib.seg_time_step()
            lib.download("u")
            lib.download("prev_u")
            assert np.array_equal(lib.arr("prev_u"), before), f"step {st} (warm {warm}): prev_u"
            if precision == 0 and warm == 0:
                uo, prev, coef = before.copy(), np.zeros(shape), zeros(shape, 8)
                zo = oracle.seg_time_step(params(shape, gsTol=tol, **KW), uo, prev, g, coef,
                                          np.ones(shape))[0]
                assert z == zo, f"step {st}: sweeps {z} vs oracle {zo}"
                assert np.array_equal(lib.arr("u"), uo), f"step {st}: u after a mode-0 step"
                exact += 1
        assert precision == 1 or exact == modes.count(0)
    finally:
        lib.set_warm_start(0)
        lib.set_precision(0)
        lib.set_resident(False)
        lib.deallocate()


@pytest.mark.parametrize("precision,warm", [(0, 1), (1, 1), (0, 3), (1, 3), (0, 4), (1, 4)])
def test_extrapolated_start_vs_oracle(seglib, oracle, precision, warm):
    """SOR started from 2 u^n - u^(n-1) (warm 1) or 3 u^n - 3 u^(n-1) +
    u^(n-2) (warm 3) in resident mode: same stopping test, trajectory within
    the stated error of the FP64 oracle, fewer sweeps."""
    lib = seglib
    dims = (40, 36, 20, 6)
    shape, img, u0 = _volume(lib, dims, 7)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(REL)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.upload("Mmask")
        lib.set_precision(precision)
        lib.set_warm_start(warm)
        u, prev, coef = u0.copy(), np.zeros(shape), zeros(shape, 8)
        cold, zs = [], []
        for st in range(6):
            p = params(shape, gsTol=REL * REL * oracle.sumsq_interior(params(shape, **KW), u), **KW)
            cold.append(oracle.seg_time_step(p, u, prev, g, coef, np.ones(shape))[0])
            z, R = lib.seg_time_step()
            zs.append(z)
            assert R <= lib.lib.seg4dLastTolerance()
        lib.download("u")
        ug = lib.arr("u")
        assert np.abs(ug - u).max() <= 1e-4 * np.abs(u).max()
        iou = oracle.iou(params(shape), oracle.mask(params(shape), u),
                         oracle.mask(params(shape), np.ascontiguousarray(ug)))
        assert iou >= 0.999
        print(f"  precision {precision}, warm {warm}: oracle sweeps {cold}, extrapolated start {zs}",
              flush=True)
        if precision == 0:
            assert sum(zs[2:]) < sum(cold[2:])
    finally:
        lib.set_warm_start(0)
        lib.set_precision(0)
        lib.set_resident(False)
        lib.set_relative_tolerance(0.0)
        lib.deallocate()


def test_fp32_refined_round_cap_and_fp64_switch(seglib, oracle):
    """A round cap of 1 still returns a correction; switching back to FP64
    restores bitwise parity with the oracle on the next step."""
    lib = seglib
    dims = (20, 18, 10, 4)
    shape, img, u0 = _volume(lib, dims, 5)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=1e-30, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        lib.arr("u")[...] = u0
        lib.set_precision(1)
        lib.set_refinement(1, 0)
        z, R = lib.seg_time_step()
        assert lib.last_rounds() == 1 and z > 0
        lib.set_refinement(8, 0)
        lib.set_precision(0)
        state = lib.arr("u").copy()
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        lib.set(gsTol=1e-13)
        z_g, _ = lib.seg_time_step()
        p = params(shape, gsTol=1e-13, **KW)
        u, prev = state.copy(), np.zeros(shape)
        z_o, _ = oracle.seg_time_step(p, u, prev, g, zeros(shape, 8), np.ones(shape))
        assert z_g == z_o
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.set_precision(0)
        lib.deallocate()
