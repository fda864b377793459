"""Failure paths of the C ABI on the GPU, every condition forced from the host
so that every kernel still runs to completion:

  - a NaN in g (and in the presmoothing input): the reference's stopping test
    `R > gsTol` is false for a NaN residual and its solve silently stops
    (segmentation4D.c:309, :405); the library stops the same way but returns
    false with the reason, and the next valid call works;
  - the sweep cap (maxNoGSIteration, segmentation4D.c:44, :309) with a
    tolerance that is never reached;
  - a step that fails after its set-up rotated the device buffers
    (seg4dInjectStepFailure): u keeps u^n on the device and the warm-start
    history is dropped, so the next step starts from u^n like the reference.
"""
import numpy as np
import pytest

from harness import params, zeros

FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")
pytestmark = pytest.mark.gpu

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0, maxNoGSIteration_sigma=1000.0)


def _random_volume(rng, shape, lo=0.0, hi=1.0):
    a = np.zeros(shape)
    a[1:-1, 1:-1, 1:-1, 1:-1] = rng.uniform(lo, hi, tuple(d - 2 for d in shape))
    return a


def _generated(lib, dims, seed):
    spec = lib.gen_spec(0, seed)
    spec.nx, spec.ny, spec.nz, spec.nt = dims
    spec.axes[0][0], spec.axes[0][1], spec.axes[0][2] = dims[0] / 4, dims[1] / 5, dims[2] / 4
    spec.axes[1][0], spec.axes[1][1], spec.axes[1][2] = dims[0] / 8, dims[1] / 8, dims[2] / 6
    shape = (dims[3] + 2, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    img, u0 = np.zeros(shape), np.zeros(shape)
    lib.generate(spec, img, u0)
    return shape, img, u0


def test_nan_in_g_reports_and_recovers(seglib, oracle):
    print("[gpu] failure: NaN in g", flush=True)
    nx, ny, nz, nt = 14, 11, 9, 4
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(31)
    u0 = _random_volume(rng, shape)
    gs = [_random_volume(rng, shape, 0.5, 1.0) for _ in range(8)]
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    try:
        lib.set(gsTol=1e-10, maxNoGSIteration=500, **KW)
        for f, g in zip(FACES, gs):
            lib.arr("g_" + f)[...] = g
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        bad = gs[0].copy()
        bad[2, 3, 4, 5] = np.nan
        lib.arr("g_e")[...] = bad
        lib.host_wrote("g_e")
        lib.arr("u")[...] = u0
        assert lib.lib._segTimeStep_() is False
        assert "non-finite" in lib.last_error()
        assert lib.lib.seg4dLastSweeps() == 1  # NaN > gsTol is false: one iteration
        # valid g again: the step from u0 is the oracle's, bit for bit
        lib.arr("g_e")[...] = gs[0]
        lib.host_wrote("g_e")
        lib.arr("u")[...] = u0
        z, _ = lib.seg_time_step()
        u, prev = u0.copy(), np.zeros(shape)
        zo, _ = oracle.seg_time_step(params(shape, gsTol=1e-10, maxNoGSIteration=500, **KW),
                                     u, prev, gs, zeros(shape, 8), np.ones(shape))
        assert z == zo
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.deallocate()


def test_nan_in_presmoothing_input(seglib, oracle):
    print("[gpu] failure: NaN in the presmoothing input", flush=True)
    nx, ny, nz, nt = 12, 10, 8, 4
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(32)
    img = np.round(_random_volume(rng, shape) * 255) / 255
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    try:
        lib.set(gsTol=1e-12, maxNoGSIteration=100, **KW)
        bad = img.copy()
        bad[1, 1, 1, 1] = np.nan
        lib.arr("i0")[...] = bad
        assert lib.lib._presmoothStep_(lib.ptr("i0")) is False
        assert "non-finite" in lib.last_error()
        lib.arr("i0")[...] = img
        z, R = lib.presmooth("i0")
        a, prev = img.copy(), np.zeros(shape)
        zo, Ro = oracle.presmooth(params(shape, gsTol=1e-12, **KW), a, prev)
        assert z == zo
        assert np.array_equal(lib.arr("i0"), a)
        assert np.array_equal(lib.arr("prev_u"), prev)
    finally:
        lib.deallocate()


@pytest.mark.parametrize("cap", [1, 3])
def test_sweep_cap_with_unreached_tolerance(seglib, oracle, cap):
    """gsTol far below what the cap allows: the solve stops at the cap, with
    the oracle's u."""
    print(f"[gpu] failure: sweep cap {cap}", flush=True)
    nx, ny, nz, nt = 13, 9, 7, 5
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(40 + cap)
    u0 = _random_volume(rng, shape)
    gs = [_random_volume(rng, shape, 0.1, 1.0) for _ in range(8)]
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    try:
        lib.set(gsTol=1e-300, maxNoGSIteration=cap, **KW)
        for f, g in zip(FACES, gs):
            lib.arr("g_" + f)[...] = g
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        lib.arr("u")[...] = u0
        z, R = lib.seg_time_step()
        u, prev = u0.copy(), np.zeros(shape)
        zo, Ro = oracle.seg_time_step(params(shape, gsTol=1e-300, maxNoGSIteration=cap, **KW),
                                      u, prev, gs, zeros(shape, 8), np.ones(shape))
        assert z == zo == cap
        assert R > 1e-300 and R == pytest.approx(Ro, rel=1e-12)
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.deallocate()


@pytest.mark.parametrize("where", [1, 2])
def test_failed_step_keeps_u_and_drops_history(seglib, oracle, where):
    """Resident mode, cubic warm start (three history levels in rotating
    buffers).  A step made to fail after its set-up (1) or at its final wait,
    after the new u went into the recycled buffer (2), returns false; the
    device u is still u^n, and the next step starts from u^n (the history is
    dropped): FP64, it equals the oracle's step from u^n bit for bit."""
    print(f"[gpu] failure: injected step failure {where}", flush=True)
    lib = seglib
    dims = (24, 20, 12, 6)
    shape, img, u0 = _generated(lib, dims, 17)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, maxNoGSIteration=100000, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in FACES]
        tol = 1e-12 * oracle.sumsq_interior(params(shape, **KW), u0)
        lib.set(gsTol=tol)
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.upload("Mmask")
        lib.set_warm_start(4)
        for _ in range(4):
            lib.seg_time_step()
        lib.download("u")
        un = lib.arr("u").copy()
        lib.inject_step_failure(where)
        assert lib.lib._segTimeStep_() is False
        assert "injected" in lib.last_error()
        lib.download("u")
        assert np.array_equal(lib.arr("u"), un)
        z, _ = lib.seg_time_step()  # still warm 4, but no history: starts from u^n
        lib.download("u")
        u, prev = un.copy(), np.zeros(shape)
        zo, _ = oracle.seg_time_step(params(shape, gsTol=tol, maxNoGSIteration=100000, **KW),
                                     u, prev, g, zeros(shape, 8), np.ones(shape))
        assert z == zo
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.set_warm_start(0)
        lib.set_resident(False)
        lib.deallocate()


@pytest.mark.parametrize("where", [1, 2])
def test_failed_mirror_step_keeps_host_arrays(seglib, oracle, where):
    """Mirror mode (the default), where the step copies u into prev_u on the
    host while the device computes and uploads Mmask beside it: a step made to
    fail after its set-up (1) or at its final wait (2) leaves the host u as it
    was, prev_u holds u^n (the reference's first statement, prev_u <- u, has
    run), and the next step equals the oracle's step from u^n."""
    print(f"[gpu] failure: injected mirror-mode step failure {where}", flush=True)
    lib = seglib
    dims = (24, 20, 12, 6)
    shape, img, u0 = _generated(lib, dims, 19)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, maxNoGSIteration=100000, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in FACES]
        tol = 1e-12 * oracle.sumsq_interior(params(shape, **KW), u0)
        lib.set(gsTol=tol)
        lib.arr("u")[...] = u0
        z1, _ = lib.seg_time_step()
        u1, p1 = lib.arr("u").copy(), lib.arr("prev_u").copy()
        assert np.array_equal(p1, u0)  # the reference's prev_u <- u
        lib.inject_step_failure(where)
        assert lib.lib._segTimeStep_() is False
        assert "injected" in lib.last_error()
        assert np.array_equal(lib.arr("u"), u1)
        assert np.array_equal(lib.arr("prev_u"), u1)
        z2, _ = lib.seg_time_step()
        u, prev = u1.copy(), np.zeros(shape)
        zo, _ = oracle.seg_time_step(params(shape, gsTol=tol, maxNoGSIteration=100000, **KW),
                                     u, prev, g, zeros(shape, 8), np.ones(shape))
        assert z2 == zo
        assert np.array_equal(lib.arr("u"), u)
        assert np.array_equal(lib.arr("prev_u"), u1)
    finally:
        lib.deallocate()
