"""GPU parity: the HIP path, called through the C ABI (libseg4d_amd.so,
libmcf4d_amd.so), against the reference's golden fixtures and the oracle.

Bar: bitwise equality of every FP64 array (colour passes, coefficients and
edge indicator evaluate the reference's expressions in the reference's
order; only the residual reductions are reordered, and they feed nothing but
the stopping test).  Where a tolerance-driven sweep count could in principle
flip on a last-bit difference of the residual, the test also records the
sweep counts and requires them equal.
"""
import os

import numpy as np
import pytest

from harness import GOLDEN, params, zeros

FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")
pytestmark = pytest.mark.gpu


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def seg_globals(fx):
    d = {k[2:]: fx[k].item() for k in fx.files if k.startswith("p_")}
    return d


def setup_seg(lib, fx):
    shape = fx["i0"].shape
    nt, nz, ny, nx = (d - 2 for d in shape)
    counts = [int(c) for c in fx["counts"]]
    lib.allocate(nx, ny, nz, nt, centres=counts)
    lib.set(ballRadius=float(fx["radius"]), **seg_globals(fx))
    cx, cy, cz = lib.centre_arrays()
    q = 0
    for l, c in enumerate(counts):
        for k in range(c):
            x, y, z = fx["centres"][q]
            cx[l][k], cy[l][k], cz[l][k] = float(x), float(y), float(z)
            q += 1
    return shape


@pytest.mark.parametrize("name", ["seg_h1.npz", "seg_h001.npz"])
def test_seg_golden_gpu(seglib, name):
    print(f"[gpu] golden {name}", flush=True)
    fx = load(name)
    lib = seglib
    setup_seg(lib, fx)
    try:
        lib.arr("i0")[...] = fx["i0"]
        lib.arr("u")[...] = fx["thr"]
        lib.arr("Mmask")[...] = fx["mask"]
        lib.gf("i0", "u")
        assert np.array_equal(lib.arr("i0"), fx["i0_s"])
        assert np.array_equal(lib.arr("u"), fx["thr_s"])
        assert np.array_equal(lib.arr("prev_u"), fx["prev_gf"])
        for f in FACES:
            assert np.array_equal(lib.arr("g_" + f), fx["g_" + f]), f
        lib.arr("u")[...] = fx["u0"]
        for st in range(int(fx["steps"])):
            lib.set(numberOfTimeStep=st + 1)
            lib.seg_time_step()
            assert np.array_equal(lib.arr("u"), fx[f"u{st + 1}"]), f"step {st + 1}"
            assert lib.l2norm("prev_u", "u") == pytest.approx(fx["l2"][st], rel=1e-13)
        assert np.array_equal(lib.arr("prev_u"), fx["prev_last"])
    finally:
        lib.deallocate()


def test_seg_golden_coefficients_gpu(seglib):
    """_gs_coefficients_ alone reproduces the reference's step-1 coefficients."""
    fx = load("seg_h1.npz")
    lib = seglib
    setup_seg(lib, fx)
    try:
        for f in FACES:
            lib.arr("g_" + f)[...] = fx["g_" + f]
        lib.arr("u")[...] = fx["u0"]
        lib.gs_coefficients()
        for f in FACES:
            assert np.array_equal(lib.arr(f), fx["coef1_" + f]), f
    finally:
        lib.deallocate()


def test_mcf_golden_gpu(mcflib):
    print("[gpu] golden mcf", flush=True)
    fx = load("mcf.npz")
    m, hh = int(fx["m"]), float(fx["h"])
    lib = mcflib
    lib.set(maxNoGSIteration=1000)
    lib.allocate(m - 2)
    lib.set(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8)
    try:
        lib.exact(0.0)
        lib.copy("exact_u", "u")
        assert np.array_equal(lib.arr("u"), fx["u0"])
        for st in range(1, int(fx["steps"]) + 1):
            lib.set(numberOfTimeStep=st)
            lib.time_step()
            assert np.array_equal(lib.arr("u"), fx[f"u{st}"]), st
            lib.exact(st * hh * hh)
            assert lib.l2norm("exact_u", "u") == pytest.approx(fx["err"][st - 1], rel=1e-13)
        assert np.array_equal(lib.arr("exact_u"), fx["exact_last"])
    finally:
        lib.deallocate()


@pytest.mark.parametrize("n", [9, 11])
def test_mcf_odd_sizes_vs_oracle(mcflib, oracle, n):
    """MCF steps on odd n (an odd number of frames: the operand blocks' last
    frame pair has one frame) and 5 steps: u after every step bitwise the
    oracle's, the error norm to 1e-13 of it (mcf_4D.c:102-205, 444-462)."""
    print(f"[gpu] mcf odd n = {n}", flush=True)
    lib = mcflib
    hh = 2.5 / n
    lib.set(maxNoGSIteration=1000)
    lib.allocate(n)
    lib.set(h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8)
    try:
        shape = (n + 2,) * 4
        p = params(shape, h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8,
                   maxNoGSIteration=1000)
        u, prev, coef, ex = np.zeros(shape), np.zeros(shape), zeros(shape, 8), np.zeros(shape)
        lib.exact(0.0)
        lib.copy("exact_u", "u")
        oracle.mcf_exact(p, u, 0.0)
        assert np.array_equal(lib.arr("u"), u)
        for st in range(1, 6):
            lib.set(numberOfTimeStep=st)
            z_g, _ = lib.time_step()
            z_o, _ = oracle.mcf_time_step(p, u, prev, coef, st)
            assert z_g == z_o, f"sweeps differ at step {st}"
            assert np.array_equal(lib.arr("u"), u), f"u differs at step {st}"
            lib.exact(st * hh * hh)
            oracle.mcf_exact(p, ex, st * hh * hh)
            assert lib.l2norm("exact_u", "u") == pytest.approx(oracle.l2norm(p, ex, u), rel=1e-13)
    finally:
        lib.deallocate()


def _random_volume(rng, shape):
    a = rng.uniform(0.0, 1.0, size=shape)
    a[0] = a[-1] = 0
    a[:, 0] = a[:, -1] = 0
    a[:, :, 0] = a[:, :, -1] = 0
    a[..., 0] = a[..., -1] = 0
    return a


@pytest.mark.parametrize("dims", [(9, 8, 7, 6), (16, 12, 10, 5), (3, 3, 3, 3), (1, 1, 1, 1),
                                  (33, 5, 4, 3)])
def test_seg_random_vs_oracle(seglib, oracle, dims):
    """Ragged and minimal sizes, tolerance-driven steps: bitwise u and equal
    sweep counts against the oracle."""
    nx, ny, nz, nt = dims
    print(f"[gpu] random seg {dims}", flush=True)
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(sum(dims))
    img = np.round(_random_volume(rng, shape) * 255) / 255
    u0 = _random_volume(rng, shape) + 0.0
    kw = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
              theta_img=1.0, rho_thr=0.0, maxNoGSIteration=400, maxNoGSIteration_sigma=60.0)
    p = params(shape, gsTol=1e-13, **kw)
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    try:
        lib.set(gsTol=1e-13, **kw)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        a, b, prev = img.copy(), np.zeros(shape), np.zeros(shape)
        g = zeros(shape, 8)
        oracle.gf(p, a, b, prev, g)
        assert np.array_equal(lib.arr("i0"), a)
        for f, gg in zip(FACES, g):
            assert np.array_equal(lib.arr("g_" + f), gg), f
        u = u0.copy()
        lib.arr("u")[...] = u0
        coef = zeros(shape, 8)
        mask = np.ones(shape)
        for st in range(3):
            z_o, _ = oracle.seg_time_step(p, u, prev, g, coef, mask)
            lib.set(numberOfTimeStep=st + 1)
            z_g, _ = lib.seg_time_step()
            assert z_g == z_o, f"sweeps differ at step {st + 1}"
            assert np.array_equal(lib.arr("u"), u), f"u differs at step {st + 1}"
            assert np.array_equal(lib.arr("prev_u"), prev)
    finally:
        lib.deallocate()


@pytest.mark.parametrize("sweeps,variant,chunk,depth",
                         [(1, 0, 0, 0), (2, 0, 0, 0), (7, 0, 0, 0), (7, 1, 0, 0), (7, 2, 0, 0),
                          (7, 3, 0, 0), (7, 7, 0, 0), (7, 11, 0, 0), (7, 15, 0, 0), (7, 19, 0, 0),
                          (7, 23, 0, 0), (7, 9, 0, 0), (7, 3, 100, 1), (7, 3, 2, 4), (7, 3, 3, 2),
                          (7, 2, 1, 16), (5, 11, 2, 3)])
def test_fixed_sweep_count(seglib, oracle, sweeps, variant, chunk, depth):
    """gsTol < 0 forces exactly maxNoGSIteration sweeps (sweep cap path);
    every SOR pass kernel variant gives the same bits (bits 8 / 16: two /
    four k planes per block; nz = 9 leaves the last block partly idle), so
    do frame chunks (0: automatic, 100: one chunk) and solve pipelines of
    any depth (0: automatic)."""
    nx, ny, nz, nt = 12, 10, 9, 4
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(sweeps)
    u0 = _random_volume(rng, shape)
    gs = [_random_volume(rng, shape) + 0.01 for _ in range(8)]
    kw = dict(h=1.0, tau=0.5, eps2=1e-6, omega_c=1.8, maxNoGSIteration=sweeps)
    p = params(shape, gsTol=-1.0, **kw)
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    lib.set_sweep_variant(variant)
    lib.lib.seg4dSetSweepChunk(chunk)
    lib.lib.seg4dSetPipelineDepth(depth)
    try:
        lib.set(gsTol=-1.0, **kw)
        for f, gg in zip(FACES, gs):
            lib.arr("g_" + f)[...] = gg
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        lib.arr("u")[...] = u0
        z_g, _ = lib.seg_time_step()
        u, prev = u0.copy(), np.zeros(shape)
        z_o, _ = oracle.seg_time_step(p, u, prev, gs, zeros(shape, 8), np.ones(shape))
        assert z_g == z_o == sweeps
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.set_sweep_variant(35)
        lib.lib.seg4dSetSweepChunk(0)
        lib.lib.seg4dSetPipelineDepth(0)
        lib.lib.seg4dSetSweepBands(0)
        lib.deallocate()


@pytest.mark.parametrize("variant,dims,chunk,depth,bands",
                         [(11, (17, 11, 7, 3), 0, 0, 0), (19, (17, 11, 7, 3), 0, 0, 0),
                          (10, (8, 6, 5, 4), 0, 0, 0), (15, (30, 9, 6, 5), 0, 0, 0),
                          (3, (31, 10, 7, 6), 100, 1, 0), (3, (31, 10, 7, 6), 2, 5, 0),
                          (3, (31, 10, 7, 6), 4, 2, 0)])
def test_tolerance_driven_variants(seglib, oracle, variant, dims, chunk, depth, bands):
    """Tolerance-driven steps (the fused residual and its per-block partials)
    with several k planes per block: same sweep count and bits as the oracle."""
    nx, ny, nz, nt = dims
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(variant)
    u0 = _random_volume(rng, shape)
    gs = [_random_volume(rng, shape) + 0.01 for _ in range(8)]
    kw = dict(h=1.0, tau=0.5, eps2=1e-6, omega_c=1.8, maxNoGSIteration=5000)
    p = params(shape, gsTol=1e-9, **kw)
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    lib.set_sweep_variant(variant)
    lib.lib.seg4dSetSweepChunk(chunk)
    lib.lib.seg4dSetPipelineDepth(depth)
    lib.lib.seg4dSetSweepBands(bands)
    try:
        lib.set(gsTol=1e-9, **kw)
        for f, gg in zip(FACES, gs):
            lib.arr("g_" + f)[...] = gg
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        lib.arr("u")[...] = u0
        u, prev = u0.copy(), np.zeros(shape)
        for st in range(2):
            z_g, _ = lib.seg_time_step()
            z_o, _ = oracle.seg_time_step(p, u, prev, gs, zeros(shape, 8), np.ones(shape))
            assert z_g == z_o
            assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.set_sweep_variant(35)
        lib.lib.seg4dSetSweepChunk(0)
        lib.lib.seg4dSetPipelineDepth(0)
        lib.lib.seg4dSetSweepBands(0)
        lib.deallocate()


def _smooth_volume(shape, centre):
    L, K, J, I = shape
    l, k, j, i = np.meshgrid(np.arange(L), np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    a = 1.0 / (np.sqrt((i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2) + 1.0)
    a[0] = a[-1] = 0
    a[:, 0] = a[:, -1] = 0
    a[:, :, 0] = a[:, :, -1] = 0
    a[..., 0] = a[..., -1] = 0
    return a


@pytest.mark.parametrize("variant,bands,depth", [(35, 0, 0), (35, 2, 0), (35, 3, 1), (43, 4, 2),
                                                 (51, 2, 0)])
def test_banded_order_vs_oracle(seglib, oracle, variant, bands, depth):
    """The automatic order's bands of runs (forced here; automatic only for
    planes too large for L2), also with two or four k planes per block:
    tolerance-driven steps with the oracle's sweep counts and bits."""
    print(f"[gpu] banded order variant {variant} bands {bands} depth {depth}", flush=True)
    nx, ny, nz, nt = 61, 40, 9, 5
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(bands + 10 * variant)
    u0 = _smooth_volume(shape, (nx / 2, ny / 3, nz / 2))
    gs = []
    for _ in range(8):
        g = np.zeros(shape)
        g[1:-1, 1:-1, 1:-1, 1:-1] = rng.uniform(0.5, 1.0, (nt, nz, ny, nx))
        gs.append(g)
    kw = dict(h=1.0, tau=0.5, eps2=1e-6, omega_c=1.8, maxNoGSIteration=5000)
    p = params(shape, gsTol=1e-9, **kw)
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    lib.set_sweep_variant(variant)
    lib.lib.seg4dSetPipelineDepth(depth)
    lib.lib.seg4dSetSweepBands(bands)
    try:
        lib.set(gsTol=1e-9, **kw)
        for f, gg in zip(FACES, gs):
            lib.arr("g_" + f)[...] = gg
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        lib.arr("u")[...] = u0
        u, prev = u0.copy(), np.zeros(shape)
        coef = zeros(shape, 8)
        for st in range(2):
            z_g, _ = lib.seg_time_step()
            z_o, _ = oracle.seg_time_step(p, u, prev, gs, coef, np.ones(shape))
            assert z_g == z_o < 5000
            assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.set_sweep_variant(35)
        lib.lib.seg4dSetPipelineDepth(0)
        lib.lib.seg4dSetSweepBands(0)
        lib.deallocate()


@pytest.mark.parametrize("dims", [(9, 8, 7, 6), (1, 1, 1, 1), (33, 5, 4, 3)])
def test_reflection_entry_point(seglib, oracle, dims):
    """_reflection4Ddata_ called on its own (mirror mode, host array in and
    out): every ghost equals the oracle's sequential X, Y, Z, W copies, and
    the interior is untouched, bitwise."""
    nx, ny, nz, nt = dims
    print(f"[gpu] _reflection4Ddata_ {dims}", flush=True)
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(sum(dims) + 101)
    a = rng.uniform(-1.0, 1.0, size=shape)
    lib = seglib
    lib.allocate(nx, ny, nz, nt)
    try:
        lib.arr("i0")[...] = a
        lib.reflect("i0")
        ref = a.copy()
        oracle.reflection(params(shape), ref)
        assert np.array_equal(lib.arr("i0").view(np.uint64), ref.view(np.uint64))
        assert np.array_equal(lib.arr("i0")[1:-1, 1:-1, 1:-1, 1:-1], a[1:-1, 1:-1, 1:-1, 1:-1])
    finally:
        lib.deallocate()
