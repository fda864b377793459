"""Fused sweeps (seg4dSetFusedSweeps): a whole SOR solve in one cooperative
launch, its red passes, black passes and stopping tests separated by
grid-wide barriers.  Forced on, the solve must give bitwise the oracle's u
and sweep counts after every step (tolerance-driven and at the sweep cap),
and bitwise the pass-per-launch solve in every form that runs SOR solves:
FP64 with a warm start, the FP32-refined form (whose correction solves run
fused), other launch orders of the pass."""
import numpy as np
import pytest

from harness import params, zeros

pytestmark = pytest.mark.gpu

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0, maxNoGSIteration=400, maxNoGSIteration_sigma=60.0)
FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")


def _vol(rng, shape):
    a = rng.uniform(0.0, 1.0, size=shape)
    a[0] = a[-1] = 0
    a[:, 0] = a[:, -1] = 0
    a[:, :, 0] = a[:, :, -1] = 0
    a[..., 0] = a[..., -1] = 0
    return a


def _inputs(dims, seed):
    nx, ny, nz, nt = dims
    shape = (nt + 2, nz + 2, ny + 2, nx + 2)
    rng = np.random.default_rng(seed)
    img = np.round(_vol(rng, shape) * 255) / 255
    return img, _vol(rng, shape)


def _run(lib, dims, img, u0, fused, steps, precision=0, warm=0, variant=None, gsTol=1e-13):
    lib.allocate(*dims)
    try:
        lib.set(gsTol=gsTol, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        lib.arr("u")[...] = u0
        lib.set_fused_sweeps(fused)
        if variant is not None:
            lib.set_sweep_variant(variant)
        lib.set_warm_start(warm)
        out = []
        for st in range(steps):
            lib.set_precision(precision[st] if isinstance(precision, (list, tuple)) else precision)
            lib.set(numberOfTimeStep=st + 1)
            z, _ = lib.seg_time_step()
            out.append((lib.arr("u").copy(), z, lib.last_rounds(), lib.last_solve_fused()))
    finally:
        lib.set_warm_start(0)
        lib.set_precision(0)
        lib.deallocate()
    return out


@pytest.mark.parametrize("dims", [(24, 20, 12, 6), (33, 5, 4, 3), (1, 1, 1, 1), (64, 64, 32, 8)])
def test_fused_fp64_vs_oracle(seglib, oracle, dims):
    """Forced fused solves, three tolerance-driven steps: bitwise u and equal
    sweep counts against the oracle after every step."""
    print(f"[gpu] fused FP64 vs oracle {dims}", flush=True)
    img, u0 = _inputs(dims, sum(dims) + 7)
    shape = img.shape
    got = _run(seglib, dims, img, u0, 1, 3)
    p = params(shape, gsTol=1e-13, **KW)
    a, b, prev = img.copy(), np.zeros(shape), np.zeros(shape)
    g = zeros(shape, 8)
    oracle.gf(p, a, b, prev, g)
    u = u0.copy()
    coef = zeros(shape, 8)
    for st in range(3):
        z_o, _ = oracle.seg_time_step(p, u, prev, g, coef, np.ones(shape))
        ug, zg, _, fused = got[st]
        assert fused == 1, "the solve did not run fused"
        assert zg == z_o, f"step {st + 1}: sweeps {zg} vs oracle {z_o}"
        assert np.array_equal(ug.view(np.uint64), u.view(np.uint64)), \
            f"step {st + 1}: max |du| {np.abs(ug - u).max():.3e}"
    print(f"  sweeps {[r[1] for r in got]}", flush=True)


def test_fused_sweep_cap(seglib, oracle):
    """gsTol < 0: exactly maxNoGSIteration sweeps, fused, bitwise the oracle."""
    dims = (12, 10, 9, 4)
    img, u0 = _inputs(dims, 5)
    shape = img.shape
    kw = dict(KW)
    kw["maxNoGSIteration"] = 7
    lib = seglib
    lib.allocate(*dims)
    try:
        lib.set(gsTol=-1.0, **kw)
        rng = np.random.default_rng(3)
        gs = [_vol(rng, shape) + 0.01 for _ in range(8)]
        for f, gg in zip(FACES, gs):
            lib.arr("g_" + f)[...] = gg
        lib.host_wrote("g_e")
        lib.arr("Mmask")[...] = 1.0
        lib.arr("u")[...] = u0
        lib.set_fused_sweeps(1)
        z_g, _ = lib.seg_time_step()
        assert lib.last_solve_fused() == 1
        u, prev = u0.copy(), np.zeros(shape)
        p = params(shape, gsTol=-1.0, **kw)
        z_o, _ = oracle.seg_time_step(p, u, prev, gs, zeros(shape, 8), np.ones(shape))
        assert z_g == z_o == 7
        assert np.array_equal(lib.arr("u"), u)
    finally:
        lib.deallocate()


@pytest.mark.parametrize("precision,warm,variant", [(0, 3, None), (0, 0, 11), (1, 0, None),
                                                    (1, 4, None)])
def test_fused_matches_pass_per_launch(seglib, precision, warm, variant):
    """Fused and pass-per-launch solves give the same bits and sweep counts in
    FP64 with the quadratic start, in another launch order (variant 11: tiled
    order, two k planes per block) and in the FP32-refined form from u^n and
    from the cubic start (five steps each)."""
    print(f"[gpu] fused vs passes precision {precision} warm {warm} variant {variant}",
          flush=True)
    dims = (40, 33, 16, 7)
    img, u0 = _inputs(dims, 21)
    on = _run(seglib, dims, img, u0, 1, 5, precision, warm, variant)
    off = _run(seglib, dims, img, u0, 0, 5, precision, warm, variant)
    for s, (a, b) in enumerate(zip(on, off)):
        assert a[3] == 1 and b[3] == 0
        assert a[1] == b[1], f"step {s + 1}: sweeps {a[1]} vs {b[1]}"
        assert a[2] == b[2], f"step {s + 1}: rounds {a[2]} vs {b[2]}"
        assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64)), f"step {s + 1}"
    print(f"  sweeps {[r[1] for r in on]}", flush=True)


def test_fused_precision_switches(seglib):
    """FP32, FP64 and FP32 again within one allocation (FP32 first, so the
    FP64 solve sizes its own cooperative grid afterwards): every solve fused,
    bitwise the pass-per-launch steps."""
    print("[gpu] fused solves across precision switches", flush=True)
    dims = (40, 33, 16, 7)
    img, u0 = _inputs(dims, 23)
    seq = [1, 0, 0, 1, 0]
    on = _run(seglib, dims, img, u0, 1, len(seq), seq)
    off = _run(seglib, dims, img, u0, 0, len(seq), seq)
    for s, (a, b) in enumerate(zip(on, off)):
        assert a[3] == 1, f"step {s + 1} did not run fused"
        assert a[1] == b[1] and a[2] == b[2], f"step {s + 1}"
        assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64)), f"step {s + 1}"
