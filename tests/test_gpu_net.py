"""Learned warm start: the BF16 MFMA convolution kernels against a plain
PyTorch FP32 reference of the same network (1, 2 and 4 layers), training on
solver runs with the linear and the cubic base, and the solver started from
the network's prediction against the FP64 oracle.

Tolerances.  The kernels store activations and gradients in BF16 (8-bit
mantissa, relative rounding 2^-9) and accumulate in FP32; the reference
rounds the forward activations at the same points, so the forward output
agrees to 5e-3 of its largest value (only FP32 summation order and rare
BF16 rounding-boundary flips differ).  Gradients additionally round every
backward activation to BF16: each parameter group agrees to a relative L2
error of 2e-2.  The solver run has the stopping test of the plain method, so
it is held to the tolerances of the other warm starts (max|u - u_FP64| <=
1e-4 max|u_FP64|, mask IoU >= 0.999)."""
import numpy as np
import pytest

from harness import params, zeros

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
F = torch.nn.functional

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0, maxNoGSIteration=100000, maxNoGSIteration_sigma=100000.0)


def layout(layers=4):
    o, off = 0, {}
    for L in range(layers):
        cin = 8 if L == 0 else 16
        off[("ws", L)] = (o, 28 * cin * 16)
        o += 28 * cin * 16
        off[("wt", L)] = (o, 4 * 256)
        o += 4 * 256
        off[("b", L)] = (o, 16)
        o += 16
    off["wp"] = (o, 16)
    o += 16
    off["bp"] = (o, 1)
    o += 1
    return off, o


def bf(x):
    return x.to(torch.bfloat16).to(torch.float32)


def to_bits(x):
    """float32 array -> BF16 bit patterns (round to nearest even)."""
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy().view(np.uint16)


def reference(p, feat, P, layers=4):
    """Forward pass in FP32 with the kernels' BF16 rounding points.  p: torch
    parameter vector; feat: (L, K, J, I, 8) torch float32 (BF16 values).
    Returns out on the interior points, shape (B, nl, nk, nj, ni)."""
    off, _ = layout(layers)
    L_, K_, J_, I_, _ = feat.shape
    B = L_ // P
    x = feat.reshape(B, P, K_, J_, I_, 8)[:, 1:-1, 1:-1, 1:-1, 1:-1, :]
    nl, nk, nj, ni = x.shape[1:5]
    for layer in range(layers):
        cin = 8 if layer == 0 else 16
        o, n = off[("ws", layer)]
        ws = bf(p[o:o + 27 * cin * 16]).view(3, 3, 3, cin, 16).permute(4, 3, 0, 1, 2)
        xb = x.permute(0, 1, 5, 2, 3, 4).reshape(B * nl, cin, nk, nj, ni)
        s = bf(F.conv3d(xb, ws, padding=1))
        o, n = off[("wt", layer)]
        wt = bf(p[o:o + 3 * 256]).view(3, 16, 16).permute(2, 1, 0)
        o, n = off[("b", layer)]
        bias = p[o:o + 16]
        s5 = s.view(B, nl, 16, nk, nj, ni).permute(0, 3, 4, 5, 2, 1).reshape(-1, 16, nl)
        y = bf(torch.relu(F.conv1d(s5, wt, padding=1) + bias.view(1, 16, 1)))
        x = y.view(B, nk, nj, ni, 16, nl).permute(0, 5, 1, 2, 3, 4)
    o, _ = off["wp"]
    ob, _ = off["bp"]
    return (x * p[o:o + 16]).sum(-1) + p[ob]


def random_params(rng, scale=1.0, layers=4):
    off, n = layout(layers)
    w = np.zeros(n, dtype=np.float32)
    for L in range(layers):
        cin = 8 if L == 0 else 16
        o, m = off[("ws", L)]
        blk = rng.normal(0, np.sqrt(2.0 / (27 * cin)), size=(28, cin, 16)).astype(np.float32)
        blk[27] = 0
        w[o:o + m] = blk.ravel() * scale
        o, m = off[("wt", L)]
        blk = rng.normal(0, np.sqrt(2.0 / 48), size=(4, 16, 16)).astype(np.float32)
        blk[3] = 0
        w[o:o + m] = blk.ravel() * scale
        o, m = off[("b", L)]
        w[o:o + m] = rng.normal(0, 0.05, size=16)
    o, m = off["wp"]
    w[o:o + m] = rng.normal(0, 0.3, size=16)
    w[off["bp"][0]] = 0.01
    return w


def batch(rng, dims, P, B):
    I_, J_, K_ = dims
    L_ = B * P
    f = np.zeros((L_, K_, J_, I_, 8), dtype=np.float32)
    for b in range(B):
        blk = f[b * P + 1:(b + 1) * P - 1, 1:-1, 1:-1, 1:-1]
        blk[..., 0] = rng.uniform(0, 1, size=blk.shape[:-1])
        blk[..., 1] = rng.normal(0, 1, size=blk.shape[:-1])
        blk[..., 2] = rng.uniform(0, 1, size=blk.shape[:-1])
        blk[..., 3] = rng.normal(0, 1, size=blk.shape[:-1])  # the cubic base's feature
    bits = to_bits(f)
    fr = torch.from_numpy(bits.view(np.int16).copy()).view(torch.bfloat16).to(torch.float32)
    return bits, fr


def interior(a, P):
    L_, K_, J_, I_ = a.shape
    return a.reshape(L_ // P, P, K_, J_, I_)[:, 1:-1, 1:-1, 1:-1, 1:-1]


@pytest.mark.parametrize("dims,P,B,layers", [((12, 11, 10), 6, 2, 4), ((35, 9, 8), 5, 1, 4),
                                             ((21, 10, 9), 7, 2, 1), ((17, 13, 6), 4, 1, 2)])
def test_forward_vs_torch(seglib, dims, P, B, layers):
    """Inference path: one fused kernel per layer, the last one with the
    projection in its epilogue."""
    print(f"[gpu] net forward {dims} P={P} B={B} layers={layers}", flush=True)
    rng = np.random.default_rng(1)
    lib = seglib
    lib.net_init(3, layers=layers, base=3)
    assert lib.net_shape() == (layers, 3)
    w = random_params(rng, layers=layers)
    lib.net_set_params(w)
    bits, fr = batch(rng, dims, P, B)
    out = lib.net_forward(bits, P)
    ref = reference(torch.from_numpy(w), fr, P, layers).numpy()
    got = interior(out, P)
    err = np.abs(got - ref).max()
    print(f"  max|out| {np.abs(ref).max():.4f}, max err {err:.2e}", flush=True)
    assert err <= 5e-3 * np.abs(ref).max()
    # ghost points are never written
    full = np.zeros_like(out)
    full.reshape(out.shape[0] // P, P, *out.shape[1:])[:, 1:-1, 1:-1, 1:-1, 1:-1] = got
    assert np.array_equal(out, full)


@pytest.mark.parametrize("layers", [4, 1])
def test_gradient_vs_torch(seglib, layers):
    print(f"[gpu] net gradient layers={layers}", flush=True)
    rng = np.random.default_rng(2)
    lib = seglib
    lib.net_init(4, layers=layers, base=1 if layers == 4 else 3)
    w = random_params(rng, layers=layers)
    lib.net_set_params(w)
    dims, P, B = (13, 12, 11), 6, 2
    bits, fr = batch(rng, dims, P, B)
    L_ = B * P
    target = np.zeros((L_, dims[2], dims[1], dims[0]), dtype=np.float32)
    lmask = np.zeros(target.shape, dtype=np.uint8)
    ti = interior(target, P)
    ti[...] = rng.normal(0, 1, size=ti.shape)
    mi = interior(lmask, P)
    mi[..., 2:-2, 2:-2, 2:-2] = 1  # a core, as in training
    g, loss = lib.net_grad(bits, target, lmask, P)
    pt = torch.from_numpy(w.copy()).requires_grad_(True)
    out = reference(pt, fr, P, layers)
    tt = torch.from_numpy(np.ascontiguousarray(ti))
    mt = torch.from_numpy(np.ascontiguousarray(mi)).bool()
    d = (out - tt)[mt]
    lref = (d * d).mean()
    lref.backward()
    gref = pt.grad.numpy()
    print(f"  loss {loss:.6f} vs {lref.item():.6f}", flush=True)
    assert loss == pytest.approx(lref.item(), rel=2e-2)
    off, _ = layout(layers)
    for key, (o, n) in off.items():
        a, b = g[o:o + n], gref[o:o + n]
        nb = np.linalg.norm(b)
        rel = np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)
        print(f"  {key}: rel L2 err {rel:.3e}", flush=True)
        assert rel <= 2e-2, key


def _volume(lib, dims, seed):
    spec = lib.gen_spec(0, seed)
    spec.nx, spec.ny, spec.nz, spec.nt = dims
    spec.axes[0][0], spec.axes[0][1], spec.axes[0][2] = dims[0] / 4, dims[1] / 5, dims[2] / 4
    spec.axes[1][0], spec.axes[1][1], spec.axes[1][2] = dims[0] / 8, dims[1] / 8, dims[2] / 6
    shape = (dims[3] + 2, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    img, u0 = np.zeros(shape), np.zeros(shape)
    lib.generate(spec, img, u0)
    return shape, img, u0


@pytest.mark.parametrize("layers,base", [(4, 1), (1, 3), (4, 0)])
def test_train_on_solver_runs(seglib, tmp_path, layers, base):
    """Pool two short runs, train, and check that the loss falls below that
    of the base extrapolation (the untrained network's prediction, loss 1 by
    the choice of the target scale); save / load round trip."""
    print(f"[gpu] net training layers={layers} base={base}", flush=True)
    lib = seglib
    lib.net_init(5, layers=layers, base=base)
    dims = (72, 72, 72, 4)
    for seed in (1, 2):
        shape, img, u0 = _volume(lib, dims, seed)
        lib.allocate(*dims)
        try:
            lib.set(gsTol=0.0, **KW)
            lib.set_relative_tolerance(1e-6)
            lib.arr("i0")[...] = img
            lib.arr("u")[...] = 0.0
            lib.arr("Mmask")[...] = 1.0
            lib.gf("i0", "u")
            lib.set_resident(True)
            lib.arr("u")[...] = u0
            lib.upload("u")
            lib.upload("Mmask")
            lib.net_record(new_run=True)
            for _ in range(6):
                lib.seg_time_step()
                lib.net_record()
        finally:
            lib.set_resident(False)
            lib.set_relative_tolerance(0.0)
            lib.deallocate()
    # levels base..5 of each run's 7 (the base's history and a successor)
    assert lib.net_pool_pairs() == 2 * (6 - base)
    loss = lib.net_train(60, batch=2, lr=1e-3, seed=1, log_every=20)
    s_d, s_t = lib.net_scales()
    print(f"  tail loss {loss:.4f}, scales {s_d:.3g} {s_t:.3g}", flush=True)
    assert np.isfinite(loss) and loss < (0.9 if base == 1 else 1.0)
    w = lib.net_params()
    path = tmp_path / "net.bin"
    lib.net_save(path)
    lib.net_init(9)
    lib.net_load(path)
    assert lib.net_shape() == (layers, base)
    assert np.array_equal(lib.net_params(), w)
    assert lib.net_scales() == (s_d, s_t)


@pytest.mark.parametrize("precision,layers,base", [(0, 4, 1), (1, 4, 1), (0, 1, 3), (1, 2, 3),
                                                   (1, 4, 0)])
def test_learned_start_vs_oracle(seglib, oracle, precision, layers, base):
    """SOR started from the network's prediction (random weights with a
    non-zero projection, so the start differs from every other mode): same
    stopping test, trajectory within the stated error of the FP64 oracle.
    The cubic base predicts once three completed steps precede, the base
    u^n (inputs u^n, g) from the first step on."""
    print(f"[gpu] learned start precision={precision} layers={layers} base={base}", flush=True)
    lib = seglib
    rng = np.random.default_rng(7)
    lib.net_init(7, layers=layers, base=base)
    w = random_params(rng, 0.5, layers)
    off, _ = layout(layers)
    o, m = off["wp"]
    w[o:o + m] *= 0.01
    lib.net_set_params(w, 10.0, 100.0)
    dims = (40, 36, 20, 6)
    shape, img, u0 = _volume(lib, dims, 7)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(1e-6)
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
        assert lib.set_warm_start(2) == 0
        u, prev, coef = u0.copy(), np.zeros(shape), zeros(shape, 8)
        zs = []
        ms0, n0 = lib.net_predict_stats()
        steps = 4 if base == 1 else 6
        for st in range(steps):
            p = params(shape, gsTol=1e-12 * oracle.sumsq_interior(params(shape, **KW), u), **KW)
            oracle.seg_time_step(p, u, prev, g, coef, np.ones(shape))
            z, R = lib.seg_time_step()
            zs.append(z)
            assert R <= lib.lib.seg4dLastTolerance()
        ms, n = lib.net_predict_stats()
        ms, n = ms - ms0, n - n0
        # from the step after the base's history exists: 2..4 (linear), 4..6
        # (cubic), every step (base u^n)
        assert n == (steps if base == 0 else 3)
        lib.download("u")
        ug = lib.arr("u")
        assert np.abs(ug - u).max() <= 1e-4 * np.abs(u).max()
        iou = oracle.iou(params(shape), oracle.mask(params(shape), u),
                         oracle.mask(params(shape), np.ascontiguousarray(ug)))
        print(f"  sweeps {zs}, IoU {iou}, predict {ms / n:.3f} ms", flush=True)
        assert iou >= 0.999
    finally:
        lib.set_warm_start(0)
        lib.set_precision(0)
        lib.set_resident(False)
        lib.set_relative_tolerance(0.0)
        lib.deallocate()


def test_prediction_vs_torch(seglib):
    """The prediction path of the cubic-base network on a solver state: the
    features (u^n, the increment, mean g, the third difference), their
    scales (sigma, kappa from device reductions, folded into the first
    layer's BF16 weights), the fused one-layer network and its projection,
    against numpy features and the PyTorch reference network with the same
    BF16 rounding points (5e-3 of max|c|, as the forward test), and the
    scale kappa against numpy (rel 1e-10)."""
    print("[gpu] learned prediction vs torch", flush=True)
    lib = seglib
    rng = np.random.default_rng(11)
    layers = 1
    lib.net_init(3, layers=layers, base=3)
    w = random_params(rng, 0.5, layers)
    s_d, s_t = 0.7, 3.0
    lib.net_set_params(w, s_d, s_t)
    dims = (37, 30, 18, 6)
    shape, img, u0 = _volume(lib, dims, 5)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(1e-6)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.upload("Mmask")
        assert lib.set_warm_start(2) == 0
        levels = [u0.copy()]
        for _ in range(3):  # the cubic base needs u^(n-3): three completed steps
            lib.seg_time_step()
            lib.download("u")
            levels.append(lib.arr("u").copy())
        c, q = lib.net_preview(shape)
    finally:
        lib.set_warm_start(0)
        lib.set_resident(False)
        lib.set_relative_tolerance(0.0)
        lib.deallocate()
    u3, u2, u1, u0_ = levels[3], levels[2], levels[1], levels[0]
    inner = (slice(1, -1),) * 4
    d = u3 - u2
    t = ((u3 - 3.0 * u2) + 3.0 * u1) - u0_
    sig = np.sqrt(np.mean(d[inner] ** 2))
    kap = np.sqrt(np.mean(t[inner] ** 2))
    assert q == pytest.approx(kap, rel=1e-10)
    gm = g[0].copy()
    for a in g[1:]:
        gm = gm + a
    gm = gm * 0.125
    # unscaled features in BF16, the scales folded into the first layer's weights
    feat = np.zeros(shape + (8,), dtype=np.float32)
    feat[..., 0] = u3
    feat[..., 1] = d
    feat[..., 2] = gm
    feat[..., 3] = t
    ghost = np.ones(shape, dtype=bool)
    ghost[inner] = False
    feat[ghost] = 0.0
    fr = torch.from_numpy(feat).to(torch.bfloat16).to(torch.float32)
    off, _ = layout(layers)
    ws = w.copy()
    o, m = off[("ws", 0)]
    blk = ws[o:o + m].reshape(28, 8, 16)
    blk[:, 1, :] = blk[:, 1, :] * np.float32(float(np.float32(s_d)) / sig)  # as k_scale_w0
    blk[:, 3, :] = blk[:, 3, :] * np.float32(1.0 / kap)  # s_e = 1 after net_init
    ws[o:o + m] = blk.ravel()
    ref = reference(torch.from_numpy(ws), fr, shape[0], layers).numpy()[0]
    got = c[inner]
    err = np.abs(got - ref).max()
    print(f"  kappa {q:.3e} (sigma {sig:.3e}), max|c| {np.abs(ref).max():.4f}, max err {err:.2e}",
          flush=True)
    assert np.abs(ref).max() > 0
    assert err <= 5e-3 * np.abs(ref).max()
    assert not np.any(c[ghost])


def test_prediction_base_un_vs_torch(seglib):
    """The configured network's shape: base u^n, inputs u^n and the mean g
    only (the increment and third-difference channels stay zero), no history:
    its correction on the initial state against the PyTorch reference with
    the same BF16 rounding points, and the scale q = 1."""
    print("[gpu] learned prediction, base u^n, vs torch", flush=True)
    lib = seglib
    rng = np.random.default_rng(12)
    layers = 2
    lib.net_init(3, layers=layers, base=0)
    w = random_params(rng, 0.5, layers)
    lib.net_set_params(w, 0.7, 3.0)
    dims = (37, 30, 18, 6)
    shape, img, u0 = _volume(lib, dims, 6)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        assert lib.set_warm_start(2) == 0
        c, q = lib.net_preview(shape)
    finally:
        lib.set_warm_start(0)
        lib.set_resident(False)
        lib.deallocate()
    assert q == 1.0
    inner = (slice(1, -1),) * 4
    gm = g[0].copy()
    for a in g[1:]:
        gm = gm + a
    gm = gm * 0.125
    feat = np.zeros(shape + (8,), dtype=np.float32)
    feat[..., 0] = u0
    feat[..., 2] = gm
    ghost = np.ones(shape, dtype=bool)
    ghost[inner] = False
    feat[ghost] = 0.0
    fr = torch.from_numpy(feat).to(torch.bfloat16).to(torch.float32)
    ref = reference(torch.from_numpy(w.copy()), fr, shape[0], layers).numpy()[0]
    got = c[inner]
    err = np.abs(got - ref).max()
    print(f"  max|c| {np.abs(ref).max():.4f}, max err {err:.2e}", flush=True)
    assert np.abs(ref).max() > 0
    assert err <= 5e-3 * np.abs(ref).max()
    assert not np.any(c[ghost])
