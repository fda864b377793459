"""Batched form (seg4dBatch*): volumes of different sizes stepped together
in one launch per colour pass.  Every volume's u after every step and its
sweep count must be bitwise those of the same steps on the single-volume
library (whose FP64 path is bitwise the oracle's, checked directly for the
smallest volume below), in FP64 and FP32-refined arithmetic, with eager or
graph-replayed iterations, any queue depth and autotuned launch tiles, and
in the concurrent form (volume groups on their own streams and host threads,
switched to and from within one batch), and with each time step captured as
one graph."""
import numpy as np
import pytest

from harness import params, zeros

pytestmark = pytest.mark.gpu

KW = dict(h=1.0, tau=0.5, eps2=1e-6, K=1000.0, omega_c=1.8, omega_sigma=1.3, sigma=0.5,
          theta_img=1.0, rho_thr=0.0, maxNoGSIteration=100000, maxNoGSIteration_sigma=100000.0)
REL = 1e-6
DIMS = [(17, 11, 7, 3), (24, 20, 12, 5), (9, 8, 6, 4), (30, 14, 9, 6), (40, 33, 16, 7)]
STEPS = 3


def _volume(lib, dims, seed):
    spec = lib.gen_spec(0, seed)
    spec.nx, spec.ny, spec.nz, spec.nt = dims
    spec.axes[0][0], spec.axes[0][1], spec.axes[0][2] = dims[0] / 4, dims[1] / 5, dims[2] / 4
    spec.axes[1][0], spec.axes[1][1], spec.axes[1][2] = dims[0] / 8, dims[1] / 8, dims[2] / 6
    shape = (dims[3] + 2, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    img, u0 = np.zeros(shape), np.zeros(shape)
    lib.generate(spec, img, u0)
    return img, u0


def _single(lib, dims, img, u0, precision, warm, steps):
    """The same steps on the single-volume library (resident mode)."""
    out = []
    lib.allocate(*dims)
    # the batch's engines run the library defaults
    lib.set_sweep_variant(35)
    lib.lib.seg4dSetSweepTile(16, 8)
    lib.lib.seg4dSetSweepChunk(0)
    lib.lib.seg4dSetPipelineDepth(0)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(REL)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        lib.set_resident(True)
        lib.arr("u")[...] = u0
        lib.upload("u")
        lib.upload("Mmask")
        lib.set_precision(precision)
        lib.set_warm_start(warm)
        for s in range(steps):
            lib.set(numberOfTimeStep=s + 1)
            z, R = lib.seg_time_step()
            lib.download("u")
            out.append((lib.arr("u").copy(), z, lib.last_rounds()))
    finally:
        lib.set_warm_start(0)
        lib.set_precision(0)
        lib.set_resident(False)
        lib.set_relative_tolerance(0.0)
        lib.deallocate()
    return out


@pytest.fixture(scope="module")
def volumes(seglib):
    return [_volume(seglib, d, 11 + q) for q, d in enumerate(DIMS)]


def _run_batch(lib, volumes, precision, warm, graph, depth, tune, steps, streams=(0,),
               plans=None, info=None):
    """streams[s]: the concurrent form's group count at step s (the last entry
    for later steps; 0 is one launch per pass).  plans[s]: (iterations,
    rounds) forced on the captured step s (graph mode 2); info (a list)
    receives seg4dBatchGraphInfo after the last step."""
    lib.set(gsTol=0.0, **KW)
    lib.set_relative_tolerance(REL)
    lib.batch_create(DIMS)
    res = []
    try:
        lib.batch_configure(precision, warm, graph, depth)
        for v, (img, u0) in enumerate(volumes):
            lib.batch_set_volume(v, img, u0, np.ones_like(img))
        if tune:
            tiles, ms = lib.batch_autotune(2)
            print(f"  tuned tiles {tiles}", flush=True)
        for s in range(steps):
            lib.set(numberOfTimeStep=s + 1)
            lib.batch_streams(streams[min(s, len(streams) - 1)])
            if plans and s in plans:
                lib.batch_set_graph_plan(*plans[s])
            lib.batch_step()
            res.append([(lib.batch_u(v),) + lib.batch_stats(v)[:2] for v in range(len(DIMS))])
        if info is not None:
            info.extend(lib.batch_graph_info())
    finally:
        lib.batch_destroy()
        lib.set_relative_tolerance(0.0)
    return res


@pytest.mark.parametrize("precision,warm,graph,depth,tune,streams",
                         [(0, 0, 0, 2, False, (0,)), (0, 0, 1, 2, False, (0,)),
                          (0, 3, 0, 1, True, (0,)), (1, 0, 0, 2, False, (0,)),
                          (1, 4, 1, 3, True, (0,)), (0, 0, 0, 2, False, (2,)),
                          (0, 3, 0, 2, True, (3, 3, 0, 5)), (1, 4, 1, 2, True, (0, 2, 4))])
def test_batch_matches_single_volume(seglib, volumes, precision, warm, graph, depth, tune,
                                     streams):
    print(f"[gpu] batch precision {precision} warm {warm} graph {graph} depth {depth} "
          f"tune {tune} streams {streams}", flush=True)
    lib = seglib
    steps = STEPS + (2 if warm == 4 else 0)  # reach the cubic start
    if len(streams) > 1:
        steps = max(steps, len(streams))
    got = _run_batch(lib, volumes, precision, warm, graph, depth, tune, steps, streams)
    for v, (img, u0) in enumerate(volumes):
        ref = _single(lib, DIMS[v], img, u0, precision, warm, steps)
        for s in range(steps):
            ub, zb, rb = got[s][v]
            us, zs, rs = ref[s]
            assert zb == zs, f"volume {v} step {s + 1}: sweeps {zb} vs {zs}"
            assert rb == rs, f"volume {v} step {s + 1}: rounds {rb} vs {rs}"
            assert np.array_equal(ub.view(np.uint64), us.view(np.uint64)), \
                f"volume {v} step {s + 1}: u differs (max {np.abs(ub - us).max():.3e})"
    print(f"  {len(DIMS)} volumes x {steps} steps bitwise equal; sweeps "
          f"{[got[-1][v][1] for v in range(len(DIMS))]}", flush=True)


@pytest.mark.parametrize("precision,warm,depth", [(0, 0, 2), (0, 3, 1), (1, 0, 2), (1, 4, 3)])
def test_batch_captured_step(seglib, volumes, precision, warm, depth):
    """Graph mode 2: every step after the first runs as one captured graph
    (set-ups, solve or correction rounds, gated epilogue), bitwise the
    single-volume steps.  Step 2 is forced to hold one iteration per solve:
    its solves overrun, the graph writes nothing and the step is redone
    eagerly with the same result; the later steps plan from it again."""
    print(f"[gpu] captured step precision {precision} warm {warm} depth {depth}", flush=True)
    lib = seglib
    steps = STEPS + 2
    info = []
    got = _run_batch(lib, volumes, precision, warm, 2, depth, False, steps, plans={1: (1, 1)},
                     info=info)
    iters, rounds, gsteps, redone, _ = info
    assert redone == 1, f"forced overrun: {redone} steps redone"
    assert gsteps == steps - 2, f"{gsteps} steps ran as one graph, expected {steps - 2}"
    assert iters >= 5 and rounds >= 1
    for v, (img, u0) in enumerate(volumes):
        ref = _single(lib, DIMS[v], img, u0, precision, warm, steps)
        for s in range(steps):
            ub, zb, rb = got[s][v]
            us, zs, rs = ref[s]
            assert zb == zs, f"volume {v} step {s + 1}: sweeps {zb} vs {zs}"
            assert rb == rs, f"volume {v} step {s + 1}: rounds {rb} vs {rs}"
            assert np.array_equal(ub.view(np.uint64), us.view(np.uint64)), \
                f"volume {v} step {s + 1}: u differs (max {np.abs(ub - us).max():.3e})"
    print(f"  {len(DIMS)} volumes x {steps} steps bitwise equal; {gsteps} captured steps, "
          f"plan {iters} iterations x {rounds} rounds", flush=True)


def test_batch_fp64_vs_oracle(seglib, oracle, volumes):
    """The smallest volume of an FP64 batch against the oracle directly (the
    oracle steps with the single-volume library's edge indicator)."""
    lib = seglib
    got = _run_batch(lib, volumes, 0, 0, 1, 2, False, 2)
    v = 2
    img, u0 = volumes[v]
    shape = img.shape
    lib.allocate(*DIMS[v])
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(REL)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        g = [lib.arr("g_" + f).copy() for f in ("e", "w", "n", "s", "t", "b", "fw", "bw")]
    finally:
        lib.set_relative_tolerance(0.0)
        lib.deallocate()
    u = u0.copy()
    prev = np.zeros(shape)
    coef = zeros(shape, 8)
    for s in range(2):
        ps = params(shape, gsTol=REL * REL * oracle.sumsq_interior(params(shape, **KW), u), **KW)
        z, _ = oracle.seg_time_step(ps, u, prev, g, coef, np.ones(shape))
        ub, zb, _ = got[s][v]
        assert zb == z, f"step {s + 1}: sweeps {zb} vs oracle {z}"
        assert np.array_equal(ub.view(np.uint64), u.view(np.uint64)), \
            f"step {s + 1}: max |du| {np.abs(ub - u).max():.3e}"


def test_autotune_single_volume(seglib, volumes):
    """The single-volume autotuner picks a tile and leaves results unchanged."""
    lib = seglib
    dims = DIMS[4]
    img, u0 = volumes[4]
    ref = _single(lib, dims, img, u0, 0, 0, 1)
    lib.allocate(*dims)
    try:
        lib.set(gsTol=0.0, **KW)
        lib.set_relative_tolerance(REL)
        lib.arr("i0")[...] = img
        lib.arr("u")[...] = 0.0
        lib.arr("Mmask")[...] = 1.0
        lib.gf("i0", "u")
        lib.arr("u")[...] = u0
        tb, tk, ms = lib.autotune(2)
        assert ms > 0.0 and tb >= 0 and tk >= 0
        lib.set(numberOfTimeStep=1)
        z, _ = lib.seg_time_step()
        assert z == ref[0][1]
        assert np.array_equal(lib.arr("u").view(np.uint64), ref[0][0].view(np.uint64))
    finally:
        lib.set_relative_tolerance(0.0)
        lib.deallocate()
