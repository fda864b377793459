"""Initialisation chain on the GPU (segmentation4D.c:1386-1395 and the
helpers of :317-335, :876-1148) against the golden fixture written by the
compiled reference and against the oracle.  Bitwise: every value is a
comparison, a max/min or the same per-point arithmetic; _mass4D_ is a sum
in another order and is held to 1e-13 relative."""
import os

import numpy as np
import pytest

from harness import GOLDEN, params

pytestmark = pytest.mark.gpu


def _split(counts, flat):
    out, q = [], 0
    for c in counts:
        out.append([tuple(x) for x in flat[q:q + c]])
        q += c
    return out


def test_init_chain_golden(seglib):
    fx = np.load(os.path.join(GOLDEN, "init.npz"), allow_pickle=False)
    L_, K_, J_, I_ = fx["i0"].shape
    counts = [int(c) for c in fx["counts"]]
    centres = _split(counts, fx["centres"])
    R = float(fx["radius"])
    lib = seglib
    print(f"[gpu] init chain golden {fx['i0'].shape}", flush=True)
    lib.allocate(I_ - 2, J_ - 2, K_ - 2, L_ - 2, centres=counts)
    try:
        lib.set_balls(centres, R)
        lib.set(alpha=float(fx["alpha"]), beta=float(fx["beta"]), h=0.5)
        lib.arr("i0")[...] = fx["i0"]
        assert lib.get4dmax("i0") == fx["gmax"]
        assert lib.ball_max("i0", 1, centres[1][0], R) == fx["bmax"]
        assert lib.ball_min("i0", 1, centres[1][0], R) == fx["bmin"]
        assert abs(lib.mass("i0") - fx["mass"]) <= 1e-13 * abs(fx["mass"])
        lib.locally_threshold("i0", "u")
        assert np.array_equal(lib.arr("u"), fx["thr"])
        lib.generate_seed("u")
        assert np.array_equal(lib.arr("u"), fx["seed"])
        assert np.array_equal(lib.arr("Mmask"), fx["mask"])
        lib.rescale("u")
        assert np.array_equal(lib.arr("u"), fx["rescaled"])
    finally:
        lib.deallocate()


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("alias", [False, True])
def test_init_chain_vs_oracle(seglib, oracle, resident, alias):
    """Random image, balls overlapping each other and the array's edges
    (boxes clamped), thresholding in place when `alias`."""
    dims = (29, 23, 14, 5)
    print(f"[gpu] init chain {dims} resident={resident} alias={alias}", flush=True)
    rng = np.random.default_rng(21 + resident + 2 * alias)
    shape = (dims[3] + 2, dims[2] + 2, dims[1] + 2, dims[0] + 2)
    img = np.round(rng.uniform(0, 255, size=shape)) / 255.0
    counts = [0] + [int(rng.integers(0, 5)) for _ in range(dims[3])] + [0]
    centres = [[(float(rng.uniform(-2, dims[0] + 3)), float(rng.uniform(-2, dims[1] + 3)),
                 float(rng.uniform(-2, dims[2] + 3))) for _ in range(c)] for c in counts]
    R, alpha, beta = 4.3, 0.7, 0.3
    p = params(shape)
    balls = oracle.make_balls(counts, centres, R)
    src = img.copy()
    thr = src if alias else np.zeros(shape)
    oracle.threshold(p, src, thr, balls, alpha, beta)
    mask = np.zeros(shape)
    oracle.generate(p, thr, mask, balls)
    seed = thr.copy()
    oracle.rescale(p, thr, mask, balls)

    lib = seglib
    lib.allocate(*dims, centres=counts)
    try:
        lib.set_balls(centres, R)
        lib.set(alpha=alpha, beta=beta)
        lib.arr("i0")[...] = img
        dst = "i0" if alias else "u"
        if resident:
            lib.set_resident(True)
            lib.upload("i0")
            lib.upload("u")
            lib.upload("Mmask")
        lib.locally_threshold("i0", dst)
        lib.generate_seed(dst)
        if resident:
            lib.download(dst)
            lib.download("Mmask")
        assert np.array_equal(lib.arr(dst), seed)
        assert np.array_equal(lib.arr("Mmask"), mask)
        lib.rescale(dst)
        if resident:
            lib.download(dst)
        assert np.array_equal(lib.arr(dst), thr)
    finally:
        lib.set_resident(False)
        lib.deallocate()


@pytest.mark.parametrize("resident", [False, True])
def test_main_flow_golden(seglib, resident):
    """The reference main()'s whole flow through the library's C ABI, in the
    reference's call order, bitwise against the compiled reference."""
    fx = np.load(os.path.join(GOLDEN, "main_flow.npz"), allow_pickle=False)
    L_, K_, J_, I_ = fx["i0"].shape
    kw = {k[2:]: fx[k].item() for k in fx.files if k.startswith("p_")}
    counts = [int(c) for c in fx["counts"]]
    centres = _split(counts, fx["centres"])
    lib = seglib
    print(f"[gpu] main flow golden resident={resident}", flush=True)
    lib.allocate(I_ - 2, J_ - 2, K_ - 2, L_ - 2, centres=counts)
    try:
        lib.set_balls(centres, kw.pop("ballRadius"))
        lib.set(**kw)
        lib.arr("i0")[...] = fx["i0"]
        lib.locally_threshold("i0", "u")
        lib.gf("i0", "u")
        assert np.array_equal(lib.arr("u"), fx["thr_s"])
        lib.generate_seed("u")
        lib.rescale("u")
        assert np.array_equal(lib.arr("u"), fx["u0"])
        assert np.array_equal(lib.arr("Mmask"), fx["mask"])
        if resident:
            lib.set_resident(True)
            lib.upload("u")
            lib.upload("Mmask")
        for st in range(int(fx["steps"])):
            lib.set(numberOfTimeStep=st + 1)
            lib.seg_time_step()
            if resident:
                lib.download("u")
            assert np.array_equal(lib.arr("u"), fx[f"u{st + 1}"]), f"step {st + 1}"
            assert abs(lib.l2norm("prev_u", "u") - fx["l2"][st]) <= 1e-13 * fx["l2"][st]
    finally:
        lib.set_resident(False)
        lib.deallocate()
