"""The oracle reproduces, bitwise, the golden fixtures that the reference's own
compiled programs wrote (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from harness import GOLDEN, params, zeros

FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def seg_params(fx):
    shape = fx["i0"].shape
    kw = {k[2:]: fx[k].item() for k in fx.files if k.startswith("p_")}
    return params(shape, **kw)


def balls_of(oracle, fx):
    counts = [int(c) for c in fx["counts"]]
    flat = [tuple(r) for r in fx["centres"]]
    centres, q = [], 0
    for c in counts:
        centres.append(flat[q:q + c])
        q += c
    return oracle.make_balls(counts, centres, float(fx["radius"]))


@pytest.mark.parametrize("name", ["seg_h1.npz", "seg_h001.npz"])
def test_seg_golden(oracle, name):
    fx = load(name)
    p = seg_params(fx)
    shape = fx["i0"].shape
    a, b, u0, mask = fx["i0"].copy(), fx["thr"].copy(), fx["u0"].copy(), fx["mask"].copy()
    prev = np.zeros(shape)
    g = zeros(shape, 8)
    oracle.gf(p, a, b, prev, g)
    assert np.array_equal(a, fx["i0_s"])
    assert np.array_equal(b, fx["thr_s"])
    assert np.array_equal(prev, fx["prev_gf"])
    for f, gf in zip(FACES, g):
        assert np.array_equal(gf, fx["g_" + f]), f
    balls = balls_of(oracle, fx)
    u = u0.copy()
    coef = zeros(shape, 8)
    for st in range(int(fx["steps"])):
        oracle.seg_time_step(p, u, prev, g, coef, mask, balls)
        if st == 0:
            for f, c in zip(FACES, coef):
                assert np.array_equal(c, fx["coef1_" + f]), f
        assert np.array_equal(u, fx[f"u{st + 1}"]), f"step {st + 1}"
        assert oracle.l2norm(p, prev, u) == fx["l2"][st]
    assert np.array_equal(prev, fx["prev_last"])


def test_mcf_golden(oracle):
    fx = load("mcf.npz")
    m, hh = int(fx["m"]), float(fx["h"])
    shape = (m, m, m, m)
    p = params(shape, h=hh, tau=hh * hh, eps2=hh * hh, gsTol=1e-12, omega_c=1.8,
               maxNoGSIteration=1000)
    u = np.zeros(shape)
    oracle.mcf_exact(p, u, 0.0)
    assert np.array_equal(u, fx["u0"])
    prev = np.zeros(shape)
    coef = zeros(shape, 8)
    ex = np.zeros(shape)
    for st in range(1, int(fx["steps"]) + 1):
        oracle.mcf_time_step(p, u, prev, coef, st)
        assert np.array_equal(u, fx[f"u{st}"]), st
        oracle.mcf_exact(p, ex, st * hh * hh)
        assert oracle.l2norm(p, ex, u) == fx["err"][st - 1]
    assert np.array_equal(ex, fx["exact_last"])


def init_balls(oracle, fx):
    return balls_of(oracle, fx)


def test_init_golden(oracle):
    """Initialisation chain (segmentation4D.c:1386-1395): local threshold, seed
    function with Mmask, rescale; plus the helper reductions."""
    fx = load("init.npz")
    shape = fx["i0"].shape
    p = params(shape)
    balls = init_balls(oracle, fx)
    R, alpha, beta = float(fx["radius"]), float(fx["alpha"]), float(fx["beta"])
    img = fx["i0"].copy()
    assert oracle.get4dmax(p, img) == fx["gmax"]
    c1 = tuple(fx["centres"][0])
    assert oracle.ball_minmax(p, img, 1, c1, R, 1) == fx["bmax"]
    assert oracle.ball_minmax(p, img, 1, c1, R, 0) == fx["bmin"]
    assert oracle.mass(params(shape, h=0.5), img) == fx["mass"]
    u = np.zeros(shape)
    oracle.threshold(p, img, u, balls, alpha, beta)
    assert np.array_equal(u, fx["thr"])
    mask = np.zeros(shape)
    oracle.generate(p, u, mask, balls)
    assert np.array_equal(u, fx["seed"])
    assert np.array_equal(mask, fx["mask"])
    oracle.rescale(p, u, np.ones(shape), balls)
    assert np.array_equal(u, fx["rescaled"])


def test_main_flow_golden(oracle):
    """The reference main()'s flow: threshold, edge indicator, seed, rescale,
    three time steps with its own parameters."""
    fx = load("main_flow.npz")
    shape = fx["i0"].shape
    kw = {k[2:]: fx[k].item() for k in fx.files if k.startswith("p_")}
    alpha, beta, R = kw.pop("alpha"), kw.pop("beta"), kw.pop("ballRadius")
    p = params(shape, **kw)
    counts = [int(c) for c in fx["counts"]]
    flat = [tuple(r) for r in fx["centres"]]
    centres, q = [], 0
    for c in counts:
        centres.append(flat[q:q + c])
        q += c
    balls = oracle.make_balls(counts, centres, R)
    i0, u = fx["i0"].copy(), np.zeros(shape)
    oracle.threshold(p, i0, u, balls, alpha, beta)
    prev = np.zeros(shape)
    g = zeros(shape, 8)
    oracle.gf(p, i0, u, prev, g)
    assert np.array_equal(u, fx["thr_s"]) and np.array_equal(i0, fx["i0_s"])
    for f, gf in zip(FACES, g):
        assert np.array_equal(gf, fx["g_" + f]), f
    mask = np.zeros(shape)
    oracle.generate(p, u, mask, balls)
    oracle.rescale(p, u, mask, balls)
    assert np.array_equal(u, fx["u0"]) and np.array_equal(mask, fx["mask"])
    coef = zeros(shape, 8)
    for st in range(int(fx["steps"])):
        oracle.seg_time_step(p, u, prev, g, coef, mask, balls)
        assert np.array_equal(u, fx[f"u{st + 1}"]), f"step {st + 1}"
        assert oracle.l2norm(p, prev, u) == fx["l2"][st]
