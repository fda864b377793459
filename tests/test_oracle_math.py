"""Pins of the oracle by what the mathematics fixes (closed forms, invariants,
brute force), independent of the reference fixtures.  Each test is chosen so
that a plausible slip (a dropped term, a wrong sign or index, a transposed
face) fails it."""
import numpy as np
import pytest

from harness import params, zeros

FACES = ("e", "w", "n", "s", "t", "b", "fw", "bw")
# face -> (axis in (l,k,j,i) array order, direction)
FACE_AXIS = {0: (3, +1), 1: (3, -1), 2: (2, +1), 3: (2, -1),
             4: (1, +1), 5: (1, -1), 6: (0, +1), 7: (0, -1)}


def interior(a):
    return a[1:-1, 1:-1, 1:-1, 1:-1]


def ghost_mask(shape):
    m = np.ones(shape, bool)
    m[1:-1, 1:-1, 1:-1, 1:-1] = False
    return m


def test_gfunction_closed_form(oracle):
    assert oracle.gfunction(0.0, 1000.0) == 1.0
    assert oracle.gfunction(3.0, 2.0) == 1.0 / 7.0


def test_reflection_is_clamped_copy(oracle):
    shape = (5, 6, 7, 8)
    rng = np.random.default_rng(0)
    a = rng.standard_normal(shape)
    p = params(shape)
    b = a.copy()
    oracle.reflection(p, b)
    idx = [np.clip(np.arange(n), 1, n - 2) for n in shape]
    want = a[np.ix_(*idx)]
    assert np.array_equal(b, want)


def test_l2norm_matches_numpy(oracle):
    shape = (4, 5, 6, 7)
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal(shape), rng.standard_normal(shape)
    p = params(shape, h=0.5)
    want = np.sqrt(np.sum((a - b) ** 2) * 0.5 ** 4)
    assert oracle.l2norm(p, a, b) == pytest.approx(want, rel=1e-13)


def test_edge_indicator_constant_and_ramp(oracle):
    """Constant image -> g == 1 on every face; a linear ramp a*x away from the
    ghosts -> g = 1/(1 + K a^2) on every face (all face gradients equal a)."""
    shape = (6, 7, 8, 12)
    p = params(shape, K=1000.0, maxNoGSIteration_sigma=1.0, sigma=0.0, gsTol=1e30)
    c = np.full(shape, 0.3)
    g = zeros(shape, 8)
    oracle.gf(p, c.copy(), np.zeros(shape), np.zeros(shape), g)
    for gf in g:
        assert np.all(interior(gf) == 1.0)
    a = 0.01
    x = np.arange(shape[3], dtype=np.float64)
    ramp = np.broadcast_to(a * x, shape).copy()
    g = zeros(shape, 8)
    # sigma = 0: the heat step is the identity on the interior
    oracle.gf(p, ramp.copy(), np.zeros(shape), np.zeros(shape), g)
    want = 1.0 / (1.0 + 1000.0 * a * a)
    for f, gf in enumerate(g):
        core = gf[2:-2, 2:-2, 2:-2, 2:-2]
        np.testing.assert_allclose(core, want, rtol=1e-12, err_msg=FACES[f])


def test_theta_rho_weights(oracle):
    """rho = 0 makes the thresholded image irrelevant; theta and rho enter
    linearly inside g = 1/(1 + K (theta A + rho B))."""
    shape = (5, 6, 7, 8)
    rng = np.random.default_rng(2)
    img, thr = rng.uniform(size=shape), rng.uniform(size=shape)
    base = dict(sigma=0.0, maxNoGSIteration_sigma=1.0, gsTol=1e30, K=7.0)
    g1, g2, g3 = zeros(shape, 8), zeros(shape, 8), zeros(shape, 8)
    oracle.gf(params(shape, theta_img=1.0, rho_thr=0.0, **base), img.copy(), thr.copy(),
              np.zeros(shape), g1)
    oracle.gf(params(shape, theta_img=1.0, rho_thr=0.0, **base), img.copy(),
              rng.uniform(size=shape), np.zeros(shape), g2)
    for a, b in zip(g1, g2):
        assert np.array_equal(interior(a), interior(b))
    oracle.gf(params(shape, theta_img=0.0, rho_thr=1.0, **base), thr.copy(), img.copy(),
              np.zeros(shape), g3)
    for a, b in zip(g1, g3):  # swapping the roles with theta<->rho is the same g
        assert np.array_equal(interior(a), interior(b))


def test_face_sharing_bitwise(oracle):
    """N_w(p) == N_e(p - x) bitwise, and likewise for every axis."""
    shape = (6, 7, 8, 9)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(shape)
    p = params(shape, eps2=1e-6)
    for (lo, hi) in ((1, 0), (3, 2), (5, 4), (7, 6)):
        axis, _ = FACE_AXIS[hi]
        for l in range(2, shape[0] - 1):
            for k in range(2, shape[1] - 1):
                for j in range(2, shape[2] - 1):
                    for i in range(2, shape[3] - 1):
                        q = [l, k, j, i]
                        back = list(q)
                        back[axis] -= 1
                        a = oracle.face_norm(p, u, i, j, k, l, lo)
                        b = oracle.face_norm(p, u, back[3], back[2], back[1], back[0], hi)
                        assert a == b


def test_face_norm_matches_numpy(oracle):
    """Face gradients against a direct numpy statement of the discrete
    gradient on each face (normal difference / h, transverse averaged
    central difference / 4h)."""
    shape = (5, 6, 7, 8)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(shape)
    h, eps2 = 0.7, 0.01
    p = params(shape, h=h, eps2=eps2)

    def U(q):
        return u[tuple(q)]

    for f in range(8):
        A, d = FACE_AXIS[f]
        for q0 in [(1, 2, 3, 4), (2, 3, 4, 5), (3, 1, 1, 1)]:
            q = np.array(q0)
            eA = np.eye(4, dtype=int)[A] * d
            comp = np.zeros(4)
            for B in range(4):
                if B == A:
                    comp[B] = (U(q + eA) - U(q)) / h
                else:
                    eB = np.eye(4, dtype=int)[B]
                    comp[B] = ((U(q + eA + eB) + U(q + eB)) - (U(q + eA - eB) + U(q - eB))) / (4 * h)
            want = np.sqrt(np.sum(comp ** 2) + eps2)
            got = oracle.face_norm(p, u, q0[3], q0[2], q0[1], q0[0], f)
            assert got == pytest.approx(want, rel=1e-13), (f, q0)


def test_coefficients_constant_and_linear(oracle):
    shape = (6, 7, 8, 9)
    eps2 = 1e-6
    p = params(shape, eps2=eps2)
    rng = np.random.default_rng(5)
    g = [rng.uniform(0.1, 1.0, size=shape) for _ in range(8)]
    coef = zeros(shape, 8)
    oracle.coefficients(p, np.full(shape, 0.25), g, coef)
    for c, gg in zip(coef, g):  # N_f = eps, vc = sqrt(2) eps
        np.testing.assert_allclose(interior(c), np.sqrt(2.0) * interior(gg), rtol=1e-14)
    a = 0.3
    lin = np.broadcast_to(a * np.arange(shape[1], dtype=np.float64)[None, :, None, None],
                          shape).copy()  # linear in z
    oracle.coefficients(p, lin, None, coef)
    want = np.sqrt(a * a + 2 * eps2) / np.sqrt(a * a + eps2)
    for c in coef:
        np.testing.assert_allclose(c[1:-1, 2:-2, 1:-1, 1:-1], want, rtol=1e-13)


def _dense_matrix(shape, coef, c):
    """The linear system of one time step restated densely (interior only,
    zero Dirichlet ghosts): (1 + c sum coef) u_p - c sum coef_f u_nb = prev."""
    L_, K_, J_, I_ = shape
    idx = -np.ones(shape, int)
    inner = [(l, k, j, i) for l in range(1, L_ - 1) for k in range(1, K_ - 1)
             for j in range(1, J_ - 1) for i in range(1, I_ - 1)]
    for n_, q in enumerate(inner):
        idx[q] = n_
    A = np.zeros((len(inner), len(inner)))
    for n_, q in enumerate(inner):
        A[n_, n_] = 1 + c * sum(cf[q] for cf in coef)
        for f in range(8):
            axis, d = FACE_AXIS[f]
            nb = list(q)
            nb[axis] += d
            m_ = idx[tuple(nb)]
            if m_ >= 0:
                A[n_, m_] -= c * coef[f][q]
    return A, inner


def test_sor_converges_to_dense_solve(oracle):
    shape = (6, 6, 6, 6)
    rng = np.random.default_rng(6)
    coef = [np.zeros(shape) for _ in range(8)]
    for cf in coef:
        interior(cf)[...] = rng.uniform(0.2, 2.0, size=(4, 4, 4, 4))
    prev = np.zeros(shape)
    interior(prev)[...] = rng.uniform(size=(4, 4, 4, 4))
    c = 0.5
    p = params(shape, tau=0.5, h=1.0, gsTol=1e-28, maxNoGSIteration=10000)
    u = np.zeros(shape)
    z, R = oracle.sor(p, u, prev, coef)
    assert R <= 1e-28 and z < 10000
    A, inner = _dense_matrix(shape, coef, c)
    b = np.array([prev[q] for q in inner])
    x = np.linalg.solve(A, b)
    np.testing.assert_allclose([u[q] for q in inner], x, rtol=1e-11, atol=1e-13)


def test_residual_closed_form(oracle):
    shape = (5, 6, 5, 6)
    rng = np.random.default_rng(7)
    coef = [np.zeros(shape) for _ in range(8)]
    for cf in coef:
        interior(cf)[...] = rng.uniform(0.2, 2.0, size=interior(cf).shape)
    u, prev = rng.standard_normal(shape), rng.standard_normal(shape)
    h = 0.5
    c = 0.3 / (h * h)
    p = params(shape, tau=0.3, h=h)
    A, inner = _dense_matrix(shape, coef, c)
    uin = np.array([u[q] for q in inner])
    # ghost neighbours contribute -c coef u_ghost
    rhs_ghost = np.zeros(len(inner))
    for n_, q in enumerate(inner):
        for f in range(8):
            axis, d = FACE_AXIS[f]
            nb = list(q)
            nb[axis] += d
            if any(x_ in (0, s_ - 1) for x_, s_ in zip(nb, shape)):
                rhs_ghost[n_] -= c * coef[f][q] * u[tuple(nb)]
    r = A @ uin + rhs_ghost - np.array([prev[q] for q in inner])
    want = np.sum(r * r) * h ** 4
    assert oracle.residual(p, u, prev, coef) == pytest.approx(want, rel=1e-11)


def test_sor_fixed_point_unchanged(oracle):
    """If prev = A u exactly (u solves the system), SOR leaves u unchanged up
    to rounding and reports a tiny residual after one sweep."""
    shape = (5, 5, 6, 6)
    rng = np.random.default_rng(8)
    coef = [np.zeros(shape) for _ in range(8)]
    for cf in coef:
        interior(cf)[...] = rng.uniform(0.5, 1.5, size=interior(cf).shape)
    u = np.zeros(shape)
    interior(u)[...] = rng.uniform(size=interior(u).shape)
    A, inner = _dense_matrix(shape, coef, 0.5)
    bvec = A @ np.array([u[q] for q in inner])
    prev = np.zeros(shape)
    for n_, q in enumerate(inner):
        prev[q] = bvec[n_]
    p = params(shape, gsTol=-1.0, maxNoGSIteration=1)
    v = u.copy()
    oracle.sor(p, v, prev, coef)
    np.testing.assert_allclose(v, u, rtol=0, atol=1e-13)


def test_presmooth_is_implicit_heat_step(oracle):
    """Converged presmoothing solves (1 + 8c) u - c sum_nb u = prev with the
    ghosts frozen at the reflected input (brute-force dense solve)."""
    shape = (5, 5, 6, 6)
    rng = np.random.default_rng(9)
    img = rng.uniform(size=shape)
    c = 0.5
    p = params(shape, sigma=0.5, h=1.0, omega_sigma=1.3, gsTol=1e-30,
               maxNoGSIteration_sigma=5000.0)
    a, prev = img.copy(), np.zeros(shape)
    z, R = oracle.presmooth(p, a, prev)
    assert np.array_equal(prev, img)
    refl = img.copy()
    oracle.reflection(p, refl)
    assert np.array_equal(a[ghost_mask(shape)], refl[ghost_mask(shape)])
    ones = [np.ones(shape) for _ in range(8)]
    A, inner = _dense_matrix(shape, ones, c)
    rhs = np.array([img[q] for q in inner])
    for n_, q in enumerate(inner):
        for f in range(8):
            axis, d = FACE_AXIS[f]
            nb = list(q)
            nb[axis] += d
            if any(x_ in (0, s_ - 1) for x_, s_ in zip(nb, shape)):
                rhs[n_] += c * refl[tuple(nb)]
    x = np.linalg.solve(A, rhs)
    np.testing.assert_allclose([a[q] for q in inner], x, rtol=1e-11)


def test_rescale_identity_and_ball(oracle):
    shape = (5, 9, 10, 11)
    rng = np.random.default_rng(10)
    u = rng.uniform(0.2, 0.8, size=shape)
    p = params(shape)
    v = u.copy()
    oracle.rescale(p, v, np.ones(shape))  # no balls, full mask: identity
    assert np.array_equal(u, v)
    mask = (rng.uniform(size=shape) > 0.3).astype(float)
    v = u.copy()
    oracle.rescale(p, v, mask)
    want = u.copy()
    inner = interior(want)
    inner[interior(mask) == 0] = 0
    assert np.array_equal(v, want)
    # one ball: brute force of min/max rescale inside the radius
    counts = [0, 1, 0, 0, 0]
    centres = [[], [(5.0, 4.0, 4.0)], [], [], []]
    R = 2.5
    balls = oracle.make_balls(counts, centres, R)
    v = u.copy()
    oracle.rescale(p, v, np.ones(shape), balls)
    kk, jj, ii = np.meshgrid(np.arange(shape[1]), np.arange(shape[2]), np.arange(shape[3]),
                             indexing="ij")
    inside = np.sqrt((ii - 5.0) ** 2 + (jj - 4.0) ** 2 + (kk - 4.0) ** 2) <= R
    frame = u[1]
    mn, mx = frame[inside].min(), frame[inside].max()
    want = u.copy()
    want[1][inside] = (frame[inside] - mn) / (mx - mn)
    np.testing.assert_allclose(v, want, rtol=0, atol=1e-14)


def test_mcf_exact_and_boundary(oracle):
    m = 7
    shape = (m, m, m, m)
    h = 2.5 / (m - 2)
    p = params(shape, h=h)
    ex = np.zeros(shape)
    oracle.mcf_exact(p, ex, 0.25)
    x = (np.arange(m) - 1) * h + h / 2 - 1.25
    L_, K_, J_, I_ = np.meshgrid(x, x, x, x, indexing="ij")
    want = (I_ ** 2 + J_ ** 2 + K_ ** 2 + L_ ** 2 - 1.0) / 6.0 + 0.25
    np.testing.assert_allclose(ex, want, rtol=1e-14, atol=1e-15)
    u = np.zeros(shape)
    oracle.mcf_boundary(p, u, 0.25)
    assert np.all(interior(u) == 0)
    np.testing.assert_allclose(u[ghost_mask(shape)], want[ghost_mask(shape)], atol=1e-14)


def test_mcf_error_decreases_under_refinement(oracle):
    """The MCF twin converges to the exact solution (u_t = |grad u| div(grad u
    /|grad u|) solved by u = (|x|^2 - 1)/6 + t): the L2 error at a fixed time
    falls under grid refinement (tau = h^2, eps = h, as the reference)."""
    errs = []
    for n_, steps in ((6, 2), (12, 8)):  # same final time: tau = h^2 scales by 1/4
        m = n_ + 2
        shape = (m, m, m, m)
        h = 2.5 / n_
        tau = h * h
        p = params(shape, h=h, tau=tau, eps2=h * h, gsTol=1e-16, maxNoGSIteration=5000)
        u = np.zeros(shape)
        oracle.mcf_exact(p, u, 0.0)
        prev, coef, ex = np.zeros(shape), zeros(shape, 8), np.zeros(shape)
        for st in range(1, steps + 1):
            oracle.mcf_time_step(p, u, prev, coef, st)
        oracle.mcf_exact(p, ex, steps * tau)
        errs.append(oracle.l2norm(p, ex, u) / np.sqrt((2.5) ** 4))
    assert errs[1] < errs[0]


def test_mask_and_iou_brute_force(oracle):
    shape = (4, 5, 6, 7)
    rng = np.random.default_rng(11)
    u = rng.uniform(size=shape)
    p = params(shape)
    m = oracle.mask(p, u)
    inner = interior(u)
    thr = 0.5 * (inner.max() + inner.min())
    want = np.zeros(shape, np.uint8)
    interior(want)[...] = inner >= thr
    assert np.array_equal(m, want)
    m2 = m.copy()
    m2[1, 1, 1, 1] ^= 1
    inter = np.sum(m & m2)
    union = np.sum(m | m2)
    assert oracle.iou(p, m, m2) == pytest.approx(inter / union)
    assert oracle.iou(p, m, m) == 1.0


def test_init_functions_brute_force(oracle):
    """Seed function and ball reductions against numpy brute force, with
    centres near and beyond the edges (the oracle clamps the ball boxes)."""
    rng = np.random.default_rng(3)
    shape = (4, 9, 10, 11)
    img = rng.uniform(size=shape)
    p = params(shape)
    counts = [0, 2, 1, 0]
    centres = [[], [(0.3, 8.7, 4.0), (5.2, 4.9, 3.5)], [(10.0, 0.0, 8.0)], []]
    R = 2.6
    balls = oracle.make_balls(counts, centres, R)
    kk, jj, ii = np.meshgrid(np.arange(shape[1]), np.arange(shape[2]), np.arange(shape[3]),
                             indexing="ij")
    u = rng.uniform(0, 0.05, size=shape)
    u_ref = u.copy()
    mask = np.zeros(shape)
    mask_ref = mask.copy()
    for l in range(4):
        for c in centres[l][: counts[l]]:
            d = np.sqrt((ii - c[0]) ** 2 + (jj - c[1]) ** 2 + (kk - c[2]) ** 2)
            ball = d <= R
            assert oracle.ball_minmax(p, img, l, c, R, 1) == img[l][ball].max()
            assert oracle.ball_minmax(p, img, l, c, R, 0) == img[l][ball].min()
            # the seed's box is the ballRadius box, its radius test is d <= 12
            lo =[max(int(x - R - 0.5), 0) for x in c]
            hi = [int(x + R + 0.5) for x in c]
            box = ((ii >= lo[0]) & (ii <= hi[0]) & (jj >= lo[1]) & (jj <= hi[1]) &
                   (kk >= lo[2]) & (kk <= hi[2]))
            nv = np.abs(1.0 / (d + 0.5) - 1.0 / 12.5)
            u_ref[l][box] = np.maximum(u_ref[l][box], nv[box])
            mask_ref[l][box] = 1.0
    oracle.generate(p, u, mask, balls)
    assert np.array_equal(mask, mask_ref)
    np.testing.assert_allclose(u, u_ref, rtol=1e-15, atol=0)
    # at the centre the seed is 1/v - 1/(Rs + v) = 2 - 2/25
    p2 = params((3, 5, 5, 5))
    u2, m2 = np.zeros((3, 5, 5, 5)), np.zeros((3, 5, 5, 5))
    oracle.generate(p2, u2, m2, oracle.make_balls([0, 1, 0], [[], [(2.0, 2.0, 2.0)], []], 1.0))
    assert u2[1, 2, 2, 2] == 2.0 - 1.0 / 12.5
    assert oracle.get4dmax(p, -img) == 0.0  # the reference's max starts at 0
