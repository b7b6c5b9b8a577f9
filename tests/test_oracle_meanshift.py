"""Pins the batched-seed mean-shift of the oracle (oracle_meanshift_batched)
with what the mathematics fixes.  The reference has no batched-seed mode (its
seeds are sequential, src/cpu/clustering.cpp:44-60), so these checks, not the
reference, anchor it.

The update is m <- m + sum_j w_j (x_j - m) / sum_j w_j over the votes of the
seed's kind, with w(u) = exp(-u / 2h^2) - exp(-c^2 / 2) for u = |x - m|^2 <
(c h)^2 and 0 outside.  That is the mean-shift of the density

    f(m) = sum_j K(|x_j - m|^2),
    K(u) = 2h^2 (exp(-u / 2h^2) - e) - (c^2 h^2 - u) e   for u < c^2 h^2, else 0,
    e = exp(-c^2 / 2),

since K'(u) = -w(u), so grad f(m) = 2 sum_j w_j (x_j - m) (Comaniciu and Meer,
IEEE PAMI 24(5), 2002, section 2.2: the weights are the negative derivative
of the profile).  K is convex and decreasing on [0, inf), C^1 at the support
boundary (w vanishes there), so the theorem of that paper (monotone ascent)
applies.  The checks:

  * stationarity: at the oracle's converged modes the FP64 central-difference
    gradient of f, evaluated from K alone, vanishes relative to f / h, and the
    finite-difference Hessian is negative definite (a local maximum);
  * monotone ascent: f never decreases along the oracle's iterates;
  * symmetry: a point set symmetric about a centre keeps a seed at the centre
    exactly, and a seed on a symmetry hyperplane stays on it;
  * support boundary: a point at (1 - 1e-9) c h is inside (the mean jumps onto
    it), one at (1 + 1e-9) c h is outside (the mean stays);
  * kinds: a seed sees only the votes of its own kind, in its kind's
    dimensions.

The stationarity check is shown to have power: the same iteration with
weights that lack the shift e converges to points where the gradient of f is
about 1e-2 f/h, against 3e-10 for the oracle.
"""
import math

import numpy as np
import pytest

H = 0.05
C = 3.0


def K(u, h=H, c=C):
    e = math.exp(-0.5 * c * c)
    inside = u < (c * h) ** 2
    return np.where(inside, 2 * h * h * (np.exp(-u / (2 * h * h)) - e) - (c * c * h * h - u) * e,
                    0.0)


def density(m, pts, h=H, c=C):
    """f(m) for one mean m (D,) over points pts (D, V)."""
    d = pts - m[:, None]
    return float(np.sum(K(np.sum(d * d, axis=0), h, c)))


def fd_gradient(m, pts, step):
    g = np.zeros(len(m))
    for k in range(len(m)):
        e = np.zeros(len(m))
        e[k] = step
        g[k] = (density(m + e, pts) - density(m - e, pts)) / (2 * step)
    return g


def fd_hessian(m, pts, step):
    D = len(m)
    Hm = np.zeros((D, D))
    for a in range(D):
        for b in range(D):
            ea = np.zeros(D)
            eb = np.zeros(D)
            ea[a] = step
            eb[b] = step
            Hm[a, b] = (density(m + ea + eb, pts) - density(m + ea - eb, pts) -
                        density(m - ea + eb, pts) + density(m - ea - eb, pts)) / (4 * step * step)
    return 0.5 * (Hm + Hm.T)


def blobs(rng, dim, counts, spreads, centres):
    cols = []
    for n, s, c in zip(counts, spreads, centres):
        cols.append(np.asarray(c)[:, None] + rng.normal(0, s, size=(dim, n)))
    return np.concatenate(cols, axis=1)


def skewed_points(seed=3, dim=6):
    """Three anisotropic, skewed clusters (so plausible mistakes in the weight
    move the fixed points measurably) plus uniform clutter."""
    rng = np.random.default_rng(seed)
    a = blobs(rng, dim, [120, 60], [0.6 * H, 0.3 * H], [np.zeros(dim), np.full(dim, 0.8 * H)])
    b = blobs(rng, dim, [150], [0.9 * H], [np.linspace(0.5, 0.7, dim)])
    tail = np.zeros((dim, 40))
    tail[0] = 0.5 + rng.exponential(0.8 * H, 40)
    tail[1:] = np.linspace(0.5, 0.7, dim)[1:, None] + rng.normal(0, 0.2 * H, (dim - 1, 40))
    clutter = rng.uniform(-0.2, 1.0, size=(dim, 40))
    return np.concatenate([a, b, tail, clutter], axis=1)


def run(oracle, pts, seeds, iterations, kind=None, seed_kind=None, dim_rigid=None,
        dim_refl=None):
    dim = pts.shape[0]
    kind = np.zeros(pts.shape[1], np.uint8) if kind is None else kind
    seed_kind = np.zeros(seeds.shape[1], np.uint8) if seed_kind is None else seed_kind
    return oracle.meanshift_batched(pts, kind, dim if dim_rigid is None else dim_rigid,
                                    dim if dim_refl is None else dim_refl, seeds, seed_kind, H,
                                    C, iterations)


def numpy_variant(pts, seeds, iterations, shift=True, strict=True):
    """The same iteration with a planted mistake (shift=False drops the
    -exp(-c^2/2) term; strict=False widens the support test to <=), to show
    that the stationarity check detects it."""
    e = math.exp(-0.5 * C * C) if shift else 0.0
    r2 = (C * H) ** 2
    out = seeds.copy()
    for s in range(seeds.shape[1]):
        m = seeds[:, s].copy()
        for _ in range(iterations):
            d = pts - m[:, None]
            u = np.sum(d * d, axis=0)
            inside = u < r2 if strict else u <= r2
            w = np.where(inside, np.exp(-u / (2 * H * H)) - e, 0.0)
            if w.sum() > 0:
                m = m + (d * w).sum(axis=1) / w.sum()
        out[:, s] = m
    return out


def converged_modes(oracle, pts, seeds):
    a, wa = run(oracle, pts, seeds, 400)
    b, _ = run(oracle, pts, seeds, 401)
    keep = (wa > 0) & (np.max(np.abs(a - b), axis=0) < 1e-12)
    assert keep.sum() >= 0.9 * seeds.shape[1]
    return a[:, keep]


def stationarity_error(m, pts):
    f = density(m, pts)
    g = fd_gradient(m, pts, 1e-4 * H)
    return float(np.linalg.norm(g) / (f / H))


def test_modes_are_stationary_points_of_the_density(oracle):
    pts = skewed_points()
    rng = np.random.default_rng(11)
    seeds = pts[:, rng.choice(pts.shape[1], 40, replace=False)].copy()
    modes = converged_modes(oracle, pts, seeds)
    errs = [stationarity_error(modes[:, s], pts) for s in range(modes.shape[1])]
    assert max(errs) < 1e-6, max(errs)
    # The modes are local maxima: the Hessian is negative definite.
    for s in range(modes.shape[1]):
        ev = np.linalg.eigvalsh(fd_hessian(modes[:, s], pts, 2e-3 * H))
        assert ev.max() < 0, ev
    # Power: weights without the shift converge elsewhere.
    wrong = numpy_variant(pts, modes, 400, shift=False)
    werr = [stationarity_error(wrong[:, s], pts) for s in range(wrong.shape[1])]
    assert np.median(werr) > 1e-3, np.median(werr)


def test_density_never_decreases_along_the_iterates(oracle):
    pts = skewed_points(seed=5)
    rng = np.random.default_rng(12)
    seeds = pts[:, rng.choice(pts.shape[1], 24, replace=False)] + \
        rng.normal(0, 0.5 * H, size=(pts.shape[0], 24))
    prev = np.array([density(seeds[:, s], pts) for s in range(seeds.shape[1])])
    rises = 0
    for it in range(1, 25):
        m, _ = run(oracle, pts, seeds, it)
        cur = np.array([density(m[:, s], pts) for s in range(seeds.shape[1])])
        assert np.all(cur >= prev - 1e-13 * np.maximum(prev, 1e-300)), (it, (cur - prev).min())
        rises += int(np.sum(cur > prev * (1 + 1e-12)))
        prev = cur
    assert rises > 24  # the iterates do move uphill, not stay put


def test_symmetric_points_keep_the_seed_at_the_centre(oracle):
    rng = np.random.default_rng(21)
    half = rng.normal(0, 0.7 * H, size=(6, 150))
    # Each point is followed by its mirror image through the origin, so every
    # pair of terms of the sums cancels exactly.
    pts = np.empty((6, 300))
    pts[:, 0::2] = half
    pts[:, 1::2] = -half
    seeds = np.zeros((6, 1))
    m, w = run(oracle, pts, seeds, 30)
    assert w[0] > 0 and np.all(m == 0.0)


def test_seed_on_a_symmetry_hyperplane_stays_on_it(oracle):
    rng = np.random.default_rng(22)
    half = rng.normal(0, 0.7 * H, size=(4, 200))
    half[1:] += 0.3 * H
    pts = np.empty((4, 400))
    pts[:, 0::2] = half
    pts[:, 1::2] = half
    pts[0, 1::2] = -half[0]       # reflection through the hyperplane x0 = 0
    seeds = np.zeros((4, 5))
    seeds[1:] = rng.normal(0, 0.5 * H, size=(3, 5))
    m, w = run(oracle, pts, seeds, 40)
    assert np.all(w > 0)
    assert np.all(m[0] == 0.0)
    assert np.max(np.abs(m[1:] - seeds[1:])) > 0.1 * H   # the others do move


@pytest.mark.parametrize("scale,inside", [(1 - 1e-9, True), (1 + 1e-9, False)])
def test_support_boundary(oracle, scale, inside):
    direction = np.array([3.0, -2.0, 6.0, 0.0, 0.0, 0.0]) / 7.0
    pts = (scale * C * H * direction)[:, None]
    seeds = np.zeros((6, 1))
    m, w = run(oracle, pts, seeds, 1)
    if inside:
        # One point inside the support: the weighted mean is that point.
        assert w[0] > 0
        np.testing.assert_allclose(m[:, 0], pts[:, 0], rtol=0, atol=1e-15)
    else:
        assert w[0] == 0 and np.all(m == 0.0)


def test_kinds_and_dimensions_are_separate(oracle):
    rng = np.random.default_rng(31)
    rigid = rng.normal(0, 0.6 * H, size=(6, 100))
    mirror = rng.normal(0, 0.6 * H, size=(6, 100))
    mirror[:4] += 0.5 * H
    mirror[4:] = rng.uniform(-5, 5, size=(2, 100))   # lanes a 4-D kind ignores
    pts = np.concatenate([rigid, mirror], axis=1)
    kind = np.concatenate([np.zeros(100, np.uint8), np.ones(100, np.uint8)])
    seeds = np.concatenate([rigid[:, :3], mirror[:, :3]], axis=1).copy()
    seed_kind = np.array([0, 0, 0, 1, 1, 1], np.uint8)
    m, _ = run(oracle, pts, seeds, 200, kind, seed_kind, 6, 4)
    m_r, _ = run(oracle, rigid, seeds[:, :3], 200)
    m_f, _ = run(oracle, mirror[:4], seeds[:4, 3:], 200)
    assert np.array_equal(m[:, :3], m_r)
    assert np.array_equal(m[:4, 3:], m_f)
    # Lanes past the kind's dimension are carried, not shifted.
    assert np.array_equal(m[4:, 3:], seeds[4:, 3:])
    # Stationarity of each kind's modes in its own space.
    for s in range(3):
        assert stationarity_error(m_r[:, s], rigid) < 1e-6
        assert stationarity_error(m_f[:, s], mirror[:4]) < 1e-6
