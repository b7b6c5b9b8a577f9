"""CPU checks of the mode-refinement reference (oracle/refine_ref.py): the
analytic FP64 gradients against central finite differences of the energy,
and the closest point against the reference's own exact point-triangle
distance (oracle/_ref) and brute force."""
import numpy as np
import pytest

from oracle import refine_ref as rr


def _data(kind, seed, n=400):
    g = np.random.default_rng(seed)
    P = g.normal(0, 0.5, (n, 3))
    Q = P + g.normal(0, 0.05, (n, 3))
    Nn = g.normal(0, 1, (n, 3))
    Nn /= np.linalg.norm(Nn, axis=1, keepdims=True)
    R0 = rr.rotation(g.normal(0, 0.7, 3))
    t0 = g.normal(0, 0.2, 3)
    return P, Q, Nn, R0, t0


def _fd(kind, theta, P, Q, Nn, R0, t0, h=1e-4):
    out = np.zeros_like(theta)
    for k in range(len(theta)):
        e = np.zeros_like(theta)
        e[k] = h
        out[k] = (rr.energy(kind, theta + e, P, Q, Nn, R0, t0) -
                  rr.energy(kind, theta - e, P, Q, Nn, R0, t0)) / (2 * h)
    return out


@pytest.mark.parametrize("w_scale", [0.0, 1e-4, 0.05, 0.8])
def test_rigid_gradient_matches_finite_differences(w_scale):
    P, Q, Nn, R0, t0 = _data(0, 3)
    g = np.random.default_rng(5)
    theta = np.concatenate([g.normal(0, 1, 3) * w_scale, g.normal(0, 0.05, 3)])
    a = rr.gradient(0, theta, P, Q, Nn, R0, t0)
    f = _fd(0, theta, P, Q, Nn, R0, t0)
    assert np.linalg.norm(a - f) <= 1e-6 * (1 + np.linalg.norm(f))


def test_mirror_gradient_matches_finite_differences():
    P, Q, Nn, _, _ = _data(1, 4)
    g = np.random.default_rng(6)
    theta = np.concatenate([g.normal(0, 1, 3), [0.1]])
    a = rr.gradient(1, theta, P, Q, Nn)
    f = _fd(1, theta, P, Q, Nn, None, None)
    assert np.linalg.norm(a - f) <= 1e-6 * (1 + np.linalg.norm(f))
    # scaling m does not move the plane: the gradient is orthogonal to m
    assert abs(a[:3] @ theta[:3]) <= 1e-9 * (1 + np.linalg.norm(a))


def test_rotation_is_orthonormal_and_matches_small_angle():
    for w in (np.zeros(3), np.array([1e-5, -2e-5, 3e-5]), np.array([0.3, -1.2, 2.0])):
        R = rr.rotation(w)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)
        assert np.allclose(R @ w, w, atol=1e-14)  # w is the axis


def test_closest_point_matches_reference_distance(ref):
    import ctypes as C
    f = ref.lib.ref_point_triangle_distance2
    f.restype = C.c_double
    arr = lambda v: (C.c_double * 3)(*v)
    g = np.random.default_rng(7)
    for k in range(3000):
        a, b, c = g.normal(0, 1, (3, 3))
        if k % 10 == 0:
            c = a + 0.5 * (b - a)  # degenerate triangle
        p = g.normal(0, 1.5, 3)
        q = rr.closest_on_triangle(p, a, b, c)
        want = f(arr(p), arr(a), arr(b), arr(c))
        assert (p - q) @ (p - q) == pytest.approx(want, rel=1e-10, abs=1e-14)


def test_closest_brute_is_the_minimum():
    g = np.random.default_rng(8)
    V = g.normal(0, 1, (30, 3))
    F = g.integers(0, 30, (40, 3))
    for _ in range(50):
        p = g.normal(0, 1, 3)
        q, d2, f = rr.closest_brute(p, V, F)
        assert d2 == pytest.approx(min(((p - rr.closest_on_triangle(p, *V[t])) ** 2).sum()
                                       for t in F))


def test_vectorised_closest_points_match_voronoi_form():
    g = np.random.default_rng(21)
    V = g.normal(0, 1, (30, 3))
    F = np.array([g.choice(30, 3, replace=False) for _ in range(40)])
    P = np.concatenate([g.normal(0, 1.5, (150, 3)), V[F[:20]].mean(1)])
    q, d2, f = rr.closest_points(P, V, F)
    for i, p in enumerate(P):
        bq, bd2, bf = rr.closest_brute(p, V, F)
        assert abs(d2[i] - bd2) <= 1e-12 * (1 + bd2)
        assert np.linalg.norm(q[i] - bq) <= 1e-9 or abs(d2[i] - bd2) <= 1e-12
