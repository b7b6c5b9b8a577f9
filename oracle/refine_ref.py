"""FP64 CPU reference of the mode-refinement energy (test infrastructure only:
tests and the benchmark's checks import it; the product never does).

No reference counterpart exists for the energy or its gradient (the detector
refines nothing); the closest-point query restates the reference's exact
point-triangle distance (src/core/math.cpp:72-112, Voronoi-region method of
Ericson, Real-Time Collision Detection, 2005, section 5.1.5) as a closest
point.  Pinning: the analytic gradients are checked against central finite
differences of the energy (tests/test_refine.py); the closest point's squared
distance against the reference's own point_triangle_distance2 from
oracle/_ref where that builds, and against brute force over all faces.

Coordinates are relative to the model centre.
  rigid:  T(p) = R(w) (R0 p + t0) + v,  theta = (w, v)
  mirror: T(p) = p - 2 (u.p + d) u,     u = m / |m|, theta = (m, d)
  E = sum_j (n_j . (T(p_j) - q_j))^2
"""
import numpy as np


def skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def rodrigues_coeffs(t):
    """A, B of R = I + A K + B K^2 and A'/t, B'/t (series below 1e-3)."""
    if t < 1e-3:
        t2 = t * t
        return (1 - t2 / 6, 0.5 - t2 / 24, -1 / 3 + t2 / 30, -1 / 12 + t2 / 180)
    s, c = np.sin(t), np.cos(t)
    return s / t, (1 - c) / t**2, (t * c - s) / t**3, (t * s - 2 * (1 - c)) / t**4


def rotation(w):
    w = np.asarray(w, dtype=np.float64)
    A, B, _, _ = rodrigues_coeffs(np.linalg.norm(w))
    K = skew(w)
    return np.eye(3) + A * K + B * K @ K


def transform(kind, theta, P, R0=None, t0=None):
    theta = np.asarray(theta, dtype=np.float64)
    if kind == 0:
        Y = P @ R0.T + t0
        return Y @ rotation(theta[:3]).T + theta[3:6]
    u = theta[:3] / np.linalg.norm(theta[:3])
    s = P @ u + theta[3]
    return P - 2 * s[:, None] * u[None, :]


def energy(kind, theta, P, Q, Nn, R0=None, t0=None):
    r = np.einsum("ij,ij->i", Nn, transform(kind, theta, P, R0, t0) - Q)
    return float(r @ r)


def gradient(kind, theta, P, Q, Nn, R0=None, t0=None):
    """Analytic dE/dtheta (FP64), written from the definitions above."""
    theta = np.asarray(theta, dtype=np.float64)
    if kind == 0:
        w, v = theta[:3], theta[3:6]
        Y = P @ R0.T + t0
        A, B, dA, dB = rodrigues_coeffs(np.linalg.norm(w))
        KY = np.cross(w[None, :], Y)
        KKY = np.cross(w[None, :], KY)
        T = Y + A * KY + B * KKY + v
        r = np.einsum("ij,ij->i", Nn, T - Q)
        g1 = np.cross(Y, Nn)
        g2 = np.cross(KY, Nn) + np.cross(Y, np.cross(Nn, w[None, :]))
        s = dA * np.einsum("ij,ij->i", Nn, KY) + dB * np.einsum("ij,ij->i", Nn, KKY)
        Jw = A * g1 + B * g2 + s[:, None] * w[None, :]
        J = np.concatenate([Jw, Nn], axis=1)
    else:
        m, d = theta[:3], theta[3]
        nm = np.linalg.norm(m)
        u = m / nm
        sp = P @ u + d
        T = P - 2 * sp[:, None] * u[None, :]
        r = np.einsum("ij,ij->i", Nn, T - Q)
        nu = Nn @ u
        gu = -2 * (nu[:, None] * P + sp[:, None] * Nn)
        gm = (gu - (gu @ u)[:, None] * u[None, :]) / nm
        J = np.concatenate([gm, (-2 * nu)[:, None]], axis=1)
    return 2 * (r[:, None] * J).sum(0)


def _closest_on_segment(p, a, b):
    d = b - a
    l = d @ d
    t = np.clip((p - a) @ d / l, 0.0, 1.0) if l > 0 else 0.0
    return a + d * t


def closest_on_triangle(p, a, b, c):
    """Closest point of triangle abc to p (Voronoi regions), FP64."""
    ab, ac = b - a, c - a
    nn = np.cross(ab, ac)
    if nn @ nn < 1e-30:
        cands = [_closest_on_segment(p, a, b), _closest_on_segment(p, a, c),
                 _closest_on_segment(p, b, c)]
        return min(cands, key=lambda q: (p - q) @ (p - q))
    ap = p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + ab * (d1 / (d1 - d3))
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + ac * (d2 / (d2 - d6))
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)))
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def closest_brute(p, V, F):
    """Closest surface point over every face (small meshes): (q, d2, face)."""
    best, bq, bf = np.inf, None, -1
    for f, (i, j, k) in enumerate(F):
        q = closest_on_triangle(p, V[i], V[j], V[k])
        d = (p - q) @ (p - q)
        if d < best:
            best, bq, bf = d, q, f
    return bq, best, bf


def _seg_vec(P, a, b):
    d = b - a
    l = np.einsum("...k,...k->...", d, d)
    t = np.clip(np.einsum("...k,...k->...", P - a, d) / np.where(l > 0, l, 1), 0, 1)
    return a + d * t[..., None]


def closest_points(P, V, F, chunk=256):
    """Closest surface point of every row of P (n x 3) over all faces of a
    non-degenerate mesh, vectorised: inside the face's prism the projection
    onto its plane, otherwise the nearest of its three edges.  Returns
    (q, d2, face); ties go to the lowest face id.  A small-mesh checker, as
    closest_brute, in a form that formulates the same minimum differently
    (projection / edges rather than Voronoi regions)."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nn = np.cross(b - a, c - a)
    n2 = np.einsum("fk,fk->f", nn, nn)
    qs, ds, fs = [], [], []
    for s in range(0, len(P), chunk):
        p = P[s:s + chunk, None, :]
        h = np.einsum("nfk,fk->nf", p - a, nn) / n2
        proj = p - h[..., None] * nn
        # barycentric signs of the projection against each edge
        e0 = np.einsum("nfk,fk->nf", np.cross(b - a, proj - a), nn)
        e1 = np.einsum("nfk,fk->nf", np.cross(c - b, proj - b), nn)
        e2 = np.einsum("nfk,fk->nf", np.cross(a - c, proj - c), nn)
        inside = (e0 >= 0) & (e1 >= 0) & (e2 >= 0)
        best = proj.copy()
        bd = np.where(inside, np.einsum("nfk,nfk->nf", p - proj, p - proj), np.inf)
        for x, y in ((a, b), (b, c), (c, a)):
            q = _seg_vec(p, x, y)
            d = np.einsum("nfk,nfk->nf", p - q, p - q)
            take = ~inside & (d < bd)
            best[take] = q[take]
            bd = np.where(take, d, bd)
        f = np.argmin(bd, axis=1)
        r = np.arange(len(f))
        qs.append(best[r, f])
        ds.append(bd[r, f])
        fs.append(f)
    return np.concatenate(qs), np.concatenate(ds), np.concatenate(fs)
