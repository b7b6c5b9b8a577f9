"""Inputs of the reference's own tests for the voting path, re-created from
their specification (file:line cited per builder).

The reference draws some inputs with std::uniform_real_distribution and
std::shuffle, whose results are implementation-defined; those fixtures keep
the same shapes, special values and thresholds but draw their values from
numpy's PCG64 with the stated seed.  Every builder returns plain numpy
arrays; the same arrays feed the oracle and the HIP path.
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------- math.cpp
def unit(v):
    n = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    if n > 1e-14:
        s = 1 / n
        return (v[0] * s, v[1] * s, v[2] * s)
    return (0.0, 0.0, 0.0)


def cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def reflection_matrix(n):
    r = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    for i in range(3):
        for j in range(3):
            r[i * 3 + j] -= 2 * n[i] * n[j]
    return r


def apply(m, v):
    return (m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
            m[6] * v[0] + m[7] * v[1] + m[8] * v[2])


# ------------------------------------------------------------- pairing
class PairFixture:
    """tests/cuda_pairing.cpp:15-29: vertex i at (i, (i%7)*.1, (i%11)*.2),
    selected = 0..n-1, kmin = .5, kmax = 1 (box size 1)."""

    def __init__(self, count: int):
        i = np.arange(count, dtype=np.float64)
        self.vertices = np.stack([i, (np.arange(count) % 7) * 0.1, (np.arange(count) % 11) * 0.2], 1)
        self.selected = np.arange(count, dtype=np.int32)
        self.feat = np.zeros((11, count))
        self.feat[9] = 0.5
        self.feat[10] = 1.0
        self.box = 1.0

    @property
    def kmin(self):
        return self.feat[9]

    @property
    def kmax(self):
        return self.feat[10]


def pairing_random(seed: int = 123) -> PairFixture:
    """tests/cuda_pairing.cpp:63-76: 777 samples, kmin/kmax uniform in
    (-2, 2), kmin = 0 when i%13 == 0, kmax = 1e-12 when i%17 == 0, selected
    order shuffled (values drawn from PCG64(seed))."""
    f = PairFixture(777)
    g = np.random.default_rng(seed)
    v = g.uniform(-2, 2, (777, 2))
    f.feat[9] = v[:, 0]
    f.feat[10] = v[:, 1]
    i = np.arange(777)
    f.feat[9][i % 13 == 0] = 0.0
    f.feat[10][i % 17 == 0] = 1e-12
    g.shuffle(f.selected)
    return f


RANDOM_THRESHOLDS = [1e-8, 0.01, 0.15, 0.7, math.nextafter(1.0, 0.0), 1.0]


def pairing_boundary() -> PairFixture:
    """tests/cuda_pairing.cpp:90-102."""
    f = PairFixture(12)
    tiny = math.nextafter(0.75, 0.0)
    k = [0.75, 1, tiny, -0.75, -1, 0, -0.0, 1e-12, -1e-12, 1e-10, 1e200, 1e200]
    f.feat[9] = k
    f.feat[10] = k
    f.vertices[0] = (0, 0, 0)
    f.vertices[1] = (0.03, 0, 0)
    f.vertices[2] = (math.nextafter(0.03, 0.0), 0, 0)
    return f


def pairing_edge(t: float) -> PairFixture:
    """tests/cuda_pairing.cpp:105-112: signatures at the window edge 1/(1-t)."""
    f = PairFixture(4)
    target = 1 / (1 - t)
    k = [1.0, target, math.nextafter(target, 0.0), math.nextafter(target, math.inf)]
    f.feat[9] = k
    f.feat[10] = k
    return f


EDGE_THRESHOLDS = [0.01, 0.15, 0.5, 0.9, math.nextafter(1.0, 0.0)]


def pairing_rounded() -> PairFixture:
    """tests/cuda_pairing.cpp:116-120: one pair at t = nextafter(1, 0)."""
    f = PairFixture(2)
    f.feat[9] = [1.0, 13510798882111490.0]
    f.feat[10] = f.feat[9]
    return f


def pairing_zeros() -> PairFixture:
    """tests/cuda_pairing.cpp:121-127: signed zeros give 190 pairs at t=.15."""
    f = PairFixture(20)
    f.feat[9] = [(-0.0 if i % 2 else 0.0) for i in range(20)]
    f.feat[10] = 0.0
    return f


def pairing_sparse(count: int = 65536) -> PairFixture:
    """tests/cuda_pairing.cpp:154-164: disjoint signatures, t=1e-10, 0 pairs."""
    f = PairFixture(count)
    f.feat[9] = np.arange(1, count + 1, dtype=np.float64)
    f.feat[10] = f.feat[9]
    return f


def pairing_benchmark(count: int = 10000, groups: int = 100) -> PairFixture:
    """benchmarks/pairing.cpp:12-36: vertex (i, i%7, i%11), k = (37 i) % 100 + 1,
    kmin = k, kmax = 2k, t = 1e-6 -> 495,000 pairs."""
    f = PairFixture(count)
    i = np.arange(count)
    f.vertices = np.stack([i.astype(np.float64), (i % 7).astype(np.float64),
                           (i % 11).astype(np.float64)], 1)
    k = ((i * 37) % groups + 1).astype(np.float64)
    f.feat[9] = k
    f.feat[10] = 2 * k
    return f


# ------------------------------------------------------- transformations
def mirror_fixture():
    """tests/cuda_backend.cpp:236-264: 514 point pairs mirrored across an
    oblique plane with a 1e-16 normal component, principal directions with
    sign flips every 4th vertex; pairs (2i, 2i+1) and (2i+1, 2i);
    context box 20, centre (3, -1, 2).  Returns vertices, features, pairs,
    box, centre, plane normal n, direction d, x = n x d."""
    n = unit((1e-16, 1, 2))
    d = unit(cross(n, (1, 0, 0)))
    x = cross(n, d)
    M = reflection_matrix(n)
    verts = []
    for i in range(514):
        p = (i * 0.01, -2 - 0.001 * i, 0.13)
        verts.append(p)
        verts.append(apply(M, p))
    V = len(verts)
    feat = np.zeros((11, V))
    for i in range(V):
        normal, minimum = n, d
        if i % 4 == 3:
            minimum = (-minimum[0], -minimum[1], -minimum[2])
        if i % 2:
            normal = apply(M, normal)
            minimum = apply(M, minimum)
        mx = cross(normal, minimum)
        feat[0:3, i] = normal
        feat[3:6, i] = minimum
        feat[6:9, i] = mx
        feat[9, i] = 0.2
        feat[10, i] = 1.0
    pa, pb = [], []
    for i in range(514):
        pa += [2 * i, 2 * i + 1]
        pb += [2 * i + 1, 2 * i]
    return (np.array(verts), feat, np.array(pa, np.int32), np.array(pb, np.int32), 20.0,
            np.array([3.0, -1.0, 2.0]), n, d, x)


def mixed_features(feat: np.ndarray, n):
    """tests/cuda_backend.cpp:278-281: every 4th vertex from 1 gets the
    unreflected normal, giving mixed 2/3 vote counts."""
    f = feat.copy()
    for i in range(1, f.shape[1], 4):
        f[0:3, i] = n
        f[6:9, i] = cross(n, tuple(f[3:6, i]))
    return f


def parallel_features(count: int, n, d, x):
    """tests/cuda_backend.cpp:297-298: all frames {n, d, x}: 2 votes per pair."""
    f = np.zeros((11, count))
    f[0:3] = np.array(n)[:, None]
    f[3:6] = np.array(d)[:, None]
    f[6:9] = np.array(x)[:, None]
    f[9] = 0.2
    f[10] = 1.0
    return f


def select_fixture(count: int):
    """tests/cuda_backend.cpp:320-329: valid unless i%7 == 0, kmin .2 (1 when
    i%5 == 0), kmax 1, both 1e-12 when i%11 == 0, weight 0 when i%13 == 0."""
    i = np.arange(count)
    valid = (i % 7 != 0).astype(np.uint8)
    kmin = np.where(i % 5 == 0, 1.0, 0.2)
    kmax = np.ones(count)
    kmin[i % 11 == 0] = 1e-12
    kmax[i % 11 == 0] = 1e-12
    w = np.ones(count)
    w[i % 13 == 0] = 0
    return w, kmin, kmax, valid


def strict_prune_fixture(count: int):
    """tests/cuda_backend.cpp:338-347: prune .75, kmax 1, kmin .75 /
    nextafter(.75, 0) / -1 by i%3: exactly 2/3 are kept."""
    i = np.arange(count)
    kmin = np.where(i % 3 == 0, 0.75, np.where(i % 3 == 1, math.nextafter(0.75, 0.0), -1.0))
    return np.ones(count), kmin, np.ones(count), np.ones(count, np.uint8)


# ------------------------------------------------------------ clustering
def cluster_votes(count: int):
    """tests/cuda_clustering.cpp:12-23 and benchmarks/clustering.cpp:17-37:
    alternating rigid/reflection votes around 4 modes (d in {3, 5} plus
    .001 sin i), bandwidth .03, box 10.  Returns kind, descriptor (12, V),
    rotation (9, V)."""
    kind = np.zeros(count, np.uint8)
    desc = np.zeros((12, count))
    rot = np.zeros((9, count))
    ident = [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    refl = reflection_matrix((1.0, 0.0, 0.0))
    for i in range(count):
        r = i % 2 == 1
        d = 3 + (i % 4 // 2) * 2 + 0.001 * math.sin(i)
        kind[i] = 1 if r else 0
        R = refl if r else ident
        rot[:, i] = R
        if r:
            desc[0:4, i] = (1, 0, 0, d / 10)
        else:
            for k in range(9):
                desc[k, i] = R[k] / 2
            desc[9, i] = d / 10
    return kind, desc, rot
