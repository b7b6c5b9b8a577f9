"""Synthetic height-field meshes for the voting-core workloads.

A mesh is a z = f(x, y) height field on an n x n vertex grid over [-L, L]^2,
triangulated with alternating diagonals: cell (i, j) (i along x, j along y)
is split on its main diagonal (i, j)-(i+1, j+1) when i + j is even and on the
anti-diagonal otherwise.  Triangles are wound counter-clockwise seen from +z.
Vertex (i, j) has index j * n + i.

f is a sum of motif copies plus iid Gaussian vertex noise.  The motif is a sum
of 8 anisotropic Gaussians drawn from numpy's PCG64 generator with seed 0:
amplitude U(.02, .08), sigma_x and sigma_y U(.03, .12), centre uniform in the
disc of radius .25, orientation U(0, pi).  A copy with centre C, rotation theta
about z and an optional x-mirror evaluates the motif at
mirror_x(R(-theta) (p - C)).  Each Gaussian term is evaluated where its squared
Mahalanobis radius is below 81 (9 sigma) and is zero beyond.

Workloads (sizes as in the reference's headline configurations):
  * cfg1_mesh(): 96 x 96 grid on [-1, 1]^2 (18,050 triangles); copies at
    (.5,.5), (-.5,.5), (-.5,-.5), (.5,-.5) rotated 0/90/180/270 degrees and an
    x-mirrored copy at the origin; noise N(0, 1e-4) from seed 2.
  * cfg2_mesh(): 1024 x 1024 grid on [-4, 4]^2 (2,093,058 triangles); 64
    copies on the 8 x 8 lattice of cell centres, each rotated U(0, 2 pi) and
    x-mirrored with probability .5 (seed 1); noise N(0, 1e-4) from seed 2.
  * batch_spec() / batch_mesh(): mesh m of a batch draws, from PCG64 seeded
    with (seed, m), a grid side log-uniform in [32, 512] (rounded), 1-16
    motif copies placed on the first cells of a k x k lattice of unit cells
    (k = ceil(sqrt(copies)), domain [-k/2, k/2]^2), each rotated U(0, 2 pi)
    and x-mirrored with probability .5, and N(0, 1e-4) noise.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Motif:
    amplitude: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    phi: np.ndarray

    @staticmethod
    def make(seed: int = 0, count: int = 8) -> "Motif":
        g = np.random.default_rng(seed)
        amp = g.uniform(0.02, 0.08, count)
        sx = g.uniform(0.03, 0.12, count)
        sy = g.uniform(0.03, 0.12, count)
        r = 0.25 * np.sqrt(g.uniform(0.0, 1.0, count))
        a = g.uniform(0.0, 2 * math.pi, count)
        phi = g.uniform(0.0, math.pi, count)
        return Motif(amp, sx, sy, r * np.cos(a), r * np.sin(a), phi)

    def extent(self) -> float:
        return float(np.max(np.hypot(self.cx, self.cy)) + 9 * max(self.sx.max(), self.sy.max()))

    def evaluate(self, u: np.ndarray, v: np.ndarray) -> np.ndarray:
        out = np.zeros_like(u)
        for k in range(len(self.amplitude)):
            du, dv = u - self.cx[k], v - self.cy[k]
            c, s = math.cos(self.phi[k]), math.sin(self.phi[k])
            a = (c * du + s * dv) / self.sx[k]
            b = (-s * du + c * dv) / self.sy[k]
            q = a * a + b * b
            m = q < 81.0
            out[m] += self.amplitude[k] * np.exp(-0.5 * q[m])
        return out


@dataclass
class Copy:
    cx: float
    cy: float
    theta: float
    mirror: bool


@dataclass
class HeightField:
    n: int
    half: float
    vertices: np.ndarray  # (N, 3) float64
    faces: np.ndarray     # (F, 3) int32
    copies: list = field(default_factory=list)

    @property
    def x(self):
        return np.ascontiguousarray(self.vertices[:, 0])

    @property
    def y(self):
        return np.ascontiguousarray(self.vertices[:, 1])

    @property
    def z(self):
        return np.ascontiguousarray(self.vertices[:, 2])


def grid_faces(n: int) -> np.ndarray:
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    i = i.ravel()
    j = j.ravel()
    v00 = j * n + i
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    even = ((i + j) % 2) == 0
    t1 = np.where(even[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    t2 = np.where(even[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    faces = np.empty((2 * len(i), 3), dtype=np.int32)
    faces[0::2] = t1
    faces[1::2] = t2
    return faces


def height_field(n: int, half: float, copies: list, motif: Motif, noise_sigma: float,
                 noise_seed: int) -> HeightField:
    ax = -half + (2 * half) * np.arange(n) / (n - 1)
    X, Y = np.meshgrid(ax, ax, indexing="xy")
    X = X.ravel()
    Y = Y.ravel()
    Z = np.zeros_like(X)
    reach = motif.extent()
    for c in copies:
        m = (np.abs(X - c.cx) < reach) & (np.abs(Y - c.cy) < reach)
        px, py = X[m] - c.cx, Y[m] - c.cy
        ct, st = math.cos(c.theta), math.sin(c.theta)
        u = ct * px + st * py
        v = -st * px + ct * py
        if c.mirror:
            u = -u
        Z[m] += motif.evaluate(u, v)
    if noise_sigma > 0:
        Z += np.random.default_rng(noise_seed).normal(0.0, noise_sigma, Z.shape)
    verts = np.stack([X, Y, Z], 1)
    return HeightField(n, half, verts, grid_faces(n), copies)


def cfg1_mesh() -> HeightField:
    copies = [Copy(0.5, 0.5, 0.0, False), Copy(-0.5, 0.5, 0.5 * math.pi, False),
              Copy(-0.5, -0.5, math.pi, False), Copy(0.5, -0.5, 1.5 * math.pi, False),
              Copy(0.0, 0.0, 0.0, True)]
    return height_field(96, 1.0, copies, Motif.make(0), 1e-4, 2)


def lattice_copies(k: int, spacing: float, seed: int) -> list:
    g = np.random.default_rng(seed)
    theta = g.uniform(0.0, 2 * math.pi, k * k)
    mirror = g.uniform(0.0, 1.0, k * k) < 0.5
    out = []
    for j in range(k):
        for i in range(k):
            c = (-(k - 1) / 2 + i) * spacing, (-(k - 1) / 2 + j) * spacing
            out.append(Copy(c[0], c[1], float(theta[j * k + i]), bool(mirror[j * k + i])))
    return out


def cfg2_mesh(n: int = 1024) -> HeightField:
    return height_field(n, 4.0, lattice_copies(8, 1.0, 1), Motif.make(0), 1e-4, 2)


def uniform_samples(count: int, nvertices: int, seed: int) -> np.ndarray:
    """`count` distinct vertex ids drawn uniformly without replacement, ascending."""
    g = np.random.default_rng(seed)
    count = min(count, nvertices)
    return np.sort(g.choice(nvertices, size=count, replace=False)).astype(np.int32)


def magnitude_keep(samples: np.ndarray, kmin: np.ndarray, kmax: np.ndarray,
                   mean_edge: float, threshold: float = 0.5) -> np.ndarray:
    """Keep samples with (|k1| + |k2|) / mean_edge >= threshold.

    The reference's curvature tensor is a dimensionless dihedral-angle average
    (src/cpu/curvature.cpp:21-37), so dividing by the mean edge length gives a
    curvature estimate in 1/length units.
    """
    s = samples
    mag = (np.abs(kmin[s]) + np.abs(kmax[s])) / mean_edge
    return s[mag >= threshold]


@dataclass
class BatchMeshSpec:
    side: int
    copies: list
    half: float
    noise_seed: int


def batch_spec(index: int, seed: int = 4, min_side: int = 32, max_side: int = 512) -> BatchMeshSpec:
    """Parameters of mesh `index` of a batch (see the module docstring)."""
    g = np.random.default_rng([seed, index])
    side = int(round(math.exp(g.uniform(math.log(min_side), math.log(max_side)))))
    count = int(g.integers(1, 17))
    k = int(math.ceil(math.sqrt(count)))
    theta = g.uniform(0.0, 2 * math.pi, count)
    mirror = g.uniform(0.0, 1.0, count) < 0.5
    copies = []
    for c in range(count):
        i, j = c % k, c // k
        copies.append(Copy(-(k - 1) / 2 + i, -(k - 1) / 2 + j, float(theta[c]), bool(mirror[c])))
    return BatchMeshSpec(side, copies, k / 2.0, int(g.integers(0, 2**31)))


def batch_mesh(spec: BatchMeshSpec) -> HeightField:
    return height_field(spec.side, spec.half, spec.copies, Motif.make(0), 1e-4, spec.noise_seed)


def batch_samples(nvertices: int, index: int, seed: int = 4) -> np.ndarray:
    """clamp(0.05 N, 200, 20000) uniform vertex samples of mesh `index`."""
    count = int(min(max(round(0.05 * nvertices), 200), 20000))
    return uniform_samples(count, nvertices, seed * 1000003 + index)
