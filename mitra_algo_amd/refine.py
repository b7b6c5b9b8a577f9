"""Refinement of detected modes by point-to-plane alignment.

Each mode (a rigid motion or a mirror plane found by mean-shift) gets a patch
of surface points P.  With q_j the closest mesh point of T(p_j) and n_j its
face normal (held fixed while the parameters move), the energy of the mode is
E = sum_j (n_j . (T(p_j) - q_j))^2, in coordinates relative to the model
centre:
    rigid  (6-DoF twist): T(p) = exp([w]x) (R0 p + t0) + v, theta = (w, v)
    mirror (plane):       T(p) = p - 2 (u.p + d) u, u = m / |m|, theta = (m, d)
The energy, its gradient and the Gauss-Newton matrix come from one FP32 HIP
kernel per kind (mvs_dev_align_energy), exposed to torch through
AlignEnergy, a torch.autograd.Function whose backward is that kernel's
adjoint.  Closest points come from the session's BVH (mvs_bvh_build /
mvs_dev_closest_points).  refine() runs damped Gauss-Newton steps,
recomputing the closest points at the start of each step.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import capi

RIGID, MIRROR = 0, 1
NPARAM = {RIGID: 6, MIRROR: 4}


def rotation_torch(w: torch.Tensor) -> torch.Tensor:
    """Batched exp([w]x) in FP64 (w: M x 3), Rodrigues with a series at small angle."""
    w = w.double()
    t = torch.linalg.norm(w, dim=1)
    small = t < 1e-6
    ts = torch.where(small, torch.ones_like(t), t)
    A = torch.where(small, 1 - t * t / 6, torch.sin(ts) / ts)
    B = torch.where(small, 0.5 - t * t / 24, (1 - torch.cos(ts)) / (ts * ts))
    K = torch.zeros(w.shape[0], 3, 3, dtype=torch.float64, device=w.device)
    K[:, 0, 1], K[:, 0, 2] = -w[:, 2], w[:, 1]
    K[:, 1, 0], K[:, 1, 2] = w[:, 2], -w[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -w[:, 1], w[:, 0]
    KK = (K[:, :, :, None] * K[:, None, :, :]).sum(2)  # K @ K without batched GEMM
    eye = torch.eye(3, dtype=torch.float64, device=w.device)
    return eye + A[:, None, None] * K + B[:, None, None] * KK


@dataclass
class Problem:
    """Modes of one kind: patch points (M x 3 x N, centre-relative FP32), the
    rigid base transforms (M x 12: R0 row-major, t0) or the mirror planes in
    theta, and the closest points of the current transforms."""
    sess: capi.Session
    kind: int
    pts: torch.Tensor
    base: torch.Tensor | None
    theta: torch.Tensor            # M x P, FP32 (rigid: the twist about the base)
    q: torch.Tensor = field(init=False)
    nrm: torch.Tensor = field(init=False)

    def __post_init__(self):
        self.q = torch.empty_like(self.pts)
        self.nrm = torch.empty_like(self.pts)

    @property
    def modes(self) -> int:
        return self.pts.shape[0]

    @property
    def npts(self) -> int:
        return self.pts.shape[2]

    def transformed(self, theta: torch.Tensor) -> torch.Tensor:
        """T(p) of every patch point (FP32, M x 3 x N)."""
        p = self.pts
        if self.kind == RIGID:
            R0 = self.base[:, :9].reshape(-1, 3, 3)
            y = (R0[:, :, :, None] * p[:, None, :, :]).sum(2) + self.base[:, 9:, None]
            R = rotation_torch(theta[:, :3]).float()
            return (R[:, :, :, None] * y[:, None, :, :]).sum(2) + theta[:, 3:6, None]
        u = theta[:, :3] / torch.linalg.norm(theta[:, :3], dim=1, keepdim=True)
        s = (u[:, :, None] * p).sum(1) + theta[:, 3:4]
        return p - 2 * s[:, None, :] * u[:, :, None]

    def update_closest(self, theta: torch.Tensor) -> torch.Tensor:
        """Closest mesh points and normals of T(p); returns the squared distances."""
        tp = self.transformed(theta).contiguous()
        M, N = self.modes, self.npts
        # the kernel takes one 3 x n SoA: lay the modes side by side
        flat = tp.permute(1, 0, 2).reshape(3, M * N).contiguous()
        q = torch.empty_like(flat)
        nrm = torch.empty_like(flat)
        d2 = torch.empty(M * N, dtype=torch.float32, device=flat.device)
        self.sess.dev_closest_points(M * N, flat, q, nrm, d2)
        self.q = q.reshape(3, M, N).permute(1, 0, 2).contiguous()
        self.nrm = nrm.reshape(3, M, N).permute(1, 0, 2).contiguous()
        return d2.reshape(M, N)

    def evaluate(self, theta: torch.Tensor, want_hessian: bool = True):
        """(E, g, H) of every mode at theta with the current closest points."""
        M, P = self.modes, NPARAM[self.kind]
        theta = theta.detach().float().contiguous()
        E = torch.empty(M, dtype=torch.float32, device=theta.device)
        g = torch.empty(M, P, dtype=torch.float32, device=theta.device)
        H = torch.empty(M, P, P, dtype=torch.float32, device=theta.device) if want_hessian \
            else None
        self.sess.dev_align_energy(self.kind, M, self.npts, theta, self.base, self.pts,
                                   self.q, self.nrm, E, g, H)
        return E, g, H


class AlignEnergy(torch.autograd.Function):
    """E(theta) per mode; the backward pass is the FP32 adjoint kernel (the
    forward launch computes the energy and its gradient together)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, problem: Problem) -> torch.Tensor:
        E, g, _ = problem.evaluate(theta, want_hessian=False)
        ctx.save_for_backward(g)
        return E

    @staticmethod
    def backward(ctx, grad_energy: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_energy[:, None].to(g.dtype) * g, None


def align_energy(theta: torch.Tensor, problem: Problem) -> torch.Tensor:
    return AlignEnergy.apply(theta, problem)


def refine(problem: Problem, steps: int = 10, damping: float = 1e-3):
    """Damped Gauss-Newton (Levenberg-Marquardt) steps; closest points are
    recomputed at the start of every step.  Rigid twists are folded into the
    base after each accepted step.  Returns the per-step energies (steps + 1
    rows: before the first step and after each, with fresh closest points)."""
    M, P = problem.modes, NPARAM[problem.kind]
    dev = problem.pts.device
    lam = torch.full((M,), damping, dtype=torch.float64, device=dev)
    history = []
    for it in range(steps + 1):
        problem.update_closest(problem.theta)
        E0, g, H = problem.evaluate(problem.theta)
        history.append(E0.double().cpu().numpy())
        if it == steps:
            break
        Hd = H.double()
        diag = torch.diagonal(Hd, dim1=1, dim2=2)
        A = Hd + torch.diag_embed(lam[:, None] * (diag + 1e-12 * (1 + diag.sum(1, keepdim=True))))
        delta = -torch.linalg.solve(A, g.double()[:, :, None])[:, :, 0]
        trial = problem.theta.double() + delta
        E1, _, _ = problem.evaluate(trial.float(), want_hessian=False)
        ok = (E1 < E0) & torch.isfinite(E1)
        lam = torch.where(ok, lam / 3, lam * 4)
        if problem.kind == RIGID:
            R = rotation_torch(delta[:, :3])
            R0 = problem.base[:, :9].double().reshape(-1, 3, 3)
            t0 = problem.base[:, 9:].double()
            R0n = (R[:, :, :, None] * R0[:, None, :, :]).sum(2)
            t0n = (R * t0[:, None, :]).sum(2) + delta[:, 3:6]
            nb = torch.cat([R0n.reshape(-1, 9), t0n], 1).float()
            problem.base = torch.where(ok[:, None], nb, problem.base).contiguous()
        else:
            m = trial[:, :3] / torch.linalg.norm(trial[:, :3], dim=1, keepdim=True)
            nt = torch.cat([m, trial[:, 3:4]], 1).float()
            problem.theta = torch.where(ok[:, None], nt, problem.theta).contiguous()
    return np.stack(history)


def surface_samples(vertices: np.ndarray, faces: np.ndarray, count: int,
                    seed: int) -> np.ndarray:
    """`count` points uniform on the surface (area-weighted faces, uniform
    barycentric coordinates), PCG64(seed); (count, 3) FP64."""
    g = np.random.default_rng(seed)
    a, b, c = (vertices[faces[:, k]] for k in range(3))
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    f = g.choice(len(faces), size=count, p=area / area.sum())
    r1, r2 = g.uniform(size=count), g.uniform(size=count)
    s = np.sqrt(r1)
    return ((1 - s)[:, None] * a[f] + (s * (1 - r2))[:, None] * b[f] +
            (s * r2)[:, None] * c[f])


def merge_modes(points: np.ndarray, weight: np.ndarray, radius: float) -> np.ndarray:
    """Indices of distinct modes: by decreasing weight, a mode is kept when no
    kept mode lies within `radius` (points: D x K)."""
    order = np.argsort(-weight, kind="stable")
    kept = []
    for i in order:
        if not np.isfinite(points[:, i]).all():
            continue
        if kept:
            d = np.linalg.norm(points[:, kept] - points[:, i:i + 1], axis=0)
            if d.min() <= radius:
                continue
        kept.append(i)
    return np.array(kept, dtype=np.int64)


def axis_angle_to_matrix(r: np.ndarray) -> np.ndarray:
    t = np.linalg.norm(r)
    if t < 1e-12:
        return np.eye(3)
    k = r / t
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K
