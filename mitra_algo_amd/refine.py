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
        # each point's face from the previous query (-1: none yet): the next
        # query starts from it
        self.face = torch.full((self.pts.shape[0], self.pts.shape[2]), -1, dtype=torch.int32,
                               device=self.pts.device)

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
        """Closest mesh points and normals of T(p) (one launch transforms and
        queries every mode's points); returns the squared distances."""
        theta = theta.detach().float().contiguous()
        d2 = torch.empty(self.modes, self.npts, dtype=torch.float32, device=self.pts.device)
        self.sess.dev_closest_modes(self.kind, self.modes, self.npts, theta, self.base,
                                    self.pts, False, self.q, self.nrm, d2, self.face)
        return d2

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
    """Damped Gauss-Newton (Levenberg-Marquardt) steps of one problem; see
    refine_all.  Returns the per-step energies (steps + 1 rows)."""
    return refine_all([problem], steps, damping)[0]


def _flat_views(buf: torch.Tensor, shapes):
    """Consecutive views of a flat tensor with the given shapes."""
    views, o = [], 0
    for shape in shapes:
        n = int(np.prod(shape))
        views.append(buf[o:o + n].view(*shape))
        o += n
    return views


def refine_all(problems, steps: int = 10, damping: float = 1e-3):
    """Damped Gauss-Newton (Levenberg-Marquardt) steps; closest points are
    recomputed at the start of every step.  A step is kept when it lowers the
    energy (with the step's closest points), else the damping grows.  Rigid
    twists are folded into the base after each accepted step.  Energies,
    gradients and Gauss-Newton matrices come from the device; the 6x6 / 4x4
    solves and parameter updates run on the host in FP64.  The problems go
    in lock step: every launch of a phase is queued before one copy brings
    all their results to the host, and one copy takes all trial parameters
    (or all updated parameters) to the device, so a step costs two
    synchronisations whatever the number of problems.  Returns, per problem,
    the per-step energies (steps + 1 rows: before the first step and after
    each, with fresh closest points)."""
    problems = list(problems)
    if not problems:
        return []
    dev = problems[0].pts.device
    Ms = [p.modes for p in problems]
    Ps = [NPARAM[p.kind] for p in problems]
    lam = [np.full(M, damping) for M in Ms]
    theta = [p.theta.double().cpu().numpy() for p in problems]
    base = [p.base.double().cpu().numpy() if p.base is not None else None for p in problems]
    # device results of a phase, all problems in one buffer: E, g, H each
    shapes = [s for M, P in zip(Ms, Ps) for s in ((M,), (M, P), (M, P, P))]
    out = torch.empty(sum(int(np.prod(s)) for s in shapes), dtype=torch.float32, device=dev)
    outv = _flat_views(out, shapes)
    history = [[] for _ in problems]

    def launch(k, th, want_hessian):
        p = problems[k]
        E, g, H = outv[3 * k:3 * k + 3]
        p.sess.dev_align_energy(p.kind, p.modes, p.npts, th, p.base, p.pts, p.q, p.nrm, E, g,
                                H if want_hessian else None)

    def upload(arrays):
        """FP64 host arrays -> FP32 device tensors of the same shapes, one copy."""
        flat = torch.from_numpy(np.concatenate([a.ravel() for a in arrays])).float().to(dev)
        return _flat_views(flat, [a.shape for a in arrays])

    for it in range(steps + 1):
        for k, p in enumerate(problems):
            p.update_closest(p.theta)
            launch(k, p.theta.detach().float().contiguous(), True)
        host = _flat_views(out.cpu().double(), shapes)
        for k in range(len(problems)):
            history[k].append(host[3 * k].numpy().copy())
        if it == steps:
            break
        trials = []
        for k in range(len(problems)):
            g, H = host[3 * k + 1].numpy(), host[3 * k + 2].numpy()
            diag = np.diagonal(H, axis1=1, axis2=2)
            A = H + (lam[k][:, None] * (diag + 1e-12 * (1 + diag.sum(1, keepdims=True))))[
                :, :, None] * np.eye(H.shape[1])[None]
            trials.append(theta[k] - np.linalg.solve(A, g[:, :, None])[:, :, 0])
        for k, th in enumerate(upload(trials)):
            launch(k, th, False)
        host1 = _flat_views(out.cpu().double(), shapes)
        updated = []
        for k, p in enumerate(problems):
            E0, E1, trial = history[k][-1], host1[3 * k].numpy(), trials[k]
            ok = (E1 < E0) & np.isfinite(E1)
            lam[k] = np.where(ok, lam[k] / 3, lam[k] * 4)
            if p.kind == RIGID:
                # fold the accepted twist into the base: the twist restarts at zero
                for j in np.nonzero(ok)[0]:
                    R = axis_angle_to_matrix(trial[j, :3])
                    R0 = base[k][j, :9].reshape(3, 3)
                    base[k][j] = np.concatenate([(R @ R0).ravel(),
                                                 R @ base[k][j, 9:] + trial[j, 3:6]])
                    theta[k][j] = 0
                updated += [base[k], theta[k]]
            else:
                for j in np.nonzero(ok)[0]:
                    theta[k][j] = np.concatenate([trial[j, :3] / np.linalg.norm(trial[j, :3]),
                                                  trial[j, 3:4]])
                updated += [theta[k]]
        dv = iter(upload(updated))
        for p in problems:
            if p.kind == RIGID:
                p.base = next(dv)
            p.theta = next(dv)
    return [np.stack(h) for h in history]


def surface_samples(vertices: np.ndarray, faces: np.ndarray, count: int,
                    seed: int):
    """`count` points uniform on the surface (area-weighted faces, uniform
    barycentric coordinates), PCG64(seed): ((count, 3) FP64 points, their
    unit face normals)."""
    g = np.random.default_rng(seed)
    a, b, c = (vertices[faces[:, k]] for k in range(3))
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    f = g.choice(len(faces), size=count, p=area / area.sum())
    r1, r2 = g.uniform(size=count), g.uniform(size=count)
    s = np.sqrt(r1)
    pts = ((1 - s)[:, None] * a[f] + (s * (1 - r2))[:, None] * b[f] +
           (s * r2)[:, None] * c[f])
    return pts, cr[f] / (2 * area[f])[:, None]


def morton_codes(p: np.ndarray, bits: int = 10) -> np.ndarray:
    """3-D Morton (Z-order) codes of points, `bits` per axis over their box."""
    lo, hi = p.min(0), p.max(0)
    q = ((p - lo) / np.where(hi > lo, hi - lo, 1) * ((1 << bits) - 1)).astype(np.uint64)
    code = np.zeros(len(p), dtype=np.uint64)
    for b in range(bits):
        for a in range(3):
            code |= ((q[:, a] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return code


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


# ---------------------------------------------------------------- modes + patches
VERIFY_THRESHOLD = 0.025  # the reference's default (include/symmetry/options.h:7)


@dataclass
class Mode:
    kind: int
    weight: float
    coords: np.ndarray   # mean-shift point: rigid (t_c, axis-angle), mirror (d_c, n)
    support: int = 0     # candidate surface points passing the patch rule


def top_modes(wl, count: int = 64, seeds: int = 20000, iterations: int = 30,
              h: float = 0.05, cutoff: float = 3.0, seed_value: int = 7) -> list:
    """Distinct FP32 modes of a workload, heaviest first: mean-shift over the
    votes of the FP32-guarded pair pass, modes within h/2 of a heavier one of
    the same kind dropped (the reference's merge radius, squared h^2/4,
    src/cpu/clustering.cpp), then the `count` heaviest over both kinds."""
    wl.pairvote(capi.FP32_GUARDED)
    ms = wl.meanshift(seeds, iterations, h, cutoff, seed_value)
    s = wl.sess
    out = []
    for kind, k, modes, weight, dim in ((RIGID, ms.rigid_seeds, ms.rigid_modes,
                                         ms.rigid_weight, 6),
                                        (MIRROR, ms.refl_seeds, ms.refl_modes,
                                         ms.refl_weight, 4)):
        if k == 0:
            continue
        pts = s.fetch(modes, (dim, k), torch.float32, k).double().cpu().numpy()
        w = s.fetch(weight, (k,), torch.float32).double().cpu().numpy()
        for i in merge_modes(pts, w, h / 2):
            out.append(Mode(kind, float(w[i]), pts[:, i].copy()))
    out.sort(key=lambda m: -m.weight)
    return out[:count]


def mode_transform(m: Mode):
    """(3x3 linear part, translation) of a mode in centre-relative coordinates."""
    if m.kind == RIGID:
        return axis_angle_to_matrix(m.coords[3:6]), m.coords[0:3].copy()
    u = m.coords[1:4] / np.linalg.norm(m.coords[1:4])
    return np.eye(3) - 2 * np.outer(u, u), -2 * m.coords[0] * u


class Patches:
    """Candidate surface points of one mesh (centre-relative FP32 on the
    device, with their face normals) and the session's BVH of that mesh."""

    def __init__(self, sess: capi.Session, vertices: np.ndarray, faces: np.ndarray,
                 center, box: float, candidates: int = 262144, seed: int = 11,
                 device: str = "cuda:0"):
        self.sess = sess
        self.box = box
        self.tau = VERIFY_THRESHOLD * box
        sess.bvh_build(vertices, faces, center)
        p, n = surface_samples(vertices, faces, candidates, seed)
        p = p - np.asarray(center, dtype=np.float64)
        # Morton order: neighbouring candidates are neighbours on the surface,
        # so a wavefront's closest-point queries walk the same BVH nodes
        order = np.argsort(morton_codes(p), kind="stable")
        p, n = p[order], n[order]
        self.pts = torch.as_tensor(np.ascontiguousarray(p.T), dtype=torch.float32).to(device)
        self.nrm = torch.as_tensor(np.ascontiguousarray(n.T), dtype=torch.float32).to(device)
        self.count = candidates

    def select(self, modes: list, npts: int, chunk: int = 16) -> list:
        """Patch point indices per mode: the candidates p whose image T(p)
        passes the reference's verification rule for the A side
        (src/cpu/verification.cpp:58-80): |T(p) - p| >= 2 tau, for mirrors
        u.p + d < -tau, the closest surface point within tau and the mapped
        normal within 60 degrees of that face's normal (tau = 0.025 box).
        The patch takes `npts` of the passing candidates evenly spread over
        their (Morton) order, or all of them repeated cyclically when fewer
        pass; None when none does."""
        dev = self.pts.device
        out = []
        for c0 in range(0, len(modes), chunk):
            ms = modes[c0:c0 + chunk]
            M, C = len(ms), self.count
            L = torch.tensor(np.stack([mode_transform(m)[0] for m in ms]),
                             dtype=torch.float32, device=dev)
            t = torch.tensor(np.stack([mode_transform(m)[1] for m in ms]),
                             dtype=torch.float32, device=dev)
            # each mode as the kernel's rigid transform (linear part, offset)
            base = torch.cat([L.reshape(M, 9), t], 1).contiguous()
            zero = torch.zeros(M, 6, dtype=torch.float32, device=dev)
            q = torch.empty(M, 3, C, dtype=torch.float32, device=dev)
            qn = torch.empty_like(q)
            d2 = torch.empty(M, C, dtype=torch.float32, device=dev)
            self.sess.dev_closest_modes(RIGID, M, C, zero, base, self.pts, True, q, qn, d2)
            tp = (L[:, :, :, None] * self.pts[None, None, :, :]).sum(2) + t[:, :, None]
            tn = (L[:, :, :, None] * self.nrm[None, None, :, :]).sum(2)
            move = torch.linalg.norm(tp - self.pts[None], dim=1)
            ok = (move >= 2 * self.tau) & (d2 <= self.tau * self.tau) & \
                ((tn * qn).sum(1) >= 0.5)
            for j, m in enumerate(ms):
                if m.kind == MIRROR:
                    u = m.coords[1:4] / np.linalg.norm(m.coords[1:4])
                    side = (torch.tensor(u, dtype=torch.float32, device=dev)[:, None]
                            * self.pts).sum(0) + float(m.coords[0])
                    ok[j] &= side < -self.tau
                idx = torch.nonzero(ok[j]).flatten()
                m.support = int(idx.numel())
                if m.support == 0:
                    out.append(None)
                elif m.support >= npts:
                    k = torch.arange(npts, device=dev, dtype=torch.int64)
                    out.append(idx[(k * m.support) // npts])
                else:
                    out.append(idx.repeat(-(-npts // m.support))[:npts])
        return out

    def problem(self, modes: list, idx: list, kind: int) -> tuple:
        """The Problem over the modes of `kind` with a patch; returns it with
        the positions (in `modes`) of the modes it holds."""
        pick = [i for i, m in enumerate(modes) if m.kind == kind and idx[i] is not None]
        if not pick:
            return None, pick
        dev = self.pts.device
        pts = torch.stack([self.pts[:, idx[i]] for i in pick]).contiguous()
        if kind == RIGID:
            base = torch.tensor(np.stack([np.concatenate([mode_transform(modes[i])[0].ravel(),
                                                          modes[i].coords[0:3]])
                                          for i in pick]), dtype=torch.float32, device=dev)
            theta = torch.zeros(len(pick), 6, dtype=torch.float32, device=dev)
        else:
            base = None
            theta = torch.tensor(np.stack([np.concatenate([modes[i].coords[1:4],
                                                           modes[i].coords[0:1]])
                                           for i in pick]), dtype=torch.float32, device=dev)
        return Problem(self.sess, kind, pts, base, theta.contiguous()), pick


def rms_residual(energy: np.ndarray, npts: int) -> np.ndarray:
    """Root-mean-square point-to-plane residual from per-mode energies."""
    return np.sqrt(np.maximum(energy, 0) / npts)
