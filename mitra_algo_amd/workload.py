"""Device-resident workloads for the voting core (used by bench.py and the
GPU tests).  A workload owns one libmvs session and torch device tensors
holding the mesh, its topology, the features and the selected samples, so a
timed pass starts with every input already in HBM.

Selection follows the reference's select (src/cpu/selection.cpp:15-29) with
the sampling replaced as the workload specifies: `samples` uniform vertex ids
without replacement (seeded), optionally the magnitude drop of
meshgen.magnitude_keep, then the reference's anisotropy prune (.85).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from . import capi, meshgen


@dataclass
class PairVoteResult:
    P: int
    R: int
    a: torch.Tensor
    b: torch.Tensor
    q4: torch.Tensor
    t3: torch.Tensor
    rp: torch.Tensor
    n3: torch.Tensor
    d1: torch.Tensor
    fp: torch.Tensor

    @property
    def V(self) -> int:
        return 2 * self.P + self.R


class Workload:
    def __init__(self, mesh: meshgen.HeightField, samples: int, threshold: float,
                 magnitude_drop: bool, sample_seed: int = 3, smoothing: int = 3,
                 device: str = "cuda:0"):
        self.mesh = mesh
        self.threshold = threshold
        self.dev = torch.device(device)
        self.sess = capi.Session(self.dev.index or 0)
        self.stream = torch.cuda.current_stream(self.dev)
        self.sess.set_stream(self.stream.cuda_stream)
        v = mesh.vertices
        self.topo = capi.prepare(v, mesh.faces)
        self.box = float(self.topo["box_size"])
        self.center = np.array(self.topo["center"])
        n = len(v)
        self.n = n
        t = lambda a, dt=None: torch.as_tensor(np.ascontiguousarray(a)).to(self.dev)
        self.x = t(v[:, 0])
        self.y = t(v[:, 1])
        self.z = t(v[:, 2])
        self.normals = t(self.topo["normals"])
        self.face_normals = t(self.topo["face_normals"])
        self.nbr_off = t(self.topo["nbr_off"].view(np.int64))
        self.nbr_idx = t(self.topo["nbr_idx"])
        self.edges = t(self.topo["edges"])
        self.feat = torch.empty((11, n), dtype=torch.float64, device=self.dev)
        self.valid = torch.empty(n, dtype=torch.uint8, device=self.dev)
        self.smoothing = smoothing
        self.curvature()
        torch.cuda.synchronize(self.dev)
        feat = self.feat.cpu().numpy()
        valid = self.valid.cpu().numpy()
        e = self.topo["edges"]
        self.mean_edge = float(np.mean(np.linalg.norm(v[e[:, 1]] - v[e[:, 0]], axis=1)))
        s = meshgen.uniform_samples(samples, n, sample_seed)
        if magnitude_drop:
            s = meshgen.magnitude_keep(s, feat[9], feat[10], self.mean_edge)
        self.selected_np = self.sess.prune(feat, valid, s, 0.85)
        self.selected = torch.as_tensor(self.selected_np).to(self.dev)
        self.S = len(self.selected_np)
        self.feat_np = feat
        self.valid_np = valid

    def curvature(self):
        nf = self.face_normals.shape[1]
        self.sess.dev_curvature(self.n, self.x, self.y, self.z, self.normals, nf,
                                self.face_normals, self.nbr_off, self.nbr_idx,
                                self.edges.shape[0], self.edges, self.smoothing, self.feat,
                                self.valid)

    def pair_votes(self, precision: int, max_pairs: int = 2**31 - 1) -> PairVoteResult:
        s = self.sess
        P = s.dev_pair_count(self.n, self.x, self.y, self.z, self.feat, self.S, self.selected,
                             self.box, self.center, self.threshold, max_pairs, precision)
        a = torch.empty(max(P, 1), dtype=torch.int32, device=self.dev)
        b = torch.empty(max(P, 1), dtype=torch.int32, device=self.dev)
        s.dev_pair_fill(a, b, True)
        R = s.dev_vote_count(P, a, b, precision)
        dt = torch.float32 if precision == capi.FP32_GUARDED else torch.float64
        q4 = torch.empty((4, max(2 * P, 1)), dtype=dt, device=self.dev)
        t3 = torch.empty((3, max(2 * P, 1)), dtype=dt, device=self.dev)
        rp = torch.empty(max(2 * P, 1), dtype=torch.int32, device=self.dev)
        n3 = torch.empty((3, max(R, 1)), dtype=dt, device=self.dev)
        d1 = torch.empty(max(R, 1), dtype=dt, device=self.dev)
        fp = torch.empty(max(R, 1), dtype=torch.int32, device=self.dev)
        s.dev_vote_emit(P, a, b, self.box, self.center, precision, q4, t3, rp, n3, d1, fp)
        return PairVoteResult(P, R, a, b, q4, t3, rp, n3, d1, fp)


    # ---------------------------------------------------------- fused step
    def pairvote(self, precision: int, max_pairs: int = 2**31 - 1,
                 max_votes: int = 2**31 - 1) -> capi.PairVoteResult:
        return self.sess.dev_pairvote(self.n, self.x, self.y, self.z, self.feat, self.S,
                                      self.selected, self.box, self.center, self.threshold,
                                      max_pairs, max_votes, precision)

    def meanshift(self, seeds: int = 20000, iterations: int = 30, h: float = 0.05,
                  cutoff: float = 3.0, seed_value: int = 7,
                  want_stats: bool = False) -> capi.MeanShiftResult:
        return self.sess.dev_meanshift_votes(seeds, seed_value, iterations, h, cutoff,
                                             want_stats)

    def step(self, precision: int, seeds: int = 20000, iterations: int = 30):
        """One pass of the voting core: pairs + votes, then batched mean-shift."""
        pv = self.pairvote(precision)
        ms = self.meanshift(seeds, iterations)
        return pv, ms

    def capture(self, precision: int, seeds: int = 20000, iterations: int = 30,
                h: float = 0.05, cutoff: float = 3.0, seed_value: int = 7,
                max_pairs: int = 2**31 - 1, max_votes: int = 2**31 - 1):
        """Captures step() as a HIP graph (see mvs_step_graph_capture); replay()
        reruns it over the same resident inputs."""
        return self.sess.step_graph_capture(
            self.n, self.x, self.y, self.z, self.feat, self.S, self.selected, self.box,
            self.center, self.threshold, max_pairs, max_votes, precision, seeds, seed_value,
            iterations, h, cutoff)

    def replay(self):
        self.sess.graph_launch()


def end_to_end(sess, soa, samples: np.ndarray, threshold: float, precision: int,
               seeds: int = 20000, iterations: int = 30, h: float = 0.05, cutoff: float = 3.0,
               seed_value: int = 7, smoothing: int = 3, stage_sync: bool = False):
    """One pass from the host mesh to modes on the device: prepare on the
    device (soa: the reference's host SoA mesh, x, y, z, f0, f1, f2),
    curvature, the prune (.85) of the host sample ids, pairs + votes, batched
    mean-shift.  Returns (prepared, S, pair-vote result, mean-shift result,
    stage seconds); with stage_sync each stage is timed to completion, else
    only the total is."""
    dev = torch.device("cuda", torch.cuda.current_device())
    times = {}
    t = [time.perf_counter()]

    def mark(name):
        if stage_sync:
            torch.cuda.synchronize(dev)
            now = time.perf_counter()
            times[name] = now - t[0]
            t[0] = now

    t0 = t[0]
    d = sess.dev_prepare_soa(*soa, device=dev)
    mark("prepare")
    n, nf = d["nv"], d["nf"]
    feat = torch.empty((11, n), dtype=torch.float64, device=dev)
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    xyz = d["xyz"]
    sess.dev_curvature(n, xyz[0], xyz[1], xyz[2], d["normals"], nf, d["face_normals"],
                       d["nbr_off"], d["nbr_idx"], d["edge_count"], d["edges"], smoothing, feat,
                       valid)
    mark("curvature")
    sample = torch.from_numpy(np.ascontiguousarray(samples, dtype=np.int32)).to(dev)
    sel = torch.empty(max(len(samples), 1), dtype=torch.int32, device=dev)
    S = sess.dev_prune(n, feat, valid, sample, 0.85, sel)
    mark("prune")
    pv = sess.dev_pairvote(n, xyz[0], xyz[1], xyz[2], feat, S, sel, d["box_size"], d["center"],
                           threshold, 2**31 - 1, 2**31 - 1, precision)
    mark("pairvote")
    ms = sess.dev_meanshift_votes(seeds, seed_value, iterations, h, cutoff, False)
    torch.cuda.synchronize(dev)
    mark("meanshift")
    times["total"] = time.perf_counter() - t0
    return d, S, pv, ms, times


def cfg1(device: str = "cuda:0") -> Workload:
    """96x96 height field, 2,000 uniform samples, magnitude drop, t = .05."""
    return Workload(meshgen.cfg1_mesh(), 2000, 0.05, True, device=device)


def cfg2(device: str = "cuda:0", n: int = 1024) -> Workload:
    """1024x1024 height field (2.09M triangles), 100,000 uniform samples, t = .02."""
    return Workload(meshgen.cfg2_mesh(n), 100000, 0.02, False, device=device)


def copy_motions(mesh: meshgen.HeightField, count: int = 20) -> dict:
    """Cluster records for verifying a height field: `count - 4` rigid
    motions carrying one motif copy onto another of the same handedness
    (copy i onto copy (37 i + 11) mod n, in copy order, skipping pairs of
    unlike handedness), then reflections through the mesh centre about the
    planes x = 0, y = 0, x = y and x = -y.  A copy at centre C_i turned by
    theta_i maps onto copy j by the turn R = Rz(theta_j - theta_i) about z and
    the translation C_j - R C_i."""
    copies = mesh.copies
    n = len(copies)
    refl, rot, tr, pn, po = [], [], [], [], []
    for i in range(n):
        if len(refl) == count - 4:
            break
        j = (37 * i + 11) % n
        a, b = copies[i], copies[j]
        if j == i or a.mirror != b.mirror:
            continue
        t = b.theta - a.theta
        c, s_ = np.cos(t), np.sin(t)
        R = np.array([[c, -s_, 0.0], [s_, c, 0.0], [0.0, 0.0, 1.0]])
        refl.append(0)
        rot.append(R.ravel())
        tr.append(np.array([b.cx, b.cy, 0.0]) - R @ np.array([a.cx, a.cy, 0.0]))
        pn.append(np.zeros(3))
        po.append(0.0)
    for nrm in ([1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.7071067811865476, -0.7071067811865476, 0.0],
                [0.7071067811865476, 0.7071067811865476, 0.0]):
        u = np.array(nrm)
        refl.append(1)
        rot.append((np.eye(3) - 2 * np.outer(u, u)).ravel())
        tr.append(np.zeros(3))
        pn.append(u)
        po.append(0.0)
    return dict(reflection=np.array(refl, np.uint8), rotation=np.array(rot),
                translation=np.array(tr), plane_normal=np.array(pn),
                plane_offset=np.array(po))


class Batch:
    """A batch of height-field meshes (meshgen.batch_spec) resident on one GPU:
    concatenated vertices and features, per-mesh sample offsets, boxes and
    centres, ready for the batch pair + vote and mean-shift entry points.
    Each mesh also keeps its own feature array for one-mesh reference runs."""

    def __init__(self, count: int = 512, seed: int = 4, threshold: float = 0.05,
                 min_side: int = 32, max_side: int = 512, smoothing: int = 3,
                 device: str = "cuda:0", keep_single: bool = False):
        self.dev = torch.device(device)
        self.threshold = threshold
        self.sess = capi.Session(self.dev.index or 0)
        self.stream = torch.cuda.current_stream(self.dev)
        self.sess.set_stream(self.stream.cuda_stream)
        xs, ys, zs, feats, sels = [], [], [], [], []
        self.vbase, self.soff, self.boxes, self.centers = [0], [0], [], []
        self.single = []
        for m in range(count):
            spec = meshgen.batch_spec(m, seed, min_side, max_side)
            mesh = meshgen.batch_mesh(spec)
            v = mesh.vertices
            topo = capi.prepare(v, mesh.faces)
            t = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(self.dev)
            x, y, z = t(v[:, 0]), t(v[:, 1]), t(v[:, 2])
            n = len(v)
            feat = torch.empty((11, n), dtype=torch.float64, device=self.dev)
            valid = torch.empty(n, dtype=torch.uint8, device=self.dev)
            self.sess.dev_curvature(n, x, y, z, t(topo["normals"]), topo["face_normals"].shape[1],
                                    t(topo["face_normals"]), t(topo["nbr_off"].view(np.int64)),
                                    t(topo["nbr_idx"]), topo["edges"].shape[0], t(topo["edges"]),
                                    smoothing, feat, valid)
            fnp = feat.cpu().numpy()
            sel = self.sess.prune(fnp, valid.cpu().numpy(), meshgen.batch_samples(n, m, seed),
                                  0.85)
            box, center = float(topo["box_size"]), np.array(topo["center"])
            xs.append(x), ys.append(y), zs.append(z), feats.append(feat)
            sels.append(torch.as_tensor(sel.astype(np.int32) + self.vbase[-1]))
            if keep_single:
                self.single.append(dict(n=n, x=x, y=y, z=z, feat=feat, S=len(sel),
                                        sel=torch.as_tensor(sel.astype(np.int32)).to(self.dev),
                                        box=box, center=center, side=spec.side))
            self.vbase.append(self.vbase[-1] + n)
            self.soff.append(self.soff[-1] + len(sel))
            self.boxes.append(box)
            self.centers.append(center)
        self.count = count
        self.n = self.vbase[-1]
        self.x = torch.cat(xs)
        self.y = torch.cat(ys)
        self.z = torch.cat(zs)
        self.feat = torch.cat(feats, dim=1).contiguous()
        self.selected = torch.cat(sels).to(self.dev)
        self.S = self.soff[-1]
        self.soff = np.array(self.soff, dtype=np.uint64)
        self.boxes = np.array(self.boxes)
        self.centers = np.array(self.centers)
        torch.cuda.synchronize(self.dev)

    def pairvote(self, precision: int, max_pairs: int = 2**31 - 1,
                 max_votes: int = 2**31 - 1) -> capi.BatchPairVoteResult:
        return self.sess.dev_batch_pairvote(self.n, self.x, self.y, self.z, self.feat, self.soff,
                                            self.selected, self.boxes, self.centers,
                                            self.threshold, max_pairs, max_votes, precision)

    def meanshift(self, seeds_per_sample: float = 0.2, iterations: int = 20, h: float = 0.05,
                  cutoff: float = 3.0, seed_value: int = 7,
                  want_stats: bool = False) -> capi.BatchMeanShiftResult:
        return self.sess.dev_batch_meanshift(seeds_per_sample, seed_value, iterations, h, cutoff,
                                             want_stats)

    def step(self, precision: int, seeds_per_sample: float = 0.2, iterations: int = 20):
        """Pairs + votes of every mesh, then mean-shift of every mesh."""
        pv = self.pairvote(precision)
        ms = self.meanshift(seeds_per_sample, iterations)
        return pv, ms

    def capture(self, precision: int, seeds_per_sample: float = 0.2, iterations: int = 20,
                h: float = 0.05, cutoff: float = 3.0, seed_value: int = 7):
        """Captures one batch step as a HIP graph (see mvs_batch_graph_capture)."""
        return self.sess.batch_graph_capture(
            self.n, self.x, self.y, self.z, self.feat, self.soff, self.selected, self.boxes,
            self.centers, self.threshold, 2**31 - 1, 2**31 - 1, precision, seeds_per_sample,
            seed_value, iterations, h, cutoff)

    def replay(self):
        self.sess.batch_graph_launch()

    def autotune(self, precision: int, seeds_per_sample: float = 0.2, iterations: int = 20,
                 h: float = 0.05, cutoff: float = 3.0, seed_value: int = 7):
        """Picks the launch table by timing graph replays (mvs_batch_autotune)."""
        return self.sess.batch_autotune(
            self.n, self.x, self.y, self.z, self.feat, self.soff, self.selected, self.boxes,
            self.centers, self.threshold, 2**31 - 1, 2**31 - 1, precision, seeds_per_sample,
            seed_value, iterations, h, cutoff)


def converged_mode_errors(sess, D: int, p64, p32, seed_idx, h: float = 0.05,
                          cutoff: float = 3.0, iterations: int = 2000) -> dict:
    """FP32 against FP64 mean-shift modes from the same seed positions.

    Both arithmetics run every seed to its fixed point (a seed stops once an
    update leaves it unchanged; `iterations` caps the updates), over their
    own points (p64: FP64 points, p32: the FP32 path's points, D x n).  The
    FP64 endpoints are merged into modes by the reference's rule (a seed joins
    the first mode within h/2 in the mean-shift coordinates, else starts a
    new one; src/cpu/clustering.cpp:64-70); a mode's position is the mean of
    its members' endpoints, in each arithmetic over the same members.  Errors
    per mode: length units (translation about the centre, or mirror offset)
    and radians (angle of R32^T R64, or between mirror normals).  A seed
    counts as converged when one more update moves it by at most 1e-12
    (FP64) or 8 units in the last place of its largest coordinate (FP32)."""
    import torch
    from scipy.spatial.transform import Rotation
    n = p64.shape[1]
    s64 = p64[:, seed_idx].contiguous()
    s32 = s64.float().contiguous()
    ns = s64.shape[1]

    def run(prec, pts, seeds, it):
        out = torch.empty_like(seeds)
        w = torch.empty(ns, dtype=seeds.dtype, device=seeds.device)
        c, i, u = sess.dev_ms_batched(prec, D, n, pts, ns, seeds, it, h, cutoff, out, w,
                                      stats="updates")
        torch.cuda.synchronize()
        return out.double().cpu().numpy(), u

    x64, u64 = run(capi.FP64, p64, s64, iterations)
    y64, _ = run(capi.FP64, p64, s64, iterations + 1)
    x32, u32 = run(capi.FP32_GUARDED, p32, s32, iterations)
    y32, _ = run(capi.FP32_GUARDED, p32, s32, iterations + 1)
    conv64 = np.max(np.abs(y64 - x64), axis=0) <= 1e-12
    ulp32 = np.spacing(np.max(np.abs(x32), axis=0).astype(np.float32)).astype(np.float64)
    conv32 = np.max(np.abs(y32 - x32), axis=0) <= 8 * ulp32
    both = conv64 & conv32
    centres, members = [], []
    for s in np.flatnonzero(both):
        for k, c in enumerate(centres):
            if np.sum((x64[:, s] - c) ** 2) < (h / 2) ** 2:
                members[k].append(s)
                break
        else:
            centres.append(x64[:, s])
            members.append([s])
    lin, ang, size = [], [], []
    for mem in members:
        a = x64[:, mem].mean(axis=1)
        b = x32[:, mem].mean(axis=1)
        if D == 6:
            lin.append(float(np.linalg.norm(b[0:3] - a[0:3])))
            ang.append(float((Rotation.from_rotvec(b[3:6]).inv() *
                              Rotation.from_rotvec(a[3:6])).magnitude()))
        else:
            lin.append(abs(float(b[0] - a[0])))
            ua, ub = a[1:4] / np.linalg.norm(a[1:4]), b[1:4] / np.linalg.norm(b[1:4])
            ang.append(float(np.arccos(np.clip(np.dot(ua, ub), -1.0, 1.0))))
        size.append(len(mem))
    lin, ang = np.array(lin), np.array(ang)
    return {"seeds": int(ns), "converged_fp64": int(conv64.sum()),
            "converged_fp32": int(conv32.sum()), "modes": len(members),
            "updates_fp64": int(u64), "updates_fp32": int(u32),
            "units_max": float(lin.max(initial=0)), "units_median": float(np.median(lin))
            if len(lin) else 0.0, "rad_max": float(ang.max(initial=0)),
            "rad_median": float(np.median(ang)) if len(ang) else 0.0,
            "modes_over_1e-6_units": int(np.sum(lin > 1e-6)),
            "modes_over_1e-5_rad": int(np.sum(ang > 1e-5)),
            "largest_mode_seeds": int(max(size, default=0))}
