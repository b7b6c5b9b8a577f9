"""ctypes binding of libmvs (include/mvs.h).

The library is built in-tree by mitra_algo_amd/build.py.  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmvs.so")

OK, EINVAL, EBUDGET, EOVERFLOW, EHIP, ENODEV, ERUNTIME = 0, 1, 2, 3, 4, 5, 6
FP64, FP32_GUARDED = 0, 1
OWN_STREAM = 2**64 - 1  # MVS_OWN_STREAM: (void *)-1

EXPORTS = [
    "mvs_prepare", "mvs_mt64_seed", "mvs_mt64_next", "mvs_sample_vertices",
    "mvs_session_create", "mvs_session_destroy", "mvs_last_error", "mvs_session_set_stream",
    "mvs_session_sync", "mvs_session_enable_timing", "mvs_session_timing_filter",
    "mvs_dev_timing",
    "mvs_curvature", "mvs_prune", "mvs_pairing", "mvs_pairing_fetch", "mvs_transformations",
    "mvs_transformations_fetch", "mvs_clustering", "mvs_clustering_fetch",
    "mvs_dev_prepare", "mvs_dev_prune",
    "mvs_dev_curvature", "mvs_dev_pair_count", "mvs_dev_pair_fill", "mvs_dev_pair_stats",
    "mvs_dev_vote_count", "mvs_dev_vote_emit", "mvs_dev_ms_points", "mvs_dev_ms_batched",
    "mvs_dev_pairvote", "mvs_dev_meanshift_votes", "mvs_dev_copy", "mvs_dev_batch_pairvote",
    "mvs_dev_batch_meanshift", "mvs_batch_graph_capture", "mvs_batch_graph_launch",
    "mvs_batch_graph_check", "mvs_batch_autotune", "mvs_session_get_tuning",
    "mvs_session_set_tuning", "mvs_bvh_build", "mvs_dev_closest_points",
    "mvs_dev_closest_modes", "mvs_dev_align_energy",
    "mvs_verification", "mvs_verification_fetch", "mvs_verification_rows",
    "mvs_verification_stats", "mvs_verification_surface", "mvs_dev_verification",
    "mvs_step_graph_capture", "mvs_graph_launch", "mvs_graph_check", "mvs_thread_metrics",
    "mvs_thread_metrics_reset", "mvs_dev_ms_batched_noalloc",
]


class MemoryMetrics(C.Structure):
    _fields_ = [("live_bytes", C.c_uint64), ("peak_bytes", C.c_uint64),
                ("h2d_bytes", C.c_uint64), ("d2h_bytes", C.c_uint64)]


def thread_metrics() -> dict:
    """Device-memory and copy counters of this thread's libmvs calls."""
    m = MemoryMetrics()
    check(lib().mvs_thread_metrics(C.byref(m)))
    return {k: getattr(m, k) for k, _ in MemoryMetrics._fields_}


class MvsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class MvsInvalidArgument(MvsError, ValueError):
    """std::invalid_argument in the reference."""


class MvsBudgetError(MvsError):
    """std::runtime_error carrying "budget" in the reference."""


class MvsOverflowError(MvsError, OverflowError):
    """std::overflow_error in the reference."""


def _raise(code: int, msg: str):
    if code == EINVAL:
        raise MvsInvalidArgument(code, msg)
    if code == EBUDGET:
        raise MvsBudgetError(code, msg)
    if code == EOVERFLOW:
        raise MvsOverflowError(code, msg)
    raise MvsError(code, msg)


class MT64(C.Structure):
    """std::mt19937_64 state, libstdc++ layout."""
    _fields_ = [("x", C.c_uint64 * 312), ("p", C.c_uint64)]

    def copy(self) -> "MT64":
        o = MT64()
        C.memmove(C.byref(o), C.byref(self), C.sizeof(MT64))
        return o

    def __eq__(self, other) -> bool:
        return bytes(self) == bytes(other)


class Edge(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("lo", "hi", "a", "b", "face0", "face1")]


class Points(C.Structure):
    _fields_ = [("count", C.c_uint64), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p)]


class TopologyView(C.Structure):
    _fields_ = [("normals", Points), ("face_normals", Points), ("neighbor_offsets", C.c_void_p),
                ("neighbor_indices", C.c_void_p), ("edge_count", C.c_uint64),
                ("edges", C.c_void_p)]


class FeaturesView(C.Structure):
    _fields_ = [("count", C.c_uint64), ("lane", C.c_void_p * 11), ("valid", C.c_void_p)]


class FeaturesOut(C.Structure):
    _fields_ = [("lane", C.c_void_p * 11), ("valid", C.c_void_p)]


class Prepared(C.Structure):
    _fields_ = [
        ("face_count", C.c_int64),
        ("f0", C.c_void_p), ("f1", C.c_void_p), ("f2", C.c_void_p),
        ("nx", C.c_void_p), ("ny", C.c_void_p), ("nz", C.c_void_p),
        ("fnx", C.c_void_p), ("fny", C.c_void_p), ("fnz", C.c_void_p),
        ("weights", C.c_void_p),
        ("nbr_off", C.c_void_p), ("nbr_idx", C.c_void_p), ("nbr_nnz", C.c_int64),
        ("inc_off", C.c_void_p), ("inc_idx", C.c_void_p), ("inc_nnz", C.c_int64),
        ("edges", C.c_void_p), ("edge_count", C.c_int64),
        ("box_size", C.c_double), ("center", C.c_double * 3),
        ("rejected_faces", C.c_int32), ("nonmanifold_edges", C.c_int32),
        ("orientation_conflicts", C.c_int32),
    ]


class PairVoteResult(C.Structure):
    _fields_ = [("pairs", C.c_uint64), ("reflections", C.c_uint64), ("a", C.c_void_p),
                ("b", C.c_void_p), ("rigid_q4", C.c_void_p), ("rigid_t3", C.c_void_p),
                ("rigid_pair", C.c_void_p), ("refl_n3", C.c_void_p), ("refl_d", C.c_void_p),
                ("refl_pair", C.c_void_p), ("refl_stride", C.c_uint64), ("precision", C.c_int)]


class BatchPairVoteResult(C.Structure):
    _fields_ = [("all", PairVoteResult), ("meshes", C.c_uint32), ("pair_offsets", C.c_void_p),
                ("refl_offsets", C.c_void_p)]


class MeanShiftResult(C.Structure):
    _fields_ = [("rigid_seeds", C.c_uint64), ("refl_seeds", C.c_uint64),
                ("seed_votes", C.c_void_p), ("rigid_modes", C.c_void_p),
                ("rigid_weight", C.c_void_p), ("refl_modes", C.c_void_p),
                ("refl_weight", C.c_void_p), ("candidates", C.c_uint64),
                ("in_support", C.c_uint64), ("precision", C.c_int),
                ("updates", C.c_uint64), ("rigid_candidates", C.c_uint64),
                ("rigid_in_support", C.c_uint64), ("refl_candidates", C.c_uint64),
                ("refl_in_support", C.c_uint64)]


class TuneEntry(C.Structure):
    _fields_ = [("walk_quota", C.c_uint32), ("walk_blocks", C.c_uint32),
                ("wide_seeds", C.c_uint64), ("ms", C.c_double), ("seed_lanes", C.c_uint32),
                ("reserved", C.c_uint32)]

    def as_dict(self):
        return {"walk_quota": self.walk_quota, "walk_blocks": self.walk_blocks,
                "wide_seeds": self.wide_seeds, "seed_lanes": self.seed_lanes, "ms": self.ms}


class VerificationCounts(C.Structure):
    _fields_ = [("side_a", C.c_uint64), ("side_b", C.c_uint64), ("patches_a", C.c_uint64),
                ("patches_b", C.c_uint64)]


class VerificationSide(C.Structure):
    _fields_ = [("side_offsets", C.c_void_p), ("side", C.c_void_p),
                ("patch_groups", C.c_void_p), ("patch_offsets", C.c_void_p),
                ("patch_members", C.c_void_p)]


class BatchMeanShiftResult(C.Structure):
    _fields_ = [("all", MeanShiftResult), ("meshes", C.c_uint32),
                ("rigid_seed_offsets", C.c_void_p), ("refl_seed_offsets", C.c_void_p)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run mitra_algo_amd/build.py "
                              "(or __graft_entry__.build())")
        _torch_first()
        L = C.CDLL(LIB_PATH)
        L.mvs_last_error.restype = C.c_char_p
        L.mvs_mt64_next.restype = C.c_uint64
        L.mvs_mt64_next.argtypes = [C.POINTER(MT64)]
        L.mvs_mt64_seed.argtypes = [C.POINTER(MT64), C.c_uint64]
        L.mvs_mt64_seed.restype = None
        L.mvs_session_destroy.restype = None
        _lib = L
    return _lib


def check(rc: int):
    if rc != OK:
        _raise(rc, lib().mvs_last_error().decode())


def ptr(a) -> C.c_void_p:
    """Address of a numpy array or a torch tensor (device or host)."""
    if a is None:
        return C.c_void_p(0)
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    if isinstance(a, int):
        return C.c_void_p(a)
    return C.c_void_p(a.data_ptr())


def rng(seed: int) -> MT64:
    g = MT64()
    lib().mvs_mt64_seed(C.byref(g), C.c_uint64(seed))
    return g


def mt_next(g: MT64) -> int:
    return int(lib().mvs_mt64_next(C.byref(g)))


def prepare(vertices: np.ndarray, faces: np.ndarray) -> dict:
    """prepare() of the reference (host code in libmvs)."""
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    f = np.ascontiguousarray(faces, dtype=np.int32)
    nv, nf = len(v), len(f)
    x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
    f0, f1, f2 = (np.ascontiguousarray(f[:, k]) for k in range(3))
    cap = max(nf, 1)
    out = {k: np.zeros(cap, np.int32) for k in ("f0", "f1", "f2")}
    out.update({k: np.zeros(max(nv, 1)) for k in ("nx", "ny", "nz", "weights")})
    out.update({k: np.zeros(cap) for k in ("fnx", "fny", "fnz")})
    out["nbr_off"] = np.zeros(nv + 1, np.uint64)
    out["nbr_idx"] = np.zeros(6 * cap, np.int32)
    out["inc_off"] = np.zeros(nv + 1, np.uint64)
    out["inc_idx"] = np.zeros(3 * cap, np.int32)
    out["edges"] = np.zeros((3 * cap, 6), np.int32)
    P = Prepared()
    for k, a in out.items():
        setattr(P, k, ptr(a))
    check(lib().mvs_prepare(C.c_int64(nv), ptr(x), ptr(y), ptr(z), C.c_int64(nf), ptr(f0),
                            ptr(f1), ptr(f2), C.byref(P)))
    F = P.face_count
    return dict(
        faces=np.stack([out["f0"][:F], out["f1"][:F], out["f2"][:F]], 1),
        normals=np.stack([out["nx"][:nv], out["ny"][:nv], out["nz"][:nv]]),
        face_normals=np.stack([out["fnx"][:F], out["fny"][:F], out["fnz"][:F]]),
        weights=out["weights"][:nv].copy(), nbr_off=out["nbr_off"],
        nbr_idx=out["nbr_idx"][:P.nbr_nnz].copy(), inc_off=out["inc_off"],
        inc_idx=out["inc_idx"][:P.inc_nnz].copy(), edges=out["edges"][:P.edge_count].copy(),
        box_size=P.box_size, center=np.array(list(P.center)),
        rejected_faces=P.rejected_faces)


def sample_vertices(weights: np.ndarray, ratio: float, g: MT64) -> np.ndarray:
    w = np.ascontiguousarray(weights, dtype=np.float64)
    out = np.zeros(max(len(w), 1), np.int32)
    cnt = C.c_uint64()
    check(lib().mvs_sample_vertices(C.c_uint64(len(w)), ptr(w), C.c_double(ratio), C.byref(g),
                                    ptr(out), C.byref(cnt)))
    return out[:cnt.value].copy()


def _torch_first():
    """Loads and initialises torch's HIP runtime, when torch is installed and
    sees a GPU, before libmvs is loaded.  torch ships its own HIP runtime
    beside the ROCm one libmvs links; both drive the same device through one
    HSA runtime (so device pointers pass between them), and they work
    together when torch's is loaded and initialised first; in the other
    orders one of the two reports no GPU."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


class Session:
    """One libmvs session (device context + stream + scratch)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().mvs_session_create(C.c_int(device), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().mvs_session_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_stream(self, stream_handle: int):
        """Run on this hipStream_t (0: the null stream, torch's default)."""
        check(lib().mvs_session_set_stream(self.h, C.c_void_p(stream_handle)))

    def use_own_stream(self):
        """Run on the session's own non-blocking stream again."""
        check(lib().mvs_session_set_stream(self.h, C.c_void_p(OWN_STREAM)))

    def sync(self):
        check(lib().mvs_session_sync(self.h))

    def enable_timing(self, on: bool = True, only=None):
        """Per-launch HIP-event timing; `only` restricts it to these kernel names."""
        check(lib().mvs_session_enable_timing(self.h, C.c_int(1 if on else 0)))
        names = ",".join(only) if only else ""
        check(lib().mvs_session_timing_filter(self.h, names.encode()))

    def timing(self, kernel: str):
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().mvs_dev_timing(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ------------------------------------------------------ host layer
    def curvature(self, vertices: np.ndarray, topo: dict, smoothing: int):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        n = len(v)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        nrm = np.ascontiguousarray(topo["normals"], dtype=np.float64)
        fn = np.ascontiguousarray(topo["face_normals"], dtype=np.float64)
        off = np.ascontiguousarray(topo["nbr_off"], dtype=np.uint64)
        idx = np.ascontiguousarray(topo["nbr_idx"], dtype=np.int32)
        edges = np.ascontiguousarray(topo["edges"], dtype=np.int32)
        T = TopologyView(Points(n, ptr(nrm[0]), ptr(nrm[1]), ptr(nrm[2])),
                         Points(fn.shape[1], ptr(fn[0]), ptr(fn[1]), ptr(fn[2])), ptr(off),
                         ptr(idx), len(edges), ptr(edges))
        feat = np.zeros((11, n))
        valid = np.zeros(n, np.uint8)
        O = FeaturesOut()
        for k in range(11):
            O.lane[k] = feat[k].ctypes.data
        O.valid = valid.ctypes.data
        check(lib().mvs_curvature(self.h, C.byref(Points(n, ptr(x), ptr(y), ptr(z))),
                                  C.byref(T), C.c_int(smoothing), C.byref(O)))
        return feat, valid

    @staticmethod
    def _features(feat: np.ndarray, valid: np.ndarray | None):
        feat = np.ascontiguousarray(feat, dtype=np.float64)
        n = feat.shape[1]
        if valid is None:
            valid = np.ones(n, np.uint8)
        valid = np.ascontiguousarray(valid, dtype=np.uint8)
        F = FeaturesView()
        F.count = n
        for k in range(11):
            F.lane[k] = feat[k].ctypes.data
        F.valid = valid.ctypes.data
        return F, (feat, valid)

    def prune(self, feat, valid, sample, prune: float) -> np.ndarray:
        F, keep = self._features(feat, valid)
        s = np.ascontiguousarray(sample, dtype=np.int32)
        out = np.zeros(max(len(s), 1), np.int32)
        cnt = C.c_uint64()
        check(lib().mvs_prune(self.h, C.byref(F), C.c_uint64(len(s)), ptr(s), C.c_double(prune),
                              ptr(out), C.byref(cnt)))
        return out[:cnt.value].copy()

    def pairing(self, vertices, feat, selected, box: float, threshold: float,
                max_pairs: int = 200000, precision: int = FP64):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        F, keep = self._features(feat, None)
        s = np.ascontiguousarray(selected, dtype=np.int32)
        cnt = C.c_uint64()
        check(lib().mvs_pairing(self.h, C.byref(Points(len(v), ptr(x), ptr(y), ptr(z))),
                                C.byref(F), C.c_uint64(len(s)), ptr(s), C.c_double(box),
                                C.c_double(threshold), C.c_uint64(max_pairs), C.c_int(precision),
                                C.byref(cnt)))
        P = cnt.value
        a = np.zeros(max(P, 1), np.int32)
        b = np.zeros(max(P, 1), np.int32)
        if P:
            check(lib().mvs_pairing_fetch(self.h, ptr(a), ptr(b)))
        return a[:P], b[:P]

    def transformations(self, vertices, feat, pa, pb, box: float, center,
                        max_votes: int = 600000):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        F, keep = self._features(feat, None)
        pa = np.ascontiguousarray(pa, dtype=np.int32)
        pb = np.ascontiguousarray(pb, dtype=np.int32)
        ctr = np.ascontiguousarray(center, dtype=np.float64)
        cnt = C.c_uint64()
        check(lib().mvs_transformations(self.h, C.byref(Points(len(v), ptr(x), ptr(y), ptr(z))),
                                        C.byref(F), C.c_uint64(len(pa)), ptr(pa), ptr(pb),
                                        C.c_double(box), ptr(ctr), C.c_uint64(max_votes),
                                        C.byref(cnt)))
        V = cnt.value
        pair = np.zeros(max(V, 1), np.int32)
        kind = np.zeros(max(V, 1), np.uint8)
        lanes = np.zeros((28, max(V, 1)))
        if V:
            arr = (C.c_void_p * 28)(*[lanes[k].ctypes.data for k in range(28)])
            check(lib().mvs_transformations_fetch(self.h, ptr(pair), ptr(kind), arr))
        return pair[:V], kind[:V], np.ascontiguousarray(lanes[:, :V])

    def clustering(self, kind, desc, rotation, box: float, center, bandwidth: float,
                   max_iterations: int, max_clusters: int, g: MT64):
        kind = np.ascontiguousarray(kind, dtype=np.uint8)
        V = len(kind)
        desc = np.ascontiguousarray(desc, dtype=np.float64).reshape(12, V)
        rotation = np.ascontiguousarray(rotation, dtype=np.float64).reshape(9, V)
        dptr = (C.c_void_p * 12)(*[desc[k].ctypes.data for k in range(12)])
        rptr = (C.c_void_p * 9)(*[rotation[k].ctypes.data for k in range(9)])
        ctr = np.ascontiguousarray(center, dtype=np.float64)
        nc, nm, lim = C.c_uint64(), C.c_uint64(), C.c_int32()
        check(lib().mvs_clustering(self.h, C.c_uint64(V), ptr(kind), dptr, rptr, C.c_double(box),
                                   ptr(ctr), C.c_double(bandwidth), C.c_int(max_iterations),
                                   C.c_int(max_clusters), C.byref(g), C.byref(nc), C.byref(nm),
                                   C.byref(lim)))
        n, m = nc.value, nm.value
        refl = np.zeros(max(n, 1), np.uint8)
        rot = np.zeros((max(n, 1), 9))
        tr = np.zeros((max(n, 1), 3))
        pn = np.zeros((max(n, 1), 3))
        po = np.zeros(max(n, 1))
        off = np.zeros(n + 1, np.uint64)
        mem = np.zeros(max(m, 1), np.int32)
        check(lib().mvs_clustering_fetch(self.h, ptr(refl), ptr(rot), ptr(tr), ptr(pn), ptr(po),
                                         ptr(off), ptr(mem)))
        return dict(reflection=refl[:n], rotation=rot[:n], translation=tr[:n],
                    plane_normal=pn[:n], plane_offset=po[:n], member_off=off,
                    members=mem[:m], limited_iterations=lim.value)

    def verification(self, vertices, faces, topo: dict, clusters: dict,
                     verify_threshold: float, box: float) -> dict:
        """Backend::verification (src/cpu/verification.cpp:50-110).  topo holds
        normals (3, N), weights (N), nbr_off, nbr_idx; clusters holds
        reflection (C), rotation (C, 9), translation (C, 3), plane_normal
        (C, 3), plane_offset (C).  Returns the Verification fields: side_a /
        side_b as (offsets, indices), patches_a / patches_b as (groups,
        row offsets, members), coverage and rms."""
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        n = len(v)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        f0, f1, f2 = (np.ascontiguousarray(f[:, k]) for k in range(3))
        nrm = np.ascontiguousarray(topo["normals"], dtype=np.float64).reshape(3, n)
        w = np.ascontiguousarray(topo["weights"], dtype=np.float64)
        off = np.ascontiguousarray(topo["nbr_off"], dtype=np.uint64)
        idx = np.ascontiguousarray(topo["nbr_idx"], dtype=np.int32)
        refl = np.ascontiguousarray(clusters["reflection"], dtype=np.uint8)
        nc = len(refl)
        rot = np.ascontiguousarray(clusters["rotation"], dtype=np.float64).reshape(nc, 9)
        tr = np.ascontiguousarray(clusters["translation"], dtype=np.float64).reshape(nc, 3)
        pn = np.ascontiguousarray(clusters["plane_normal"], dtype=np.float64).reshape(nc, 3)
        po = np.ascontiguousarray(clusters["plane_offset"], dtype=np.float64).reshape(nc)
        cnt = VerificationCounts()
        check(lib().mvs_verification(
            self.h, C.byref(Points(n, ptr(x), ptr(y), ptr(z))), C.c_uint64(len(f)), ptr(f0),
            ptr(f1), ptr(f2), C.byref(Points(n, ptr(nrm[0]), ptr(nrm[1]), ptr(nrm[2]))), ptr(w),
            ptr(off), ptr(idx), C.c_uint64(nc), ptr(refl), ptr(rot), ptr(tr), ptr(pn), ptr(po),
            C.c_double(verify_threshold), C.c_double(box), C.byref(cnt)))
        return self._verification_fetch(cnt, nc)

    def _verification_fetch(self, cnt, nc: int) -> dict:
        out, sides = {}, []
        for name, total, patches in (("a", cnt.side_a, cnt.patches_a),
                                     ("b", cnt.side_b, cnt.patches_b)):
            # every element is written by the fetch
            so = np.empty(nc + 1, np.uint64)
            sv = np.empty(max(total, 1), np.int32)
            pg = np.empty(nc + 1, np.uint64)
            pof = np.empty(patches + 1, np.uint64)
            pm = np.empty(max(total, 1), np.int32)
            sides.append(VerificationSide(ptr(so), ptr(sv), ptr(pg), ptr(pof), ptr(pm)))
            out["side_" + name] = (so, sv[:total])
            out["patches_" + name] = (pg, pof, pm[:total])
        cov = np.zeros(max(nc, 1))
        rms = np.zeros(max(nc, 1))
        check(lib().mvs_verification_fetch(self.h, C.byref(sides[0]), C.byref(sides[1]),
                                           ptr(cov), ptr(rms)))
        out["coverage"], out["rms"] = cov[:nc], rms[:nc]
        return out

    def verification_surface(self, vertices, faces):
        """mvs_verification_surface: keep the mesh's surface tree on the session."""
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        f0, f1, f2 = (np.ascontiguousarray(f[:, k]) for k in range(3))
        check(lib().mvs_verification_surface(self.h, C.byref(Points(len(v), ptr(x), ptr(y), ptr(z))),
                                             C.c_uint64(len(f)), ptr(f0), ptr(f1), ptr(f2)))

    def dev_verification(self, nv: int, pts6, nbr_off, nbr_idx, weights, clusters: dict,
                         verify_threshold: float, box: float) -> dict:
        """mvs_dev_verification: pts6 (6, nv) float64, nbr_off, nbr_idx device
        tensors; weights a host array; results as verification()."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        refl = np.ascontiguousarray(clusters["reflection"], dtype=np.uint8)
        nc = len(refl)
        rot = np.ascontiguousarray(clusters["rotation"], dtype=np.float64).reshape(nc, 9)
        tr = np.ascontiguousarray(clusters["translation"], dtype=np.float64).reshape(nc, 3)
        pn = np.ascontiguousarray(clusters["plane_normal"], dtype=np.float64).reshape(nc, 3)
        po = np.ascontiguousarray(clusters["plane_offset"], dtype=np.float64).reshape(nc)
        cnt = VerificationCounts()
        check(lib().mvs_dev_verification(self.h, C.c_uint64(nv), ptr(pts6), ptr(nbr_off),
                                         ptr(nbr_idx), ptr(w), C.c_uint64(nc), ptr(refl), ptr(rot),
                                         ptr(tr), ptr(pn), ptr(po), C.c_double(verify_threshold),
                                         C.c_double(box), C.byref(cnt)))
        return self._verification_fetch(cnt, nc)

    def verification_stats(self) -> dict:
        """Triangle tests, tree nodes visited and FP64 operations of the
        surface searches of the last verification."""
        out = np.zeros(6, np.uint64)
        check(lib().mvs_verification_stats(self.h, ptr(out)))
        return {"triangle_tests": int(out[0]), "node_visits": int(out[1]),
                "fp64_ops": int(out[2]), "device_ms": out[3] / 1e3, "download_ms": out[4] / 1e3,
                "assembly_ms": out[5] / 1e3}

    def verification_rows(self, clusters: int, n: int):
        """Per-row results of the last verification: accepted (clusters, 2, n)
        and the nearest squared distance (NaN where the search did not run)."""
        acc = np.zeros(max(2 * clusters * n, 1), np.uint8)
        d2 = np.zeros(max(2 * clusters * n, 1))
        check(lib().mvs_verification_rows(self.h, ptr(acc), ptr(d2)))
        m = 2 * clusters * n
        return acc[:m].reshape(clusters, 2, n), d2[:m].reshape(clusters, 2, n)

    # ---------------------------------------------------- device layer
    def dev_prepare(self, vertices: np.ndarray, faces: np.ndarray, device=None) -> dict:
        """mvs_dev_prepare: the host mesh in, the prepared mesh and topology
        as device tensors (capacities as mvs_prepare's; face lanes have stride
        nf, the input face count) plus the host scalars."""
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        return self.dev_prepare_soa(*(np.ascontiguousarray(v[:, k]) for k in range(3)),
                                    *(np.ascontiguousarray(f[:, k]) for k in range(3)),
                                    device=device)

    def dev_prepare_soa(self, x, y, z, f0, f1, f2, device=None) -> dict:
        """dev_prepare over the reference's SoA mesh layout (contiguous
        float64 x, y, z and int32 face corner lanes)."""
        import torch
        dev = device or torch.device("cuda", torch.cuda.current_device())
        nv, nf = len(x), len(f0)
        cap = max(nf, 1)
        e = lambda *shape, dt=torch.float64: torch.empty(shape, dtype=dt, device=dev)
        out = dict(xyz=e(3, max(nv, 1)), faces3=e(3, cap, dt=torch.int32),
                   normals=e(3, max(nv, 1)), face_normals=e(3, cap), weights=e(max(nv, 1)),
                   nbr_off=e(nv + 1, dt=torch.int64), nbr_idx=e(6 * cap, dt=torch.int32),
                   inc_off=e(nv + 1, dt=torch.int64), inc_idx=e(3 * cap, dt=torch.int32),
                   edges=e(3 * cap, 6, dt=torch.int32))
        P = Prepared()
        P.f0, P.f1, P.f2 = (out["faces3"][k].data_ptr() for k in range(3))
        P.nx, P.ny, P.nz = (out["normals"][k].data_ptr() for k in range(3))
        P.fnx, P.fny, P.fnz = (out["face_normals"][k].data_ptr() for k in range(3))
        for k in ("weights", "nbr_off", "nbr_idx", "inc_off", "inc_idx", "edges"):
            setattr(P, k, out[k].data_ptr())
        check(lib().mvs_dev_prepare(self.h, C.c_int64(nv), ptr(x), ptr(y), ptr(z), C.c_int64(nf),
                                    ptr(f0), ptr(f1), ptr(f2), ptr(out["xyz"]), C.byref(P)))
        out.update(nv=nv, nf=nf, face_count=P.face_count, nbr_nnz=P.nbr_nnz,
                   inc_nnz=P.inc_nnz, edge_count=P.edge_count, box_size=P.box_size,
                   center=np.array(list(P.center)), rejected_faces=P.rejected_faces)
        return out

    def dev_prune(self, n, feat, valid, sample, prune: float, out) -> int:
        cnt = C.c_uint64()
        check(lib().mvs_dev_prune(self.h, C.c_uint64(n), ptr(feat), ptr(valid),
                                  C.c_uint64(sample.numel()), ptr(sample), C.c_double(prune),
                                  ptr(out), C.byref(cnt)))
        return cnt.value

    def dev_curvature(self, n, x, y, z, normals3, nf, fn3, off, idx, ne, edges, smoothing,
                      feat, valid):
        check(lib().mvs_dev_curvature(self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(normals3),
                                      C.c_uint64(nf), ptr(fn3), ptr(off), ptr(idx), C.c_uint64(ne),
                                      ptr(edges), C.c_int(smoothing), ptr(feat), ptr(valid)))

    def dev_pair_count(self, n, x, y, z, feat, S, sel, box, center, t, max_pairs,
                       precision) -> int:
        P = C.c_uint64()
        ctr = (C.c_double * 3)(*[float(c) for c in center])
        check(lib().mvs_dev_pair_count(self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat),
                                       C.c_uint64(S), ptr(sel), C.c_double(box), ctr,
                                       C.c_double(t), C.c_uint64(max_pairs), C.c_int(precision),
                                       C.byref(P)))
        return P.value

    def dev_pair_fill(self, a, b, ordinals: bool):
        check(lib().mvs_dev_pair_fill(self.h, ptr(a), ptr(b), C.c_int(1 if ordinals else 0)))

    def dev_pair_stats(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().mvs_dev_pair_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(window_candidates=a.value, index_candidates=b.value, fp64_rechecks=c.value)

    def dev_batch_pairvote(self, n, x, y, z, feat, sample_offsets, sel, boxes, centers, t,
                           max_pairs, max_votes, precision) -> BatchPairVoteResult:
        """sample_offsets (meshes + 1), boxes (meshes), centers (meshes x 3): host arrays."""
        off = np.ascontiguousarray(sample_offsets, dtype=np.uint64)
        bx = np.ascontiguousarray(boxes, dtype=np.float64)
        ct = np.ascontiguousarray(centers, dtype=np.float64)
        meshes = len(off) - 1
        assert len(bx) == meshes and ct.size == 3 * meshes
        r = BatchPairVoteResult()
        check(lib().mvs_dev_batch_pairvote(
            self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat), C.c_uint32(meshes),
            ptr(off), ptr(sel), ptr(bx), ptr(ct), C.c_double(t), C.c_uint64(max_pairs),
            C.c_uint64(max_votes), C.c_int(precision), C.byref(r)))
        return r

    def dev_batch_meanshift(self, seeds_per_sample, seed_value, iterations, h, cutoff,
                            want_stats=False) -> BatchMeanShiftResult:
        r = BatchMeanShiftResult()
        check(lib().mvs_dev_batch_meanshift(self.h, C.c_double(seeds_per_sample),
                                            C.c_uint64(seed_value), C.c_int(iterations),
                                            C.c_double(h), C.c_double(cutoff),
                                            C.c_int(1 if want_stats else 0), C.byref(r)))
        return r

    def batch_graph_capture(self, n, x, y, z, feat, sample_offsets, sel, boxes, centers, t,
                            max_pairs, max_votes, precision, seeds_per_sample, seed_value,
                            iterations, h, cutoff):
        off = np.ascontiguousarray(sample_offsets, dtype=np.uint64)
        bx = np.ascontiguousarray(boxes, dtype=np.float64)
        ct = np.ascontiguousarray(centers, dtype=np.float64)
        pv, ms = BatchPairVoteResult(), BatchMeanShiftResult()
        check(lib().mvs_batch_graph_capture(
            self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat), C.c_uint32(len(off) - 1),
            ptr(off), ptr(sel), ptr(bx), ptr(ct), C.c_double(t), C.c_uint64(max_pairs),
            C.c_uint64(max_votes), C.c_int(precision), C.c_double(seeds_per_sample),
            C.c_uint64(seed_value), C.c_int(iterations), C.c_double(h), C.c_double(cutoff),
            C.byref(pv), C.byref(ms)))
        return pv, ms

    def batch_autotune(self, n, x, y, z, feat, sample_offsets, sel, boxes, centers, t,
                       max_pairs, max_votes, precision, seeds_per_sample, seed_value,
                       iterations, h, cutoff):
        """Returns (table of candidate dicts with their replay ms, best index, pv, ms)."""
        off = np.ascontiguousarray(sample_offsets, dtype=np.uint64)
        bx = np.ascontiguousarray(boxes, dtype=np.float64)
        ct = np.ascontiguousarray(centers, dtype=np.float64)
        table = (TuneEntry * 16)()
        n_ent, best = C.c_uint32(), C.c_uint32()
        pv, ms = BatchPairVoteResult(), BatchMeanShiftResult()
        check(lib().mvs_batch_autotune(
            self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat), C.c_uint32(len(off) - 1),
            ptr(off), ptr(sel), ptr(bx), ptr(ct), C.c_double(t), C.c_uint64(max_pairs),
            C.c_uint64(max_votes), C.c_int(precision), C.c_double(seeds_per_sample),
            C.c_uint64(seed_value), C.c_int(iterations), C.c_double(h), C.c_double(cutoff),
            table, C.c_uint32(16), C.byref(n_ent), C.byref(best), C.byref(pv), C.byref(ms)))
        return [table[i].as_dict() for i in range(n_ent.value)], best.value, pv, ms

    def get_tuning(self) -> dict:
        e = TuneEntry()
        check(lib().mvs_session_get_tuning(self.h, C.byref(e)))
        return e.as_dict()

    def set_tuning(self, walk_quota: int, walk_blocks: int, wide_seeds: int,
                   seed_lanes: int = 32):
        e = TuneEntry(walk_quota, walk_blocks, wide_seeds, 0.0, seed_lanes, 0)
        check(lib().mvs_session_set_tuning(self.h, C.byref(e)))

    # ------------------------------------------------------ mode refinement
    def bvh_build(self, vertices: np.ndarray, faces: np.ndarray, center):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        f0, f1, f2 = (np.ascontiguousarray(f[:, k]) for k in range(3))
        ctr = (C.c_double * 3)(*[float(c) for c in center])
        check(lib().mvs_bvh_build(self.h, C.c_uint64(len(v)), ptr(x), ptr(y), ptr(z),
                                  C.c_uint64(len(f)), ptr(f0), ptr(f1), ptr(f2), ctr))

    def dev_closest_points(self, n, pts, q, nrm, dist2, face=None):
        check(lib().mvs_dev_closest_points(self.h, C.c_uint64(n), ptr(pts), ptr(q), ptr(nrm),
                                           ptr(dist2), ptr(face)))

    def dev_closest_modes(self, kind, modes, npts, theta, base, pts, shared, q, nrm, dist2,
                          face=None):
        check(lib().mvs_dev_closest_modes(self.h, C.c_int(kind), C.c_uint32(modes),
                                          C.c_uint32(npts), ptr(theta), ptr(base), ptr(pts),
                                          C.c_int(1 if shared else 0), ptr(q), ptr(nrm),
                                          ptr(dist2), ptr(face)))

    def dev_align_energy(self, kind, modes, npts, theta, base, pts, q, nrm, energy, grad,
                         hess=None):
        check(lib().mvs_dev_align_energy(self.h, C.c_int(kind), C.c_uint32(modes),
                                         C.c_uint32(npts), ptr(theta), ptr(base), ptr(pts),
                                         ptr(q), ptr(nrm), ptr(energy), ptr(grad), ptr(hess)))

    def step_graph_capture(self, n, x, y, z, feat, S, sel, box, center, t, max_pairs, max_votes,
                           precision, seeds, seed_value, iterations, h, cutoff):
        """Captures one pair + vote + mean-shift step as a HIP graph (see
        mvs_step_graph_capture); returns the recorded run's results, whose
        buffers every replay rewrites."""
        pv, ms = PairVoteResult(), MeanShiftResult()
        ctr = (C.c_double * 3)(*[float(c) for c in center])
        check(lib().mvs_step_graph_capture(
            self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat), C.c_uint64(S), ptr(sel),
            C.c_double(box), ctr, C.c_double(t), C.c_uint64(max_pairs), C.c_uint64(max_votes),
            C.c_int(precision), C.c_uint64(seeds), C.c_uint64(seed_value), C.c_int(iterations),
            C.c_double(h), C.c_double(cutoff), C.byref(pv), C.byref(ms)))
        return pv, ms

    def graph_launch(self):
        check(lib().mvs_graph_launch(self.h))

    def graph_check(self):
        check(lib().mvs_graph_check(self.h))

    def batch_graph_launch(self):
        check(lib().mvs_batch_graph_launch(self.h))

    def batch_graph_check(self):
        check(lib().mvs_batch_graph_check(self.h))

    def dev_pairvote(self, n, x, y, z, feat, S, sel, box, center, t, max_pairs, max_votes,
                     precision) -> PairVoteResult:
        r = PairVoteResult()
        ctr = (C.c_double * 3)(*[float(c) for c in center])
        check(lib().mvs_dev_pairvote(self.h, C.c_uint64(n), ptr(x), ptr(y), ptr(z), ptr(feat),
                                     C.c_uint64(S), ptr(sel), C.c_double(box), ctr, C.c_double(t),
                                     C.c_uint64(max_pairs), C.c_uint64(max_votes),
                                     C.c_int(precision), C.byref(r)))
        return r

    def dev_copy(self, dst, src_addr: int, nbytes: int):
        """Stream-ordered device-to-device copy out of session-owned memory."""
        check(lib().mvs_dev_copy(self.h, ptr(dst), C.c_void_p(src_addr), C.c_uint64(nbytes)))

    def fetch(self, addr: int, shape, dtype, stride_elems: int | None = None):
        """Copy a session-owned SoA block into a new torch tensor of `shape`
        (rows of shape[-1] elements; source rows `stride_elems` apart)."""
        import torch
        out = torch.empty(shape, dtype=dtype, device="cuda:0")
        if out.numel() == 0 or not addr:
            return out
        es = out.element_size()
        if stride_elems is None or len(shape) == 1:
            self.dev_copy(out, addr, out.numel() * es)
        else:
            for r in range(shape[0]):
                self.dev_copy(out[r], addr + r * stride_elems * es, shape[-1] * es)
        return out

    def dev_meanshift_votes(self, seeds, seed_value, iterations, h, cutoff,
                            want_stats=False) -> MeanShiftResult:
        r = MeanShiftResult()
        check(lib().mvs_dev_meanshift_votes(self.h, C.c_uint64(seeds), C.c_uint64(seed_value),
                                            C.c_int(iterations), C.c_double(h),
                                            C.c_double(cutoff), C.c_int(1 if want_stats else 0),
                                            C.byref(r)))
        return r

    def dev_ms_points(self, precision, nrigid, q4, t3, nrefl, n3, d1, rigid_pts, refl_pts):
        check(lib().mvs_dev_ms_points(self.h, C.c_int(precision), C.c_uint64(nrigid), ptr(q4),
                                      ptr(t3), C.c_uint64(nrefl), ptr(n3), ptr(d1),
                                      ptr(rigid_pts), ptr(refl_pts)))

    def dev_ms_batched(self, precision, dim, n, pts, ns, seeds, iterations, h, cutoff, out,
                       weight, stats: bool = False):
        cand, ins, upd = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(lib().mvs_dev_ms_batched(self.h, C.c_int(precision), C.c_int(dim), C.c_uint64(n),
                                       ptr(pts), C.c_uint64(ns), ptr(seeds), C.c_int(iterations),
                                       C.c_double(h), C.c_double(cutoff), ptr(out), ptr(weight),
                                       C.byref(cand) if stats else None,
                                       C.byref(ins) if stats else None,
                                       C.byref(upd) if stats else None))
        if stats == "updates":
            return cand.value, ins.value, upd.value
        return (cand.value, ins.value) if stats else None

    def dev_ms_batched_noalloc(self, precision, dim, n, pts, ns, seeds, iterations, h, cutoff,
                               out, weight):
        check(lib().mvs_dev_ms_batched_noalloc(
            self.h, C.c_int(precision), C.c_int(dim), C.c_uint64(n), ptr(pts), C.c_uint64(ns),
            ptr(seeds), C.c_int(iterations), C.c_double(h), C.c_double(cutoff), ptr(out),
            ptr(weight)))

    def dev_vote_count(self, P, a, b, precision) -> int:
        R = C.c_uint64()
        check(lib().mvs_dev_vote_count(self.h, C.c_uint64(P), ptr(a), ptr(b), C.c_int(precision),
                                       C.byref(R)))
        return R.value

    def dev_vote_emit(self, P, a, b, box, center, precision, q4, t3, rp, n3, d1, fp):
        ctr = (C.c_double * 3)(*[float(c) for c in center])
        check(lib().mvs_dev_vote_emit(self.h, C.c_uint64(P), ptr(a), ptr(b), C.c_double(box), ctr,
                                      C.c_int(precision), ptr(q4), ptr(t3), ptr(rp), ptr(n3),
                                      ptr(d1), ptr(fp)))
