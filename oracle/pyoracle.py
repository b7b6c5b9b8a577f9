"""ctypes front end of the CPU oracle (test infrastructure only).

`Oracle("oracle")` loads oracle/liboracle.so (the C restatement);
`Oracle("ref")` loads oracle/_ref/libref.so (the reference's own CPU sources
compiled by oracle/Makefile).  Both expose the same functions; results come
back as numpy arrays.  Only tests, the smoke entry and the CPU-baseline leg of
the benchmark use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OK, EINVAL, EBUDGET, EOVERFLOW, ERUNTIME = 0, 1, 2, 3, 6


class MT64(C.Structure):
    _fields_ = [("x", C.c_uint64 * 312), ("p", C.c_uint64)]

    def copy(self) -> "MT64":
        o = MT64()
        C.memmove(C.byref(o), C.byref(self), C.sizeof(MT64))
        return o

    def __eq__(self, other) -> bool:
        return bytes(self) == bytes(other)


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


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Topology:
    faces: np.ndarray          # (F, 3) int32 usable faces
    normals: np.ndarray        # (3, N)
    face_normals: np.ndarray   # (3, F)
    weights: np.ndarray        # (N,)
    nbr_off: np.ndarray
    nbr_idx: np.ndarray
    inc_off: np.ndarray
    inc_idx: np.ndarray
    edges: np.ndarray          # (E, 6) int32: lo, hi, a, b, face0, face1
    box_size: float
    center: np.ndarray         # (3,)
    rejected_faces: int

    @property
    def mean_edge(self) -> float:
        return float("nan")


@dataclass
class Clusters:
    reflection: np.ndarray
    rotation: np.ndarray       # (C, 9)
    translation: np.ndarray    # (C, 3)
    plane_normal: np.ndarray   # (C, 3)
    plane_offset: np.ndarray   # (C,)
    member_off: np.ndarray     # (C+1,)
    members: np.ndarray
    limited_iterations: int

    def __len__(self):
        return len(self.reflection)


class Oracle:
    def __init__(self, which: str = "oracle"):
        self.which = which
        path = os.path.join(HERE, "liboracle.so") if which == "oracle" else os.path.join(HERE, "_ref", "libref.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.px = "oracle_" if which == "oracle" else "ref_"
        err = getattr(self.lib, "oracle_last_error" if which == "oracle" else "ref_last_error")
        err.restype = C.c_char_p
        self._err = err
        seed = getattr(self.lib, "or_mt64_seed" if which == "oracle" else "ref_mt64_seed")
        seed.argtypes = [C.POINTER(MT64), C.c_uint64]
        nxt = getattr(self.lib, "or_mt64_next" if which == "oracle" else "ref_mt64_next")
        nxt.argtypes = [C.POINTER(MT64)]
        nxt.restype = C.c_uint64
        self._seed, self._next = seed, nxt

    def _fn(self, name):
        f = getattr(self.lib, self.px + name)
        f.restype = C.c_int
        return f

    def _check(self, rc):
        if rc != OK:
            raise OracleError(rc, self._err().decode())

    # ------------------------------------------------------------- rng
    def rng(self, seed: int) -> MT64:
        g = MT64()
        self._seed(C.byref(g), C.c_uint64(seed))
        return g

    def next(self, g: MT64) -> int:
        return int(self._next(C.byref(g)))

    # ----------------------------------------------------------- stages
    def prepare(self, vertices: np.ndarray, faces: np.ndarray) -> Topology:
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
            setattr(P, k, _p(a))
        rc = self._fn("prepare")(C.c_int64(nv), _p(x), _p(y), _p(z), C.c_int64(nf),
                                 _p(f0), _p(f1), _p(f2), C.byref(P))
        self._check(rc)
        F = P.face_count
        return Topology(
            faces=np.stack([out["f0"][:F], out["f1"][:F], out["f2"][:F]], 1),
            normals=np.stack([out["nx"][:nv], out["ny"][:nv], out["nz"][:nv]]),
            face_normals=np.stack([out["fnx"][:F], out["fny"][:F], out["fnz"][:F]]),
            weights=out["weights"][:nv].copy(),
            nbr_off=out["nbr_off"], nbr_idx=out["nbr_idx"][:P.nbr_nnz].copy(),
            inc_off=out["inc_off"], inc_idx=out["inc_idx"][:P.inc_nnz].copy(),
            edges=out["edges"][:P.edge_count].copy(), box_size=P.box_size,
            center=np.array(list(P.center)), rejected_faces=P.rejected_faces)

    def curvature(self, vertices: np.ndarray, topo: Topology, smoothing: int = 3):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        nv = len(v)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        n = np.ascontiguousarray(topo.normals)
        fn = np.ascontiguousarray(topo.face_normals)
        feat = np.zeros((11, nv))
        valid = np.zeros(nv, np.uint8)
        edges = np.ascontiguousarray(topo.edges, dtype=np.int32)
        rc = self._fn("curvature")(
            C.c_int64(nv), _p(x), _p(y), _p(z), _p(n[0]), _p(n[1]), _p(n[2]),
            _p(fn[0]), _p(fn[1]), _p(fn[2]), _p(topo.nbr_off), _p(topo.nbr_idx),
            C.c_int64(len(edges)), _p(edges), C.c_int(smoothing), _p(feat), _p(valid))
        self._check(rc)
        return feat, valid

    def sample_vertices(self, weights: np.ndarray, ratio: float, g: MT64) -> np.ndarray:
        assert self.which == "oracle"
        w = np.ascontiguousarray(weights, dtype=np.float64)
        out = np.zeros(max(len(w), 1), np.int32)
        cnt = C.c_int64()
        rc = self._fn("sample_vertices")(C.c_int64(len(w)), _p(w), C.c_double(ratio),
                                         C.byref(g), _p(out), C.byref(cnt))
        self._check(rc)
        return out[:cnt.value].copy()

    def prune(self, kmin, kmax, valid, sample, prune: float = 0.85) -> np.ndarray:
        assert self.which == "oracle"
        s = np.ascontiguousarray(sample, dtype=np.int32)
        out = np.zeros(max(len(s), 1), np.int32)
        cnt = C.c_int64()
        rc = self._fn("prune")(C.c_int64(len(kmin)), _p(np.ascontiguousarray(kmin)),
                               _p(np.ascontiguousarray(kmax)),
                               _p(np.ascontiguousarray(valid, dtype=np.uint8)),
                               C.c_int64(len(s)), _p(s), C.c_double(prune), _p(out),
                               C.byref(cnt))
        self._check(rc)
        return out[:cnt.value].copy()

    def select(self, weights, kmin, kmax, valid, ratio, prune, g: MT64):
        """Backend::select = sampling + prune (src/cpu/selection.cpp:15-29)."""
        if self.which == "oracle":
            sample = self.sample_vertices(weights, ratio, g)
            return sample, self.prune(kmin, kmax, valid, sample, prune)
        nv = len(weights)
        sample = np.zeros(max(nv, 1), np.int32)
        pruned = np.zeros(max(nv, 1), np.int32)
        ns, npr = C.c_int64(), C.c_int64()
        rc = self._fn("select")(C.c_int64(nv), _p(np.ascontiguousarray(weights, dtype=np.float64)),
                                _p(np.ascontiguousarray(kmin)), _p(np.ascontiguousarray(kmax)),
                                _p(np.ascontiguousarray(valid, dtype=np.uint8)),
                                C.c_double(ratio), C.c_double(prune), C.byref(g),
                                _p(sample), C.byref(ns), _p(pruned), C.byref(npr))
        self._check(rc)
        return sample[:ns.value].copy(), pruned[:npr.value].copy()

    def _mt(self, name, threads):
        """The single-thread entry, or (threads > 1, this oracle only) its
        multi-threaded counterpart with the thread count bound first."""
        if threads <= 1:
            return self._fn(name)
        assert self.which == "oracle", "only the oracle has multi-threaded runs"
        f = self._fn(name + "_mt")
        return lambda *a: f(C.c_int(threads), *a)

    def pairing(self, vertices, kmin, kmax, selected, box: float, threshold: float,
                max_pairs: int = 200000, cap: int | None = None, threads: int = 1):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        s = np.ascontiguousarray(selected, dtype=np.int32)
        ns = len(s)
        if cap is None:
            cap = int(min(max_pairs, ns * (ns - 1) // 2 if ns > 1 else 0))
        cap = max(cap, 1)
        a = np.zeros(cap, np.int32)
        b = np.zeros(cap, np.int32)
        cnt = C.c_int64()
        rc = self._mt("pairing", threads)(C.c_int64(len(v)), _p(x), _p(y), _p(z),
                                 _p(np.ascontiguousarray(kmin, dtype=np.float64)),
                                 _p(np.ascontiguousarray(kmax, dtype=np.float64)),
                                 C.c_int64(ns), _p(s), C.c_double(box), C.c_double(threshold),
                                 C.c_uint64(max_pairs), C.c_int64(cap), _p(a), _p(b),
                                 C.byref(cnt))
        self._check(rc)
        return a[:cnt.value].copy(), b[:cnt.value].copy()

    def transformations(self, vertices, feat, pa, pb, box: float, center,
                        max_votes: int = 600000, cap: int | None = None, threads: int = 1):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        x, y, z = (np.ascontiguousarray(v[:, k]) for k in range(3))
        feat = np.ascontiguousarray(feat, dtype=np.float64)
        pa = np.ascontiguousarray(pa, dtype=np.int32)
        pb = np.ascontiguousarray(pb, dtype=np.int32)
        if cap is None:
            cap = int(min(max_votes, 3 * len(pa)))
        cap = max(cap, 1)
        pair = np.zeros(cap, np.int32)
        kind = np.zeros(cap, np.uint8)
        lanes = np.zeros((28, cap))
        ctr = np.ascontiguousarray(center, dtype=np.float64)
        cnt = C.c_int64()
        rc = self._mt("transformations", threads)(C.c_int64(len(v)), _p(x), _p(y), _p(z), _p(feat),
                                         C.c_int64(len(pa)), _p(pa), _p(pb), C.c_double(box),
                                         _p(ctr), C.c_uint64(max_votes), C.c_int64(cap),
                                         _p(pair), _p(kind), _p(lanes), C.byref(cnt))
        self._check(rc)
        V = cnt.value
        return pair[:V].copy(), kind[:V].copy(), np.ascontiguousarray(lanes[:, :V])

    def clustering(self, kind, desc, rotation, box: float, center, bandwidth: float,
                   max_iterations: int, max_clusters: int, g: MT64) -> Clusters:
        kind = np.ascontiguousarray(kind, dtype=np.uint8)
        V = len(kind)
        desc = np.ascontiguousarray(desc, dtype=np.float64).reshape(12, V)
        rotation = np.ascontiguousarray(rotation, dtype=np.float64).reshape(9, V)
        mc = max(max_clusters, 1)
        ncl = C.c_int32()
        lim = C.c_int32()
        refl = np.zeros(mc, np.uint8)
        rot = np.zeros((mc, 9))
        tr = np.zeros((mc, 3))
        pn = np.zeros((mc, 3))
        po = np.zeros(mc)
        moff = np.zeros(mc + 1, np.uint64)
        mem = np.zeros(max(V, 1), np.int32)
        ctr = np.ascontiguousarray(center, dtype=np.float64)
        rc = self._fn("clustering")(C.c_int64(V), _p(kind), _p(desc), _p(rotation),
                                    C.c_double(box), _p(ctr), C.c_double(bandwidth),
                                    C.c_int(max_iterations), C.c_int(max_clusters),
                                    C.byref(g), C.byref(ncl), _p(refl), _p(rot), _p(tr),
                                    _p(pn), _p(po), _p(moff), _p(mem), C.byref(lim))
        self._check(rc)
        n = ncl.value
        return Clusters(refl[:n].copy(), rot[:n].copy(), tr[:n].copy(), pn[:n].copy(),
                        po[:n].copy(), moff[:n + 1].copy(), mem[:int(moff[n])].copy(),
                        lim.value)

    def meanshift_batched(self, points, kind, dim_rigid, dim_reflection, seeds,
                          seed_kind, h, cutoff, iterations, threads: int = 1):
        assert self.which == "oracle"
        pts = np.ascontiguousarray(points, dtype=np.float64)
        dim, nv = pts.shape
        sd = np.ascontiguousarray(seeds, dtype=np.float64)
        ns = sd.shape[1]
        out = np.zeros((dim, ns))
        w = np.zeros(ns)
        rc = self._mt("meanshift_batched", threads)(
            C.c_int64(nv), C.c_int(dim), _p(pts), _p(np.ascontiguousarray(kind, dtype=np.uint8)),
            C.c_int(dim_rigid), C.c_int(dim_reflection), C.c_int64(ns), _p(sd),
            _p(np.ascontiguousarray(seed_kind, dtype=np.uint8)), C.c_double(h),
            C.c_double(cutoff), C.c_int(iterations), _p(out), _p(w))
        self._check(rc)
        return out, w

    # ------------------------------------------------------ verification
    def point_triangle_distance2(self, p, a, b, c) -> float:
        f = getattr(self.lib, self.px + "point_triangle_distance2")
        f.restype = C.c_double
        arr = [np.ascontiguousarray(v, dtype=np.float64) for v in (p, a, b, c)]
        return float(f(*[_p(v) for v in arr]))

    @staticmethod
    def _verify_args(vertices, faces, normals, clusters):
        v = np.ascontiguousarray(vertices, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        n = np.ascontiguousarray(normals, dtype=np.float64).reshape(3, len(v))
        cols = [np.ascontiguousarray(v[:, k]) for k in range(3)]
        fc = [np.ascontiguousarray(f[:, k]) for k in range(3)]
        nc = len(clusters["reflection"])
        cl = [np.ascontiguousarray(clusters["reflection"], dtype=np.uint8)] + [
            np.ascontiguousarray(clusters[k], dtype=np.float64).reshape(nc, w)
            for k, w in (("rotation", 9), ("translation", 3), ("plane_normal", 3),
                         ("plane_offset", 1))]
        return len(v), len(f), nc, cols, fc, [n[0], n[1], n[2]], cl

    def verification_rows(self, vertices, faces, normals, clusters, verify_threshold, box):
        """oracle_verification_rows: (state, dist2), each (clusters, 2, nv)."""
        assert self.which == "oracle"
        nv, nf, nc, cols, fc, nrm, cl = self._verify_args(vertices, faces, normals, clusters)
        state = np.zeros(max(2 * nc * nv, 1), np.uint8)
        d2 = np.zeros(max(2 * nc * nv, 1))
        rc = self._fn("verification_rows")(
            C.c_int64(nv), *[_p(a) for a in cols], C.c_int64(nf), *[_p(a) for a in fc],
            *[_p(np.ascontiguousarray(a)) for a in nrm], C.c_int64(nc), *[_p(a) for a in cl],
            C.c_double(verify_threshold), C.c_double(box), _p(state), _p(d2))
        self._check(rc)
        m = 2 * nc * nv
        return state[:m].reshape(nc, 2, nv), d2[:m].reshape(nc, 2, nv)

    def components(self, nbr_off, nbr_idx, ids):
        """oracle_components: list of sorted patches of `ids`."""
        assert self.which == "oracle"
        off = np.ascontiguousarray(nbr_off, dtype=np.uint64)
        idx = np.ascontiguousarray(nbr_idx, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        po = np.zeros(len(ids) + 1, np.uint64)
        pm = np.zeros(max(len(ids), 1), np.int32)
        npat = C.c_int64()
        self._check(self._fn("components")(C.c_int64(len(off) - 1), _p(off), _p(idx),
                                           C.c_int64(len(ids)), _p(ids), _p(po), _p(pm),
                                           C.byref(npat)))
        return [pm[po[k]:po[k + 1]].copy() for k in range(npat.value)]

    def verification(self, vertices, faces, topo: Topology, clusters, verify_threshold, box):
        """The compiled reference's Backend::verification (ref only): dict of
        side_a / side_b (offsets, indices), patches_a / patches_b (groups, row
        offsets, members), coverage, rms."""
        assert self.which == "ref"
        nv, nf, nc, cols, fc, nrm, cl = self._verify_args(vertices, faces, topo.normals, clusters)
        cap = max(nc * nv, 1)
        bufs = []
        for _ in range(2):
            bufs.append([np.zeros(nc + 1, np.uint64), np.zeros(cap, np.int32)])
        pats = []
        for _ in range(2):
            pats.append([np.zeros(nc + 1, np.uint64), np.zeros(cap + 1, np.uint64),
                         np.zeros(cap, np.int32)])
        cov, rms = np.zeros(max(nc, 1)), np.zeros(max(nc, 1))
        counts = np.zeros(4, np.uint64)
        rc = self._fn("verification")(
            C.c_int64(nv), *[_p(a) for a in cols], C.c_int64(nf), *[_p(a) for a in fc],
            *[_p(np.ascontiguousarray(a)) for a in nrm],
            _p(np.ascontiguousarray(topo.weights, dtype=np.float64)),
            _p(np.ascontiguousarray(topo.nbr_off, dtype=np.uint64)),
            _p(np.ascontiguousarray(topo.nbr_idx, dtype=np.int32)), C.c_int64(nc),
            *[_p(a) for a in cl], C.c_double(verify_threshold), C.c_double(box),
            _p(bufs[0][0]), _p(bufs[0][1]), _p(bufs[1][0]), _p(bufs[1][1]),
            *[_p(a) for a in pats[0]], *[_p(a) for a in pats[1]], _p(cov), _p(rms), _p(counts))
        self._check(rc)
        sa, sb, pa, pb = (int(c) for c in counts)
        return dict(side_a=(bufs[0][0], bufs[0][1][:sa].copy()),
                    side_b=(bufs[1][0], bufs[1][1][:sb].copy()),
                    patches_a=(pats[0][0], pats[0][1][:pa + 1].copy(), pats[0][2][:sa].copy()),
                    patches_b=(pats[1][0], pats[1][1][:pb + 1].copy(), pats[1][2][:sb].copy()),
                    coverage=cov[:nc].copy(), rms=rms[:nc].copy())


def read_off(path: str):
    """Minimal OFF reader (triangles only) for the example fixtures."""
    with open(path) as f:
        tokens = f.read().split()
    assert tokens[0] == "OFF"
    nv, nf = int(tokens[1]), int(tokens[2])
    pos = 4
    v = np.array(tokens[pos:pos + 3 * nv], dtype=np.float64).reshape(nv, 3)
    pos += 3 * nv
    faces = []
    for _ in range(nf):
        k = int(tokens[pos])
        faces.append([int(t) for t in tokens[pos + 1:pos + 1 + k]])
        pos += 1 + k
    return v, np.array(faces, dtype=np.int32)
