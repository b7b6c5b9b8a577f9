"""Device-layer parity: the fused fast path (pair count/fill on device-resident
features, compact vote emission, batched mean-shift) against the oracle.

Bars: pair set byte-equal for both arithmetics; FP64 compact votes equal
the oracle's rotation / centred translation / plane within 1e-12 (the
quaternion round trip costs a few ulps); FP32-guarded votes carry exactly
the FP64 mirror decisions and agree within 2e-5 (rotation entries, plane
normal) and 2e-5 box (translations, plane offset); batched mean-shift FP64
within 1e-9 of the oracle, FP32 within 2e-5.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi, meshgen, workload  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

pytestmark = pytest.mark.gpu


def quat_to_R(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


@pytest.fixture(scope="module")
def cfg1():
    return workload.cfg1()


def _oracle_votes(oracle, wl, pa, pb):
    v = wl.mesh.vertices
    return oracle.transformations(v, wl.feat_np, pa, pb, wl.box, wl.center, 10**9)


# Per-source pair counts on this workload: t = .05 keeps every source at
# most 15 partners (slab, one pass of the bitonic sort), t = .15 has sources
# with 65..72 partners (slab, LDS rank sort), t = .3 has sources beyond 128
# partners (chunked long path).
@pytest.mark.parametrize("t", [0.05, 0.15, 0.3])
def test_compact_votes_match_oracle(cfg1, oracle, t):
    wl = cfg1
    wl.threshold = t
    v = wl.mesh.vertices
    pa, pb = oracle.pairing(v, wl.feat_np[9], wl.feat_np[10], wl.selected_np, wl.box, t, 10**8)
    _, per_source = np.unique(pa, return_counts=True)
    if t >= 0.15:
        assert np.any((per_source > 64) & (per_source <= 128))
    if t >= 0.3:
        assert np.any(per_source > 128)
    pair, kind, lanes = _oracle_votes(oracle, wl, pa, pb)
    ref = {}
    for prec in (capi.FP64, capi.FP32_GUARDED):
        r = wl.pair_votes(prec)
        sel = wl.selected_np
        a = sel[r.a[:r.P].cpu().numpy()]
        b = sel[r.b[:r.P].cpu().numpy()]
        assert np.array_equal(a, pa) and np.array_equal(b, pb)
        q = r.q4[:, :2 * r.P].double().cpu().numpy()
        tc = r.t3[:, :2 * r.P].double().cpu().numpy()
        n3 = r.n3[:, :r.R].double().cpu().numpy()
        d1 = r.d1[:r.R].double().cpu().numpy()
        fp = r.fp[:r.R].cpu().numpy()
        rp = r.rp[:2 * r.P].cpu().numpy()
        assert np.array_equal(rp, np.repeat(np.arange(r.P), 2))
        # oracle votes per pair: [mirror], rigid(+), rigid(-)
        refl_pairs = pair[kind == 1]
        assert np.array_equal(fp, refl_pairs)
        rig = np.where(kind == 0)[0]
        assert len(rig) == 2 * r.P
        # the documented bars: FP64 2e-12 (tests/cuda_backend.cpp), FP32 2e-5
        tol = 2e-12 if prec == capi.FP64 else 2e-5
        R_or = lanes[0:9, rig].T.reshape(-1, 3, 3)
        R_gpu = np.stack([quat_to_R(q[:, i]) for i in range(2 * r.P)])
        assert np.max(np.abs(R_gpu - R_or)) < tol
        tc_or = lanes[25:28, rig] * wl.box
        assert np.max(np.abs(tc - tc_or)) < tol * wl.box
        mir = np.where(kind == 1)[0]
        assert np.max(np.abs(n3 - lanes[12:15, mir]), initial=0) < tol
        assert np.max(np.abs(d1 - lanes[19, mir] * wl.box), initial=0) < tol * wl.box
        ref[prec] = r
    st = wl.sess.dev_pair_stats()
    assert st["index_candidates"] <= st["window_candidates"]


def test_pair_stats_and_identity_of_arithmetics():
    wl = workload.Workload(meshgen.cfg2_mesh(256), 20000, 0.02, False)
    r64 = wl.pair_votes(capi.FP64)
    s64 = wl.sess.dev_pair_stats()
    r32 = wl.pair_votes(capi.FP32_GUARDED)
    s32 = wl.sess.dev_pair_stats()
    assert r64.P == r32.P and r64.R == r32.R and r64.P > 0
    assert torch.equal(r64.a[:r64.P], r32.a[:r32.P]) and torch.equal(r64.b[:r64.P], r32.b[:r32.P])
    assert torch.equal(r64.fp[:r64.R], r32.fp[:r32.R])
    assert s32["fp64_rechecks"] < s32["index_candidates"]
    assert s64["fp64_rechecks"] == 0
    o = Oracle("oracle")
    pa, pb = o.pairing(wl.mesh.vertices, wl.feat_np[9], wl.feat_np[10], wl.selected_np, wl.box,
                       0.02, 10**9)
    sel = wl.selected_np
    assert np.array_equal(sel[r64.a[:r64.P].cpu().numpy()], pa)
    assert np.array_equal(sel[r64.b[:r64.P].cpu().numpy()], pb)


def _clusters(D, seed, spread=1.0):
    g = np.random.default_rng(seed)
    centers = g.uniform(-spread, spread, (3, D))
    pts = [c + g.normal(0, 0.02, (300, D)) for c in centers]
    pts.append(g.uniform(-1.2 * spread, 1.2 * spread, (200, D)))
    P = np.concatenate(pts).T.copy()
    seeds = P[:, g.choice(P.shape[1], 64, replace=False)].copy()
    return P, seeds


# spread 1: half-width cells (reach 2); spread 8: cells of side 3h (reach 1);
# spread 200: reach 1 with cells doubled until the grid fits 2^22 cells
# lanes 256: a block of four waves per seed (rows cut to the support ball,
# the first wave finding the rows); 64, 32, 16: lanes per narrow seed
@pytest.mark.parametrize("lanes", [256, 64, 32, 16])
@pytest.mark.parametrize("spread", [1.0, 8.0, 200.0])
@pytest.mark.parametrize("D", [6, 4])
def test_batched_meanshift_matches_oracle(oracle, D, spread, lanes):
    P, seeds = _clusters(D, 7 + D, spread)
    h, cut, iters = 0.05, 3.0, 25
    want, wsum = oracle.meanshift_batched(P, np.zeros(P.shape[1], np.uint8), D, D, seeds,
                                          np.zeros(seeds.shape[1], np.uint8), h, cut, iters)
    s = capi.Session(0)
    s.set_tuning(0, 2048, 1 << 40 if lanes == 256 else 0, 64 if lanes == 256 else lanes)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    for prec, dt, tol in ((capi.FP64, torch.float64, 1e-9), (capi.FP32_GUARDED, torch.float32,
                                                             2e-5)):
        pts = torch.as_tensor(P, dtype=dt, device=dev).contiguous()
        sd = torch.as_tensor(seeds, dtype=dt, device=dev).contiguous()
        out = torch.empty_like(sd)
        w = torch.empty(seeds.shape[1], dtype=dt, device=dev)
        cand, ins = s.dev_ms_batched(prec, D, P.shape[1], pts, seeds.shape[1], sd, iters, h, cut,
                                     out, w, stats=True)
        torch.cuda.synchronize()
        got = out.double().cpu().numpy()
        assert np.max(np.abs(got - want)) < tol, float(np.max(np.abs(got - want)))
        assert np.max(np.abs(w.double().cpu().numpy() - wsum) / (1 + wsum)) < tol
        assert 0 < ins <= cand
    s.close()


@pytest.mark.parametrize("lanes", [256, 64, 32, 16])
@pytest.mark.parametrize("spread", [1.0, 8.0])
@pytest.mark.parametrize("D", [6, 4])
def test_meanshift_support_visits_match_brute_force(D, spread, lanes):
    """One update from each seed visits exactly the points inside its support
    (counted by brute force in FP64): the grid walk, with its rows cut to the
    support ball, drops no point of the support."""
    P, seeds = _clusters(D, 3 + D, spread)
    h, cut = 0.05, 3.0
    s = capi.Session(0)
    s.set_tuning(0, 2048, 1 << 40 if lanes == 256 else 0, 64 if lanes == 256 else lanes)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    pts = torch.as_tensor(P, dtype=torch.float64, device=dev).contiguous()
    sd = torch.as_tensor(seeds, dtype=torch.float64, device=dev).contiguous()
    out = torch.empty_like(sd)
    w = torch.empty(seeds.shape[1], dtype=torch.float64, device=dev)
    cand, ins = s.dev_ms_batched(capi.FP64, D, P.shape[1], pts, seeds.shape[1], sd, 1, h, cut,
                                 out, w, stats=True)
    torch.cuda.synchronize()
    s.close()
    d2 = ((P[:, None, :] - seeds[:, :, None]) ** 2).sum(0)
    inside = int(np.sum(d2 < (cut * h) ** 2))
    # points within rounding of the boundary could go either way; there are none
    near = int(np.sum(np.abs(d2 - (cut * h) ** 2) < 1e-12))
    assert near == 0
    assert ins == inside, (ins, inside)
    assert cand >= ins


@pytest.mark.parametrize("lanes", [32, 16])
@pytest.mark.parametrize("D", [6, 4])
def test_narrow_query_long_candidate_lists(oracle, D, lanes):
    """Seeds inside a dense cluster have more candidates than a seed group's
    LDS map holds (512 entries at 32 lanes per seed, 256 at 16), so the
    narrow query walks its rows in several chunks; seeds between clusters
    move far enough to cut new rows.  Modes and weights against the oracle."""
    g = np.random.default_rng(11 + D)
    dense = g.normal(0.0, 0.03, (2500, D))
    other = np.array([0.4] + [0.0] * (D - 1)) + g.normal(0.0, 0.05, (800, D))
    P = np.concatenate([dense, other, g.uniform(-1.0, 1.0, (400, D))]).T.copy()
    seeds = P[:, g.choice(P.shape[1], 200, replace=False)].copy()
    h, cut, iters = 0.05, 3.0, 25
    # most dense points lie inside the support of the cluster centre
    assert np.sum(np.sum(dense ** 2, axis=1) < (cut * h) ** 2) > 4 * 512
    want, wsum = oracle.meanshift_batched(P, np.zeros(P.shape[1], np.uint8), D, D, seeds,
                                          np.zeros(seeds.shape[1], np.uint8), h, cut, iters)
    s = capi.Session(0)
    s.set_tuning(0, 2048, 0, lanes)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    for prec, dt, tol in ((capi.FP64, torch.float64, 1e-9), (capi.FP32_GUARDED, torch.float32,
                                                             2e-5)):
        pts = torch.as_tensor(P, dtype=dt, device=dev).contiguous()
        sd = torch.as_tensor(seeds, dtype=dt, device=dev).contiguous()
        out = torch.empty_like(sd)
        w = torch.empty(seeds.shape[1], dtype=dt, device=dev)
        cand, ins = s.dev_ms_batched(prec, D, P.shape[1], pts, seeds.shape[1], sd, iters, h, cut,
                                     out, w, stats=True)
        torch.cuda.synchronize()
        got = out.double().cpu().numpy()
        assert np.max(np.abs(got - want)) < tol, float(np.max(np.abs(got - want)))
        assert np.max(np.abs(w.double().cpu().numpy() - wsum) / (1 + wsum)) < tol
        assert 0 < ins <= cand
    s.close()


def test_batched_meanshift_empty_and_degenerate(oracle):
    s = capi.Session(0)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    pts = torch.zeros((6, 0), dtype=torch.float64, device=dev)
    sd = torch.ones((6, 3), dtype=torch.float64, device=dev)
    out = torch.empty_like(sd)
    w = torch.empty(3, dtype=torch.float64, device=dev)
    s.dev_ms_batched(capi.FP64, 6, 0, pts, 3, sd, 5, 0.05, 3.0, out, w)
    torch.cuda.synchronize()
    assert torch.equal(out, sd) and torch.all(w == 0)
    # all points identical: one iteration lands exactly on them
    P = np.tile(np.array([[0.3], [0.1], [-0.2], [0.05], [0.0], [0.01]]), (1, 50))
    pts = torch.as_tensor(P, device=dev).contiguous()
    sd = torch.as_tensor(P[:, :2] + 0.01, device=dev).contiguous()
    out = torch.empty_like(sd)
    w = torch.empty(2, dtype=torch.float64, device=dev)
    s.dev_ms_batched(capi.FP64, 6, 50, pts, 2, sd, 3, 0.05, 3.0, out, w)
    torch.cuda.synchronize()
    assert np.allclose(out.cpu().numpy(), P[:, :2], atol=1e-15)
    s.close()


def _fetch_pairvote(sess, r, prec):
    dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
    P, R, st = r.pairs, r.reflections, r.refl_stride
    return dict(
        a=sess.fetch(r.a, (P,), torch.int32), b=sess.fetch(r.b, (P,), torch.int32),
        q4=sess.fetch(r.rigid_q4, (4, 2 * P), dt, 2 * P),
        t3=sess.fetch(r.rigid_t3, (3, 2 * P), dt, 2 * P),

        n3=sess.fetch(r.refl_n3, (3, R), dt, st), d1=sess.fetch(r.refl_d, (R,), dt),
        fp=sess.fetch(r.refl_pair, (R,), torch.int32))


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32_GUARDED])
def test_fused_pairvote_matches_stepwise(cfg1, prec):
    wl = cfg1
    wl.threshold = 0.05
    step = wl.pair_votes(prec)
    fused = _fetch_pairvote(wl.sess, wl.pairvote(prec), prec)
    torch.cuda.synchronize()
    P, R = step.P, step.R
    assert fused["a"].numel() == P and fused["fp"].numel() == R
    assert torch.equal(fused["a"], step.a[:P]) and torch.equal(fused["b"], step.b[:P])
    assert torch.equal(fused["q4"], step.q4[:, :2 * P]) and torch.equal(fused["t3"],
                                                                       step.t3[:, :2 * P])
    assert torch.equal(fused["n3"], step.n3[:, :R]) and torch.equal(fused["d1"], step.d1[:R])
    assert torch.equal(fused["fp"], step.fp[:R])
    assert r_fused_rigid_pair_is_implicit(wl, prec)
    with pytest.raises(capi.MvsBudgetError):
        wl.pairvote(prec, max_pairs=max(P - 1, 0))
    with pytest.raises(capi.MvsBudgetError):
        wl.pairvote(prec, max_votes=2 * P + R - 1)


def r_fused_rigid_pair_is_implicit(wl, prec):
    # the fused step writes no rigid pair ids: vote v belongs to pair v / 2
    return wl.pairvote(prec).rigid_pair is None


def _rodrigues(r):
    th = np.linalg.norm(r)
    if th < 1e-15:
        return np.eye(3)
    k = r / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def test_fused_meanshift_matches_oracle(cfg1, oracle):
    wl = cfg1
    wl.threshold = 0.05
    prec = capi.FP64
    r = wl.pairvote(prec)
    f = _fetch_pairvote(wl.sess, r, prec)
    P, R = r.pairs, r.reflections
    rp = torch.empty((6, 2 * P), dtype=torch.float64, device="cuda:0")
    fpts = torch.empty((4, R), dtype=torch.float64, device="cuda:0")
    wl.sess.dev_ms_points(prec, 2 * P, f["q4"], f["t3"], R, f["n3"].contiguous(), f["d1"], rp,
                          fpts)
    ms = wl.meanshift(seeds=300, iterations=12, want_stats=True)
    seeds = wl.sess.fetch(ms.seed_votes, (ms.rigid_seeds + ms.refl_seeds,), torch.int32)
    got_r = wl.sess.fetch(ms.rigid_modes, (6, ms.rigid_seeds), torch.float64, ms.rigid_seeds)
    got_f = wl.sess.fetch(ms.refl_modes, (4, ms.refl_seeds), torch.float64, ms.refl_seeds)
    torch.cuda.synchronize()
    assert ms.rigid_seeds + ms.refl_seeds == 300 and ms.refl_seeds > 0
    seeds = seeds.cpu().numpy()
    # stratified draws over the cell order: distinct, inside each stream
    sr, sf = seeds[:ms.rigid_seeds], seeds[ms.rigid_seeds:]
    assert len(np.unique(sr)) == len(sr) and sr.min() >= 0 and sr.max() < 2 * P
    assert len(np.unique(sf)) == len(sf) and sf.min() >= 0 and sf.max() < R
    P6 = rp.cpu().numpy()
    P4 = fpts.cpu().numpy()
    # axis-angle points reproduce the vote rotations (closed form: Rodrigues)
    q = f["q4"].cpu().numpy()
    for i in range(0, 2 * P, max(1, (2 * P) // 50)):
        w, x, y, z = q[:, i]
        Rq = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                       [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                       [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(_rodrigues(P6[3:6, i]), Rq, atol=1e-9)
    kr = ms.rigid_seeds
    want_r, _ = oracle.meanshift_batched(P6, np.zeros(P6.shape[1], np.uint8), 6, 6,
                                         P6[:, seeds[:kr]], np.zeros(kr, np.uint8), 0.05, 3.0, 12)
    want_f, _ = oracle.meanshift_batched(P4, np.zeros(P4.shape[1], np.uint8), 4, 4,
                                         P4[:, seeds[kr:]], np.zeros(len(seeds) - kr, np.uint8),
                                         0.05, 3.0, 12)
    assert np.max(np.abs(got_r.cpu().numpy() - want_r)) < 1e-9
    assert np.max(np.abs(got_f.cpu().numpy() - want_f)) < 1e-9
    assert 0 < ms.in_support <= ms.candidates
