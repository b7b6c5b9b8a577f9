"""Batches of meshes: one batch pair + vote pass over several meshes of
uneven size equals, mesh by mesh, the one-mesh pass (pairs, their order and
the compact votes bit for bit), for both arithmetics."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi, workload  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch():
    return workload.Batch(count=10, seed=11, min_side=24, max_side=96, keep_single=True)


def _fetch(sess, r, prec):
    dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
    P, R, st = r.pairs, r.reflections, r.refl_stride
    f = dict(a=sess.fetch(r.a, (P,), torch.int32), b=sess.fetch(r.b, (P,), torch.int32),
             q4=sess.fetch(r.rigid_q4, (4, 2 * P), dt, 2 * P),
             t3=sess.fetch(r.rigid_t3, (3, 2 * P), dt, 2 * P),
             n3=sess.fetch(r.refl_n3, (3, R), dt, st), d1=sess.fetch(r.refl_d, (R,), dt),
             fp=sess.fetch(r.refl_pair, (R,), torch.int32))
    return {k: v.cpu().clone() for k, v in f.items()}


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32_GUARDED])
def test_batch_pairvote_equals_single_meshes(batch, prec):
    r = batch.pairvote(prec)
    allr = _fetch(batch.sess, r.all, prec)
    M = batch.count
    poff = batch.sess.fetch(r.pair_offsets, (M + 1,), torch.int64).cpu().numpy()
    roff = batch.sess.fetch(r.refl_offsets, (M + 1,), torch.int64).cpu().numpy()
    torch.cuda.synchronize()
    assert poff[0] == 0 and poff[-1] == r.all.pairs and roff[-1] == r.all.reflections
    single = capi.Session(0)
    single.set_stream(torch.cuda.current_stream().cuda_stream)
    total = 0
    for m, s in enumerate(batch.single):
        p0, p1, r0, r1 = poff[m], poff[m + 1], roff[m], roff[m + 1]
        o = int(batch.soff[m])
        one = _fetch(single, single.dev_pairvote(s["n"], s["x"], s["y"], s["z"], s["feat"],
                                                s["S"], s["sel"], s["box"], s["center"],
                                                batch.threshold, 2**31 - 1, 2**31 - 1, prec),
                     prec)
        torch.cuda.synchronize()
        assert one["a"].numel() == p1 - p0 and one["fp"].numel() == r1 - r0, m
        assert torch.equal(allr["a"][p0:p1] - o, one["a"]) and torch.equal(
            allr["b"][p0:p1] - o, one["b"]), m
        assert torch.equal(allr["q4"][:, 2 * p0:2 * p1], one["q4"]), m
        assert torch.equal(allr["t3"][:, 2 * p0:2 * p1], one["t3"]), m
        assert torch.equal(allr["fp"][r0:r1] - int(p0), one["fp"]), m
        assert torch.equal(allr["n3"][:, r0:r1], one["n3"]) and torch.equal(
            allr["d1"][r0:r1], one["d1"]), m
        total += p1 - p0
    assert total == r.all.pairs > 0
    single.close()


@pytest.mark.parametrize("prec", [capi.FP64, capi.FP32_GUARDED])
def test_batch_meanshift_equals_single_meshes(batch, prec):
    """Per mesh: the same seeds as the one-mesh call with seeds = min(V, ceil(
    .2 S)), and the same modes (FP64 1e-9, FP32 2e-5: a launch may serve a
    seed with one or four waves, which changes the summation order)."""
    dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
    tol = 1e-9 if prec == capi.FP64 else 2e-5
    ratio, iters = 0.2, 12
    r = batch.pairvote(prec)
    M = batch.count
    poff = batch.sess.fetch(r.pair_offsets, (M + 1,), torch.int64).cpu().numpy()
    roff = batch.sess.fetch(r.refl_offsets, (M + 1,), torch.int64).cpu().numpy()
    ms = batch.meanshift(ratio, iters)
    cr, cf = ms.all.rigid_seeds, ms.all.refl_seeds
    so_r = batch.sess.fetch(ms.rigid_seed_offsets, (M + 1,), torch.int64).cpu().numpy()
    so_f = batch.sess.fetch(ms.refl_seed_offsets, (M + 1,), torch.int64).cpu().numpy()
    ids = batch.sess.fetch(ms.all.seed_votes, (cr + cf,), torch.int32).cpu().numpy()
    mr = batch.sess.fetch(ms.all.rigid_modes, (6, cr), dt, cr).double().cpu().numpy()
    mf = batch.sess.fetch(ms.all.refl_modes, (4, cf), dt, cf).double().cpu().numpy()
    torch.cuda.synchronize()
    single = capi.Session(0)
    single.set_stream(torch.cuda.current_stream().cuda_stream)
    checked = 0
    for m, s in enumerate(batch.single):
        S = int(batch.soff[m + 1] - batch.soff[m])
        P, R = int(poff[m + 1] - poff[m]), int(roff[m + 1] - roff[m])
        V = 2 * P + R
        ns = min(V, int(np.ceil(ratio * S)))
        assert (so_r[m + 1] - so_r[m]) + (so_f[m + 1] - so_f[m]) == ns, m
        if ns == 0:
            continue
        single.dev_pairvote(s["n"], s["x"], s["y"], s["z"], s["feat"], s["S"], s["sel"],
                            s["box"], s["center"], batch.threshold, 2**31 - 1, 2**31 - 1, prec)
        one = single.dev_meanshift_votes(ns, 7, iters, 0.05, 3.0, False)
        kr, kf = one.rigid_seeds, one.refl_seeds
        oid = single.fetch(one.seed_votes, (kr + kf,), torch.int32).cpu().numpy()
        orig = single.fetch(one.rigid_modes, (6, kr), dt, kr).double().cpu().numpy()
        ofl = single.fetch(one.refl_modes, (4, kf), dt, kf).double().cpu().numpy()
        torch.cuda.synchronize()
        assert kr == so_r[m + 1] - so_r[m] and kf == so_f[m + 1] - so_f[m], m
        assert np.array_equal(ids[so_r[m]:so_r[m + 1]] - 2 * poff[m], oid[:kr]), m
        assert np.array_equal(ids[cr + so_f[m]:cr + so_f[m + 1]] - roff[m], oid[kr:]), m
        assert np.max(np.abs(mr[:, so_r[m]:so_r[m + 1]] - orig), initial=0) < tol, m
        assert np.max(np.abs(mf[:, so_f[m]:so_f[m + 1]] - ofl), initial=0) < tol, m
        checked += 1
    assert checked > 0
    single.close()


def test_batch_graph_replays_the_captured_step(batch):
    """A captured batch step replays bit for bit; a replay over changed inputs
    is reported by the check (and stays inside the captured buffers)."""
    prec = capi.FP32_GUARDED
    pv, ms = batch.capture(prec, 0.2, 12)
    M = batch.count
    first = _fetch(batch.sess, pv.all, prec)
    # the seeds written: [0, offsets[M]) of the capacity all.rigid_seeds
    cap = ms.all.rigid_seeds
    cr = int(batch.sess.fetch(ms.rigid_seed_offsets, (M + 1,), torch.int64).cpu()[M])
    assert 0 < cr <= cap
    modes = batch.sess.fetch(ms.all.rigid_modes, (6, cr), torch.float32, cap).cpu().clone()
    assert torch.isfinite(modes).all()
    poff = batch.sess.fetch(pv.pair_offsets, (M + 1,), torch.int64).cpu().clone()
    torch.cuda.synchronize()
    for it in range(3):
        batch.replay()
        batch.sess.batch_graph_check()
        again = _fetch(batch.sess, pv.all, prec)
        torch.cuda.synchronize()
        for k in first:
            assert torch.equal(first[k], again[k]), (it, k)
        got = batch.sess.fetch(ms.all.rigid_modes, (6, cr), torch.float32, cap).cpu()
        diff = (got != modes).any(0)
        assert not diff.any(), (it, int(diff.sum()), cr, torch.nonzero(diff).flatten()[:8].tolist(),
                                float((got - modes).abs().max()))
        assert torch.equal(poff, batch.sess.fetch(pv.pair_offsets, (M + 1,), torch.int64).cpu())


def test_batch_graph_replay_over_changed_inputs_is_reported(batch):
    """A replay over changed inputs is reported by the check, and the graph
    then replays the captured inputs bit for bit again.  (Its earlier fault
    came from points past the last mesh segment reading the grid past the
    last one; tests/test_replay_index.py shows that read on the CPU and that
    the sentinel key keeps every index of the set-up inside its buffer.)"""
    prec = capi.FP32_GUARDED
    pv, ms = batch.capture(prec, 0.2, 12)
    first = _fetch(batch.sess, pv.all, prec)
    torch.cuda.synchronize()
    # other curvatures on one mesh (kmin sign flipped on every other vertex:
    # the ratio test is invariant under a common scale or sign, not under
    # this): other pair counts, reported by the check
    saved = batch.feat.clone()
    lo, hi = batch.vbase[0], batch.vbase[1]
    batch.feat[9, lo:hi:2] *= -1
    torch.cuda.synchronize()
    batch.replay()
    with pytest.raises(capi.MvsError):
        batch.sess.batch_graph_check()
    batch.feat.copy_(saved)
    torch.cuda.synchronize()
    batch.replay()
    batch.sess.batch_graph_check()
    again = _fetch(batch.sess, pv.all, prec)
    torch.cuda.synchronize()
    for k in first:
        assert torch.equal(first[k], again[k]), k


def test_batch_autotune_keeps_results(batch):
    """Every candidate launch table gives the same pairs and votes; the tuned
    table stays on the session with its graph captured."""
    prec = capi.FP32_GUARDED
    batch.sess.set_tuning(0, 2048, 8192)
    eager = _fetch(batch.sess, batch.pairvote(prec).all, prec)
    torch.cuda.synchronize()
    table, best, pv, ms = batch.autotune(prec, 0.2, 8)
    assert len(table) >= 8 and 0 <= best < len(table)
    assert all(t["ms"] > 0 for t in table)
    assert min(t["ms"] for t in table) == table[best]["ms"]
    tuned = batch.sess.get_tuning()
    assert (tuned["walk_quota"], tuned["walk_blocks"], tuned["wide_seeds"]) == (
        table[best]["walk_quota"], table[best]["walk_blocks"], table[best]["wide_seeds"])
    batch.replay()
    batch.sess.batch_graph_check()
    got = _fetch(batch.sess, pv.all, prec)
    torch.cuda.synchronize()
    for k in eager:
        assert torch.equal(eager[k], got[k]), k
    # every persistent quota gives the same pairs (eager path)
    for quota in (1, 3, 16):
        batch.sess.set_tuning(quota, 64, 8192)
        again = _fetch(batch.sess, batch.pairvote(prec).all, prec)
        torch.cuda.synchronize()
        for k in eager:
            assert torch.equal(eager[k], again[k]), (quota, k)
    batch.sess.set_tuning(0, 2048, 8192)


@pytest.fixture(scope="module")
def batch512():
    """Config 4 at its size: 512 meshes, grid sides log-uniform in [32, 512],
    1-16 motif copies, samples clamp(0.05 N, 200, 20000)."""
    return workload.Batch(count=512, keep_single=True)


def test_config4_full_batch_equals_single_meshes(batch512, oracle):
    """Per mesh, over a spread of 16 meshes (the smallest and the largest
    among them): the batch's pairs, their order and the compact votes equal a
    one-mesh pass bit for bit, and its mean-shift seeds and modes equal the
    one-mesh call's; the smallest mesh's pairs equal the oracle's."""
    b = batch512
    prec = capi.FP32_GUARDED
    M = b.count
    sides = np.array([s["side"] for s in b.single])
    assert M == 512 and sides.min() >= 32 and sides.max() <= 512 and sides.max() > 4 * sides.min()
    r = b.pairvote(prec)
    allr = _fetch(b.sess, r.all, prec)
    poff = b.sess.fetch(r.pair_offsets, (M + 1,), torch.int64).cpu().numpy()
    roff = b.sess.fetch(r.refl_offsets, (M + 1,), torch.int64).cpu().numpy()
    ratio, iters = 0.2, 20
    ms = b.meanshift(ratio, iters)
    cr, cf = ms.all.rigid_seeds, ms.all.refl_seeds
    so_r = b.sess.fetch(ms.rigid_seed_offsets, (M + 1,), torch.int64).cpu().numpy()
    so_f = b.sess.fetch(ms.refl_seed_offsets, (M + 1,), torch.int64).cpu().numpy()
    ids = b.sess.fetch(ms.all.seed_votes, (cr + cf,), torch.int32).cpu().numpy()
    mr = b.sess.fetch(ms.all.rigid_modes, (6, cr), torch.float32, cr).double().cpu().numpy()
    mf = b.sess.fetch(ms.all.refl_modes, (4, cf), torch.float32, cf).double().cpu().numpy()
    torch.cuda.synchronize()
    assert poff[-1] == r.all.pairs > 0
    picks = set(np.linspace(0, M - 1, 14).astype(int).tolist())
    picks |= {int(np.argmin(sides)), int(np.argmax(sides))}
    single = capi.Session(0)
    single.set_stream(torch.cuda.current_stream().cuda_stream)
    # the batch's seed count puts every seed in the narrow regime (its own
    # grid reach and query shape); the one-mesh calls are put there too
    tune = b.sess.get_tuning()
    assert cr + cf >= tune["wide_seeds"]
    single.set_tuning(0, 2048, 0, tune["seed_lanes"])
    for m in sorted(picks):
        s = b.single[m]
        p0, p1, r0, r1 = poff[m], poff[m + 1], roff[m], roff[m + 1]
        o = int(b.soff[m])
        one = _fetch(single, single.dev_pairvote(s["n"], s["x"], s["y"], s["z"], s["feat"],
                                                s["S"], s["sel"], s["box"], s["center"],
                                                b.threshold, 2**31 - 1, 2**31 - 1, prec), prec)
        torch.cuda.synchronize()
        assert torch.equal(allr["a"][p0:p1] - o, one["a"]), m
        assert torch.equal(allr["b"][p0:p1] - o, one["b"]), m
        assert torch.equal(allr["q4"][:, 2 * p0:2 * p1], one["q4"]), m
        assert torch.equal(allr["t3"][:, 2 * p0:2 * p1], one["t3"]), m
        assert torch.equal(allr["fp"][r0:r1] - int(p0), one["fp"]), m
        assert torch.equal(allr["n3"][:, r0:r1], one["n3"]), m
        assert torch.equal(allr["d1"][r0:r1], one["d1"]), m
        S = int(b.soff[m + 1] - b.soff[m])
        ns = min(2 * int(p1 - p0) + int(r1 - r0), int(np.ceil(ratio * S)))
        if ns:
            one_ms = single.dev_meanshift_votes(ns, 7, iters, 0.05, 3.0, False)
            kr, kf = one_ms.rigid_seeds, one_ms.refl_seeds
            oid = single.fetch(one_ms.seed_votes, (kr + kf,), torch.int32).cpu().numpy()
            orig = single.fetch(one_ms.rigid_modes, (6, kr), torch.float32, kr).double().cpu().numpy()
            ofl = single.fetch(one_ms.refl_modes, (4, kf), torch.float32, kf).double().cpu().numpy()
            torch.cuda.synchronize()
            assert kr == so_r[m + 1] - so_r[m] and kf == so_f[m + 1] - so_f[m], m
            assert np.array_equal(ids[so_r[m]:so_r[m + 1]] - 2 * poff[m], oid[:kr]), m
            assert np.array_equal(ids[cr + so_f[m]:cr + so_f[m + 1]] - roff[m], oid[kr:]), m
            # same grid, seeds and query shape: the same sums in the same order
            assert np.array_equal(mr[:, so_r[m]:so_r[m + 1]], orig), m
            assert np.array_equal(mf[:, so_f[m]:so_f[m + 1]], ofl), m
    single.close()
    # the smallest mesh's pairs against the oracle (FP64 predicate)
    m = int(np.argmin(sides))
    s = b.single[m]
    feat = s["feat"].cpu().numpy()
    v = np.stack([s["x"].cpu().numpy(), s["y"].cpu().numpy(), s["z"].cpu().numpy()], axis=1)
    sel = s["sel"].cpu().numpy()
    pa, pb = oracle.pairing(v, feat[9], feat[10], sel, s["box"], b.threshold, 10**9)
    a = allr["a"][poff[m]:poff[m + 1]].numpy() - int(b.soff[m])
    bb = allr["b"][poff[m]:poff[m + 1]].numpy() - int(b.soff[m])
    assert np.array_equal(sel[a], pa) and np.array_equal(sel[bb], pb)
