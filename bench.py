"""Benchmark of the transform-space voting core on one MI355X.

Workload: the 2.1M-triangle height field (1024 x 1024 grid, 64 motif copies,
noise 1e-4), 100,000 uniform vertex samples pruned by the reference's
anisotropy test, curvature-ratio tolerance .02.  Inputs (mesh, topology,
features, selected samples) are resident in HBM before timing starts.

One step = one pass of the voting core over the selected samples:
pairing (signature sort, two-level index, walk, scan, placement) + vote
emission (mirror offsets, compact records) + batched mean-shift (20,000 seed
votes x 30 iterations, h = .05, kernel cut at 3h), in the FP32-guarded
arithmetic.  `value` is pair votes per second of the whole step; the plain
FP64 path is timed the same way and reported beside it, with per-kernel
device times (HIP events on the session stream), the dominant kernel's
roofline and the reference's own CPU code (oracle/_ref) timed on a bounded
sample on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
SEEDS, ITERS, H, CUT = 20000, 30, 0.05, 3.0
KERNELS = ("batch_seg", "batch_offsets", "batch_seeds",
           "pair_gather", "radix_hist", "radix_scatter", "scan_reduce", "scan_apply", "scan_zero",
           "pair_index", "pair_walk", "pair_place", "pair_long", "vote_count",
           "vote_emit", "zero", "ms_seeds", "ms_points", "ms_bbox", "ms_grid",
           "ms_cell_keys", "ms_cell_offsets", "ms_gather", "ms_after_sort", "ms_seed_keys",
           "ms_query", "ms_stats")
# kernels timed with HIP events inside the timed region (the roofline kernels)
TIMED = ("pair_walk", "vote_emit", "ms_query")


def load_traffic():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor), or {}."""
    path = os.path.join(ROOT, "profiles", "traffic_cfg2.json")
    try:
        with open(path) as f:
            return json.load(f).get("kernels", {})
    except (OSError, ValueError):
        return {}


def load_counters():
    """Per-kernel SQ counters of the cfg2 bench from the committed rocprofv3
    --pmc summary (profiles/counters_cfg2.json), or {}: wave-level VALU
    instructions and the fraction of the SIMDs' issue cycles they fill
    (4 cycles per wave64 VALU instruction on a 16-lane SIMD)."""
    try:
        with open(os.path.join(ROOT, "profiles", "counters_cfg2.json")) as f:
            return json.load(f).get("kernels", {})
    except (OSError, ValueError):
        return {}


def load_curvature_traffic():
    """HBM bytes per curvature call from the committed rocprofv3 --pmc
    summary of the curvature kernels, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_curvature.json")) as f:
            return json.load(f)["bytes_per_call"]
    except (OSError, ValueError, KeyError):
        return None


def curvature_bytes(N, nnz, E, smoothing, s=8):
    """Algorithmic HBM bytes of one curvature call: per smoothing pass N*9s
    written, (N + nnz)*9s read (each ring entry's tensor once: an upper
    bound, as caches serve repeats) and the ring CSR 8(N+1) + 4 nnz; the
    edge pass E*(24 + 9s); the frames N*(9s + 3s + 11s + 1)."""
    per = N * 9 * s + (N + nnz) * 9 * s + 8 * (N + 1) + 4 * nnz
    return smoothing * per + E * (24 + 9 * s) + N * (9 * s + 3 * s + 11 * s + 1)


def bench_curvature(wl, steps):
    """mvs_dev_curvature on the resident mesh and topology: wall time per
    call (synchronised), per-kernel device times, and the HBM roofline of
    the whole call."""
    import torch
    names = ("curv_sum", "curv_mean", "curv_greater", "curv_gather", "curv_smooth", "curv_frames")
    wl.curvature()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        wl.curvature()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    wl.sess.enable_timing(True, only=names)
    for _ in range(steps):
        wl.curvature()
    torch.cuda.synchronize()
    kern = {}
    for k in names:
        ms_, n = wl.sess.timing(k)
        if n:
            kern[k] = {"ms_per_call": ms_ / steps, "launches_per_call": n / steps}
    wl.sess.enable_timing(False)
    N, nnz, E = wl.n, int(wl.nbr_idx.numel()), int(wl.edges.shape[0])
    B = curvature_bytes(N, nnz, E, wl.smoothing)
    gbs = B / dt / 1e9
    return {"ms_per_call": dt * 1e3, "vertices": N, "ring_entries": nnz, "edges": E,
            "smoothing": wl.smoothing, "kernels": kern,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": load_curvature_traffic(),
                         "algorithmic_bytes": B, "kernel": "whole call (gather, 3 smoothing "
                                                           "passes, frames, edge-length sum)"}}


def bench_end_to_end(mesh, precision, runs=3):
    """Host mesh to modes on the device: device prepare, curvature, prune,
    pairs + votes, mean-shift (workload.end_to_end), median of `runs` after a
    warm-up, once more with a synchronisation after every stage for the
    split; the host prepare (mvs_prepare) is timed beside it."""
    import numpy as np
    import torch
    from mitra_algo_amd import capi, meshgen, workload
    soa = [np.ascontiguousarray(mesh.vertices[:, k]) for k in range(3)] + \
          [np.ascontiguousarray(mesh.faces[:, k].astype(np.int32)) for k in range(3)]
    samples = meshgen.uniform_samples(100000, len(mesh.vertices), 3)
    sess = capi.Session(0)
    sess.set_stream(torch.cuda.current_stream().cuda_stream)
    run = lambda sync=False: workload.end_to_end(sess, soa, samples, 0.02, precision, SEEDS,
                                                 ITERS, H, CUT, stage_sync=sync)
    run()
    totals = []
    for _ in range(runs):
        d, S, pv, ms, t = run()
        totals.append(t["total"])
    stages = run(True)[4]
    t0 = time.perf_counter()
    capi.prepare(mesh.vertices, mesh.faces)
    host_prepare = time.perf_counter() - t0
    sess.close()
    return {"ms": float(np.median(totals)) * 1e3, "runs_ms": [x * 1e3 for x in totals],
            "stages_ms": {k: v * 1e3 for k, v in stages.items()},
            "selected": S, "pairs": pv.pairs, "reflection_votes": pv.reflections,
            "host_prepare_ms": host_prepare * 1e3,
            "what": "host SoA mesh -> mvs_dev_prepare -> mvs_dev_curvature -> mvs_dev_prune "
                    "(100,000 uniform samples) -> mvs_dev_pairvote -> mvs_dev_meanshift_votes "
                    "(20,000 seeds x 30 iterations), modes on the device"}


def emission_bytes(P, R, S, s):
    """Algorithmic bytes of vote emission: pair ordinals in, compact records
    out (rigid q4+t3 at s bytes per scalar, the pair being vote / 2; mirror
    n3+d and an int32 pair id), the 14-scalar sample records read once."""
    return 8 * P + 2 * P * 7 * s + R * (4 * s + 4) + 14 * s * S


def meanshift_flops(cand, inside, D=6):
    """Algorithmic flops of the mean-shift queries: 3D per candidate visit
    (difference, square, accumulate) and 2D+4 per in-support visit (weight,
    weighted sums)."""
    return cand * 3 * D + inside * (2 * D + 4)


def meanshift_flops_kinds(ms):
    """The same count with each kind's dimension: D = 6 for rigid points,
    D = 4 for mirror points (needs a want_stats result)."""
    return (meanshift_flops(ms.rigid_candidates, ms.rigid_in_support, 6) +
            meanshift_flops(ms.refl_candidates, ms.refl_in_support, 4))


def host_threads():
    """Threads for the all-cores leg: the host's CPUs, at most 16 (the host
    share of one GPU on a shared node)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n or 1))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, runs=3):
    """Median wall time of `runs` calls after one warm-up call (the
    reference's benchmark rule, benchmarks/pairing.cpp:39-52)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2], out


def cpu_baseline(wl, ms_points=None, votes=None, budget_s=20.0):
    """The voting step on this host's CPU, on bounded samples of the same
    workload:
      * pairing + transformations over the first S_sub selected samples (S_sub
        grown until one single-thread run takes about budget/8): the
        reference's own code (oracle/_ref, single thread, as the reference
        runs) and the oracle's port on all host cores (threads over sources,
        then over pairs; outputs equal the single-thread ones); medians of 3
        runs after a warm-up; votes/s is size-independent to first order
        (pairs and tests both grow as S^2);
      * batched mean-shift (the port, brute force over every vote of the kind
        per update, as the reference's evaluator, src/cpu/clustering.cpp:
        186-208) on all cores for a sample of the seeds over the full FP64
        vote set, its seed-iteration rate extrapolated to the step's seeds x
        iterations.
    value: whole-step votes/s of the all-cores leg, V / (V / votes_rate +
    extrapolated mean-shift seconds)."""
    import numpy as np
    from oracle.pyoracle import Oracle
    port = Oracle("oracle")
    try:
        ref, ref_kind = Oracle("ref"), "reference"
    except FileNotFoundError:
        ref, ref_kind = port, "port"
    thr = host_threads()
    v = wl.mesh.vertices
    feat = wl.feat_np
    S_sub = min(wl.S, 8000)

    def pv(lib, sel, threads=1):
        kw = {"threads": threads} if threads > 1 else {}
        pa, pb = lib.pairing(v, feat[9], feat[10], sel, wl.box, wl.threshold, 2**31 - 1,
                             cap=max(1, len(sel) * (len(sel) - 1) // 2 // 40 + 1000000), **kw)
        pair, kd, lanes = lib.transformations(v, feat, pa, pb, wl.box, wl.center, 2**31 - 1,
                                              **kw)
        return len(pa), len(pair)

    while True:
        sel = wl.selected_np[:S_sub]
        t0 = time.perf_counter()
        pv(ref, sel)
        dt = time.perf_counter() - t0
        if dt > budget_s / 8 or S_sub >= wl.S:
            break
        S_sub = min(wl.S, int(S_sub * 1.5))
    t1, (P1, V1) = median_time(lambda: pv(ref, sel))
    tN, (PN, VN) = median_time(lambda: pv(port, sel, thr))
    assert (P1, V1) == (PN, VN)
    legs = {"pairvote_1thread": {"votes_per_s": V1 / t1, "seconds": t1, "kind": ref_kind,
                                 "cores": 1},
            "pairvote_all_cores": {"votes_per_s": VN / tN, "seconds": tN, "kind": "port",
                                   "cores": thr}}
    ms_s = None
    if ms_points is not None:
        # seeds: every k-th vote of each kind, 2 x `thr` seeds per kind
        seconds, per_iter = 0.0, {}
        for name, (pts, ns_full) in ms_points.items():
            n = pts.shape[1]
            k = max(1, n // (2 * thr))
            idx = np.arange(0, n, k)[:2 * thr]
            seeds = np.ascontiguousarray(pts[:, idx])
            zeros_v, zeros_s = np.zeros(n, np.uint8), np.zeros(len(idx), np.uint8)
            D = pts.shape[0]
            t, _ = median_time(lambda: port.meanshift_batched(
                pts, zeros_v, D, D, seeds, zeros_s, H, CUT, ITERS, threads=thr), runs=1)
            rate = len(idx) * ITERS / t  # seed-iterations per second
            per_iter[name] = {"seed_iterations_per_s": rate, "sample_seeds": int(len(idx)),
                              "step_seeds": int(ns_full), "votes": int(n)}
            seconds += ns_full * ITERS / rate
        ms_s = seconds
        legs["meanshift_all_cores"] = {"extrapolated_s": seconds, "kinds": per_iter,
                                       "kind": "port", "cores": thr}
    V = votes
    rate = VN / tN
    value = rate if ms_s is None or not V else V / (V / rate + ms_s)
    return {"value": value, "unit": "votes/s", "cores": thr, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"pairing + transformations over the first {S_sub} of {wl.S} selected "
                      f"samples ({P1} pairs, {V1} votes), medians of 3 runs after a warm-up; "
                      + (f"mean-shift: {2 * thr} seeds per kind x {ITERS} iterations over all "
                         f"FP64 votes, extrapolated to {SEEDS} seeds; value = whole-step "
                         f"votes/s on {thr} threads" if ms_s is not None else
                         "pairing + transformations only"),
            "legs": legs}


def bench_batch(args):
    """Config 4: 512 height-field meshes (grid side log-uniform in [32, 512],
    1-16 motif copies, per-mesh seeds), samples clamp(0.05 N, 200, 20000),
    FP32-guarded pairing + votes + 20-iteration mean-shift of every mesh as one
    captured HIP graph with the autotuned launch table; meshes per second of
    graph replays, and per-mesh pairs compared with one-mesh runs."""
    import numpy as np
    import torch
    from mitra_algo_amd import capi, workload
    t0 = time.perf_counter()
    wl = workload.Batch(count=args.meshes, keep_single=True)
    print(f"[bench] batch set-up {time.perf_counter() - t0:.1f}s meshes={wl.count} "
          f"vertices={wl.n} samples={wl.S}", file=sys.stderr, flush=True)
    prec, iters, ratio = capi.FP32_GUARDED, 20, 0.2
    # eager step (host-synchronised calls) for comparison
    for _ in range(2):
        wl.step(prec, ratio, iters)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pv_e, ms_e = wl.step(prec, ratio, iters)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t1) * 1e3
    P, R = pv_e.all.pairs, pv_e.all.reflections
    t1 = time.perf_counter()
    table, best, pv, ms = wl.autotune(prec, ratio, iters)
    tune_s = time.perf_counter() - t1
    print(f"[bench] autotune {tune_s:.1f}s best {table[best]}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        wl.replay()
    wl.sess.batch_graph_check()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        wl.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t1) / args.steps
    wl.sess.batch_graph_check()
    # untimed breakdown: one eager step with an event pair around every launch
    wl.sess.enable_timing(True)
    wl.step(prec, ratio, iters)
    torch.cuda.synchronize()
    breakdown = {}
    for k in KERNELS:
        ms_, n = wl.sess.timing(k)
        if n:
            breakdown[k] = {"ms": ms_, "launches": n}
    wl.sess.enable_timing(False)
    # per-mesh pairs against one-mesh runs (a spread of meshes)
    M = wl.count
    poff = wl.sess.fetch(pv.pair_offsets, (M + 1,), torch.int64).cpu().numpy()
    a_all = wl.sess.fetch(pv.all.a, (P,), torch.int32).cpu().numpy()
    b_all = wl.sess.fetch(pv.all.b, (P,), torch.int32).cpu().numpy()
    single = capi.Session(0)
    single.set_stream(torch.cuda.current_stream().cuda_stream)
    picks = sorted(set(np.linspace(0, M - 1, min(M, args.check_meshes)).astype(int).tolist()))
    equal = True
    for m in picks:
        s = wl.single[m]
        r = single.dev_pairvote(s["n"], s["x"], s["y"], s["z"], s["feat"], s["S"], s["sel"],
                                s["box"], s["center"], wl.threshold, 2**31 - 1, 2**31 - 1, prec)
        a = single.fetch(r.a, (r.pairs,), torch.int32).cpu().numpy()
        b = single.fetch(r.b, (r.pairs,), torch.int32).cpu().numpy()
        o = int(wl.soff[m])
        equal &= bool(np.array_equal(a_all[poff[m]:poff[m + 1]] - o, a)
                      and np.array_equal(b_all[poff[m]:poff[m + 1]] - o, b))
    single.close()
    sides = [s["side"] for s in wl.single]
    out = {
        "metric": "meshes/s: batched FP32-guarded pairing + 20-iteration mean-shift (512 meshes)",
        "value": M / dt,
        "unit": "meshes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 recheck in guard bands)",
        "data": "synthetic",
        "config": {"workload": f"{M} height fields, grid side log-uniform in [32, 512] "
                               f"(sides {min(sides)}-{max(sides)}), 1-16 motif copies, "
                               f"{wl.n} vertices, samples clamp(0.05 N, 200, 20000) -> {wl.S} "
                               f"after prune, t={wl.threshold}; mean-shift {iters} iterations, "
                               f"seeds 0.2 per sample",
                   "pairs": P, "votes": 2 * P + R, "parallelism": "single GPU"},
        "eager_step_ms": eager_ms,
        "graph_speedup_over_eager": eager_ms / (dt * 1e3),
        "autotune": {"table": table, "best": best, "seconds": tune_s},
        "breakdown_eager_step": breakdown,
        "per_mesh_pairs_bit_exact": {"meshes_checked": len(picks), "equal": equal},
    }
    print(json.dumps(out), flush=True)


def bench_refine(args):
    """Config 5: the 64 heaviest distinct FP32 modes of the 96x96 height field
    (2,000 samples, t = .05, mean-shift h = .05), 20,000 patch points each
    (the reference's verification rule at the unrefined mode), 10 damped
    Gauss-Newton steps per mode on the point-to-plane energy with closest
    points recomputed each step.  One step = the 10 iterations of all modes
    from the unrefined state; value = modes refined per second."""
    import numpy as np
    import torch
    from mitra_algo_amd import refine, workload
    npts, top, iters = 20000, 64, 10
    wl = workload.cfg1()
    modes = refine.top_modes(wl, top)
    patches = refine.Patches(wl.sess, wl.mesh.vertices, wl.mesh.faces, wl.center, wl.box)
    idx = patches.select(modes, npts)
    probs = [p for p in (patches.problem(modes, idx, refine.RIGID)[0],
                         patches.problem(modes, idx, refine.MIRROR)[0]) if p is not None]
    start = [(p.base.clone() if p.base is not None else None, p.theta.clone()) for p in probs]

    def reset():
        for p, (b, t) in zip(probs, start):
            p.base = b.clone() if b is not None else None
            p.theta = t.clone()

    def step():
        reset()
        return refine.refine_all(probs, steps=iters)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    wl.sess.enable_timing(True, only=("closest", "align", "align_sum"))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hist = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kern = {}
    for k in ("closest", "align", "align_sum"):
        ms_, n = wl.sess.timing(k)
        if n:
            kern[k] = {"ms_total_per_step": ms_ / args.steps, "ms_avg": ms_ / n,
                       "launches_per_step": n / args.steps}
    wl.sess.enable_timing(False)
    M = sum(p.modes for p in probs)
    # bytes per launch: closest reads 12 and writes 28 bytes per point (all
    # modes of a kind in one launch); align reads 36 bytes per point
    launches = 2 * (iters + 1) * len(probs)
    per_launch_pts = M * npts / len(probs)
    roofs = {}
    for k, bpp in (("closest", 40), ("align", 36)):
        if k in kern:
            b = bpp * per_launch_pts
            gbs = b / (kern[k]["ms_avg"] * 1e-3) / 1e9
            roofs[k] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "traffic": None, "algorithmic_bytes": b,
                        "ms_avg": kern[k]["ms_avg"]}
    dom = max(roofs, key=lambda k: kern[k]["ms_total_per_step"])
    roof = dict(roofs[dom], kernel=dom)
    resid = {}
    for p, h in zip(probs, hist):
        before, after = refine.rms_residual(h[0], npts), refine.rms_residual(h[-1], npts)
        resid["rigid" if p.kind == refine.RIGID else "mirror"] = {
            "modes": p.modes, "rms_before_median": float(np.median(before)),
            "rms_after_median": float(np.median(after)), "rms_before_max": float(before.max()),
            "rms_after_max": float(after.max()),
            "modes_lowered": int(np.sum(after < before))}
    out = {
        "metric": "modes/s: 10 Gauss-Newton steps of point-to-plane refinement (top-64 modes)",
        "value": M / dt, "unit": "modes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"96x96 height field (18050 triangles), 2000 samples, t=0.05; "
                               f"top {top} FP32 modes ({M} with support), {npts} patch points "
                               f"each, {iters} Gauss-Newton steps", "parallelism": "single GPU",
                   "support": [m.support for m in modes]},
        "residual": resid, "kernels": kern, "roofline": roof, "rooflines": roofs,
        "launches_per_step": launches,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_refine(probs)
    print(json.dumps(out), flush=True)


def bench_precision(args):
    """Config 3: the 96x96 height field (2,000 samples, t = .05) in both
    arithmetics.  The FP32-guarded pair and mirror sets are compared with
    the FP64 ones (they must be equal), FP64 rechecks counted, and the
    mean-shift of both paths is run from the same seed positions (every
    k-th FP64 vote, at most 2,000 per kind, 30 iterations, h = .05, cut
    3h): max / median error of the FP32 modes against the FP64 GPU modes in
    radians (rotation angle of R32^T R64; angle between mirror normals) and
    length units (translation, mirror offset)."""
    import numpy as np
    import torch
    from scipy.spatial.transform import Rotation
    from mitra_algo_amd import capi, workload
    wl = workload.cfg1()
    res = {}
    for prec in (capi.FP64, capi.FP32_GUARDED):
        dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
        for _ in range(args.warmup):
            wl.pairvote(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            pv = wl.pairvote(prec)
        torch.cuda.synchronize()
        ms_pv = (time.perf_counter() - t0) / args.steps * 1e3
        st = wl.sess.dev_pair_stats()
        P, R = pv.pairs, pv.reflections
        f = wl.sess.fetch
        pairs = (f(pv.a, (P,), torch.int32).cpu(), f(pv.b, (P,), torch.int32).cpu(),
                 f(pv.refl_pair, (R,), torch.int32).cpu())
        q4 = f(pv.rigid_q4, (4, 2 * P), dt, 2 * P)
        t3 = f(pv.rigid_t3, (3, 2 * P), dt, 2 * P)
        n3 = f(pv.refl_n3, (3, R), dt, pv.refl_stride)
        d1 = f(pv.refl_d, (R,), dt)
        rp = torch.empty((6, 2 * P), dtype=dt, device="cuda")
        fpts = torch.empty((4, R), dtype=dt, device="cuda")
        wl.sess.dev_ms_points(prec, 2 * P, q4, t3, R, n3, d1, rp, fpts)
        torch.cuda.synchronize()
        res[prec] = {"pairs": pairs, "stats": st, "ms": ms_pv, "pts": {6: rp, 4: fpts}}
    a64, a32 = res[capi.FP64], res[capi.FP32_GUARDED]
    same = all(torch.equal(x, y) for x, y in zip(a64["pairs"], a32["pairs"]))
    modes = {}
    for D, name in ((6, "rigid"), (4, "mirror")):
        p64, p32 = a64["pts"][D], a32["pts"][D]
        n = p64.shape[1]
        if n == 0:
            continue
        idx = torch.arange(0, n, max(1, n // 2000), device="cuda")[:2000]
        s64 = p64[:, idx].contiguous()
        s32 = s64.float().contiguous()
        ns = s64.shape[1]
        o64, o32 = torch.empty_like(s64), torch.empty_like(s32)
        w64 = torch.empty(ns, dtype=torch.float64, device="cuda")
        w32 = torch.empty(ns, dtype=torch.float32, device="cuda")
        wl.sess.dev_ms_batched(capi.FP64, D, n, p64, ns, s64, 30, 0.05, 3.0, o64, w64)
        wl.sess.dev_ms_batched(capi.FP32_GUARDED, D, n, p32, ns, s32, 30, 0.05, 3.0, o32, w32)
        torch.cuda.synchronize()
        x64, x32 = o64.cpu().numpy(), o32.double().cpu().numpy()
        if D == 6:
            lin = np.linalg.norm(x32[0:3] - x64[0:3], axis=0)
            R64 = Rotation.from_rotvec(x64[3:6].T)
            R32 = Rotation.from_rotvec(x32[3:6].T)
            ang = (R32.inv() * R64).magnitude()
        else:
            lin = np.abs(x32[0] - x64[0])
            u64 = x64[1:4] / np.linalg.norm(x64[1:4], axis=0)
            u32 = x32[1:4] / np.linalg.norm(x32[1:4], axis=0)
            ang = np.arccos(np.clip((u64 * u32).sum(0), -1, 1))
        # the bar: modes within 1e-5 rad and 1e-6 length units of FP64
        modes[name] = {"seeds": int(ns), "units_max": float(lin.max()),
                       "units_p99": float(np.percentile(lin, 99)),
                       "units_median": float(np.median(lin)), "rad_max": float(ang.max()),
                       "rad_p99": float(np.percentile(ang, 99)),
                       "rad_median": float(np.median(ang)),
                       "seeds_over_1e-6_units": int(np.sum(lin > 1e-6)),
                       "seeds_over_1e-5_rad": int(np.sum(ang > 1e-5))}
    out = {
        "metric": "config 3: FP32-guarded vs FP64 pair sets and mean-shift modes (96x96 mesh)",
        "value": a64["ms"] / a32["ms"], "unit": "x (FP64 pairvote ms / FP32-guarded pairvote ms)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": a32["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 recheck in guard bands) vs f64", "data": "synthetic",
        "config": {"workload": "96x96 height field (18050 triangles), 2000 uniform samples, "
                               "magnitude drop, t=0.05; mean-shift 30 iterations, h=0.05, cut 3h",
                   "pairs": len(a64["pairs"][0]), "reflections": len(a64["pairs"][2]),
                   "parallelism": "single GPU"},
        "pair_set_identical": bool(same),
        "pair_stats_fp32g": a32["stats"],
        "pairvote_ms": {"fp64": a64["ms"], "fp32g": a32["ms"]},
        "mode_error_fp32_vs_fp64": modes,
    }
    # the bar on converged modes: every 500th-of-n seed run to its fixed
    # point in both arithmetics, endpoints merged by the reference's h/2
    # rule, largest error over the modes (the GPU precision test's check)
    conv = {}
    for D, name in ((6, "rigid"), (4, "mirror")):
        p64, p32 = a64["pts"][D], a32["pts"][D]
        n = p64.shape[1]
        if n == 0:
            continue
        idx = torch.arange(0, n, max(1, n // 500), device="cuda")[:500]
        conv[name] = workload.converged_mode_errors(wl.sess, D, p64, p32, idx)
    out["converged_modes_fp32_vs_fp64"] = conv
    out["converged_modes_within_bar"] = all(
        c["modes_over_1e-6_units"] == 0 and c["modes_over_1e-5_rad"] == 0 for c in conv.values())
    print(json.dumps(out), flush=True)


def bench_verify(args):
    """Verification (Backend::verification) of 20 cluster records on the
    2.1M-triangle mesh: 16 rigid copy-to-copy motions and 4 reflections
    through the centre, verify threshold .025 box.  The surface tree is built
    once (timed apart); one step = one mvs_dev_verification call with the
    mesh, normals and rings resident in HBM: 2 x 20 x N row searches, patches,
    compaction, and the sides, patches, rms and coverage handed back on the
    host.  value = verified rows (cluster x side x vertex) per second."""
    import numpy as np
    import torch
    from mitra_algo_amd import capi, meshgen, workload
    mesh = meshgen.cfg2_mesh(args.grid)
    topo = capi.prepare(mesh.vertices, mesh.faces)
    v, faces = mesh.vertices, topo["faces"]
    nv = len(v)
    cl = workload.copy_motions(mesh)
    nc = len(cl["reflection"])
    verify, box = 0.025, topo["box_size"]
    sess = capi.Session(0)
    t0 = time.perf_counter()
    sess.verification_surface(v, faces)
    surface_s = time.perf_counter() - t0
    pts = torch.from_numpy(np.ascontiguousarray(np.concatenate([v.T, topo["normals"]]))).cuda()
    off = torch.from_numpy(topo["nbr_off"].astype(np.int64)).cuda()
    idx = torch.from_numpy(topo["nbr_idx"]).cuda()
    w = topo["weights"]
    torch.cuda.synchronize()

    def step():
        return sess.dev_verification(nv, pts, off, idx, w, cl, verify, box)

    for _ in range(args.warmup):
        step()
    names = ("verify_morton", "verify_rows", "verify_cc_init", "verify_cc_jump",
             "verify_cc_hook", "verify_cc_flatten", "verify_compact")
    sess.enable_timing(True, only=names)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kern = {}
    for k in names:
        ms_, n = sess.timing(k)
        if n:
            kern[k] = {"ms_avg": ms_ / n, "launches_per_step": n / args.steps,
                       "ms_per_step": ms_ / args.steps}
    sess.enable_timing(False)
    stats = sess.verification_stats()
    rows = 2 * nc * nv
    # the whole drop-in call (surface build, uploads, searches, results) once
    t0 = time.perf_counter()
    full = sess.verification(v, faces, topo, cl, verify, box)
    full_s = time.perf_counter() - t0
    assert all(np.array_equal(a, b) for a, b in zip(full["side_a"], r["side_a"]))
    tflops = stats["fp64_ops"] / (kern["verify_rows"]["ms_avg"] * 1e-3) / 1e12
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic_verify.json")) as f:
            t = json.load(f)["kernels"]["k_verify"]
        traffic = (2.0 * t["fetch_kib"] + t["write_kib"]) * 1024
    except (OSError, ValueError, KeyError):
        pass
    roof = {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": "verify_rows",
            "ms_avg": kern["verify_rows"]["ms_avg"], "algorithmic_ops": stats["fp64_ops"]}
    out = {
        "metric": "verified rows/s (cluster x side x vertex) of Backend::verification",
        "value": rows / dt, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"height field {args.grid}x{args.grid} ({len(faces)} triangles, "
                               f"{nv} vertices), {nc} clusters (16 copy-to-copy rigid motions, "
                               "4 reflections through the centre), verify threshold .025",
                   "parallelism": "single GPU"},
        "accepted": {"side_a": int(r["side_a"][0][-1]), "side_b": int(r["side_b"][0][-1]),
                     "patches_a": int(len(r["patches_a"][1]) - 1)},
        "search": {"triangle_tests_per_row": stats["triangle_tests"] / rows,
                   "node_visits_per_row": stats["node_visits"] / rows,
                   "fp64_ops": stats["fp64_ops"]},
        "phases_ms": {"device": stats["device_ms"], "download": stats["download_ms"],
                      "assembly": stats["assembly_ms"]},
        "surface_build_s": surface_s, "host_call_s": full_s,
        "host_call_rows_per_s": rows / full_s, "kernels": kern, "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_verify(v, faces, topo, cl, verify, box)
    print(json.dumps(out), flush=True)


def bench_fixtures(args):
    """The reference's own benchmark fixtures through the drop-in calls, timed
    as the reference times them (host inputs to host outputs, transfers
    included; median of 3 after a warm-up, benchmarks/pairing.cpp:39-52 and
    benchmarks/clustering.cpp:39-58), beside its published RTX 4070 rows and
    the reference's CPU code (oracle/_ref) on this host:
      * pairing: 10,000 selected samples in 100 curvature groups, t = 1e-6
        -> 495,000 pairs (benchmarks/pairing.cpp:12-36), FP64 and
        FP32-guarded, plus the device-resident count + fill;
      * clustering: V = 1,024 / 4,096 / 16,384 votes around 4 modes,
        h = .03, 4 clusters, 100 iterations, rng seed 42
        (benchmarks/clustering.cpp:17-37)."""
    import numpy as np
    import torch
    from mitra_algo_amd import capi
    from tests import fixtures as fx
    try:
        from oracle.pyoracle import Oracle
        ref = Oracle("ref")
    except (ImportError, FileNotFoundError, OSError):
        ref = None
    sess = capi.Session(0)

    def med(fn, runs=3):
        fn()
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3, out

    # ------------------------------------------------------------- pairing
    b = fx.pairing_benchmark()
    pairing = {"workload": "benchmarks/pairing.cpp:12-36 (10,000 selected, 100 groups, "
                           "t=1e-6, max_pairs 600,000)",
               "published": {"cpu_median_ms": 93.54, "cuda_median_ms": 11.47,
                             "hardware": "RTX 4070 Laptop + host (rtx4070_wsl_pairing.txt)"}}
    want = None
    if ref is not None:
        ms_, want = med(lambda: ref.pairing(b.vertices, b.kmin, b.kmax, b.selected, b.box, 1e-6,
                                            600000))
        pairing["reference_cpu_here_ms"] = ms_
    for name, prec in (("fp64", capi.FP64), ("fp32g", capi.FP32_GUARDED)):
        ms_, got = med(lambda: sess.pairing(b.vertices, b.feat, b.selected, b.box, 1e-6, 600000,
                                            prec))
        entry = {"median_ms": ms_, "pairs": int(len(got[0])),
                 "pairs_per_s": len(got[0]) / (ms_ * 1e-3)}
        if want is not None:
            entry["equal_to_reference"] = bool(np.array_equal(got[0], want[0]) and
                                               np.array_equal(got[1], want[1]))
        pairing["drop_in_" + name] = entry
    # device-resident count + fill (inputs already in HBM)
    dev = torch.device("cuda:0")
    sess.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    v = torch.as_tensor(np.ascontiguousarray(b.vertices.T)).to(dev)
    feat = torch.as_tensor(np.ascontiguousarray(b.feat)).to(dev)
    sel = torch.as_tensor(b.selected).to(dev)
    n = len(b.vertices)

    def resident():
        P = sess.dev_pair_count(n, v[0], v[1], v[2], feat, len(b.selected), sel, b.box,
                                [0.0, 0.0, 0.0], 1e-6, 600000, capi.FP64)
        a_ = torch.empty(P, dtype=torch.int32, device=dev)
        b_ = torch.empty(P, dtype=torch.int32, device=dev)
        sess.dev_pair_fill(a_, b_, True)
        torch.cuda.synchronize()
        return P
    ms_, P = med(resident)
    pairing["device_resident_fp64"] = {"median_ms": ms_, "pairs": int(P)}
    # ---------------------------------------------------------- clustering
    clustering = {"workload": "benchmarks/clustering.cpp:17-37 (alternating rigid/reflection "
                              "votes, 4 modes, h=.03, box 10, 4 clusters, rng 42)",
                  "published": {"votes": [1024, 4096, 16384],
                                "cpu_median_ms": [0.121, 0.381, 2.033],
                                "cuda_median_ms": [5.459, 10.383, 32.670],
                                "hardware": "RTX 4070 Laptop + host "
                                            "(rtx4070_wsl_clustering.json)"},
                  "rows": []}
    for count in (1024, 4096, 16384):
        kind, desc, rot = fx.cluster_votes(count)
        row = {"votes": count}
        ms_, got = med(lambda: sess.clustering(kind, desc, rot, 10.0, [0, 0, 0], 0.03, 100, 4,
                                               capi.rng(42)))
        row.update(hip_median_ms=ms_, clusters=int(len(got["reflection"])))
        if ref is not None:
            ms_r, want_c = med(lambda: ref.clustering(kind, desc, rot, 10.0, [0, 0, 0], 0.03, 100,
                                                      4, ref.rng(42)))
            row["reference_cpu_here_ms"] = ms_r
            row["members_equal_to_reference"] = bool(
                np.array_equal(got["members"], want_c.members) and
                np.array_equal(got["member_off"], want_c.member_off))
        clustering["rows"].append(row)
    sess.close()
    out = {
        "metric": "reference benchmark fixtures: pairing median ms (495,000 pairs) and "
                  "clustering median ms (V = 1,024 / 4,096 / 16,384)",
        "value": pairing["drop_in_fp64"]["median_ms"], "unit": "ms (pairing, drop-in FP64)",
        "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": pairing["drop_in_fp64"]["median_ms"], "higher_is_better": False,
        "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "the reference's benchmarks/pairing.cpp and "
                               "benchmarks/clustering.cpp inputs", "parallelism": "single GPU"},
        "pairing": pairing, "clustering": clustering,
        "cpu_model": cpu_model(), "host_threads": host_threads(),
    }
    print(json.dumps(out), flush=True)


def cpu_baseline_verify(v, faces, topo, cl, verify, box):
    """The reference's CPU verification (oracle/_ref, single thread) of the
    first cluster on the full mesh, its own tree build included."""
    import numpy as np
    from oracle.pyoracle import Oracle
    try:
        lib, kind = Oracle("ref"), "reference"
    except FileNotFoundError:
        return None
    T = type("T", (), {})()
    T.normals, T.weights = topo["normals"], topo["weights"]
    T.nbr_off, T.nbr_idx = topo["nbr_off"], topo["nbr_idx"]
    one = {k: np.asarray(cl[k])[:1] for k in cl}
    t0 = time.perf_counter()
    lib.verification(v, faces, T, one, verify, box)
    dt = time.perf_counter() - t0
    return {"value": 2 * len(v) / dt, "unit": "rows/s", "cores": 1, "kind": kind,
            "sample": f"Backend::verification of 1 of {len(cl['reflection'])} clusters on the "
                      f"full mesh ({2 * len(v)} rows, tree build included) in {dt:.2f} s, "
                      "single thread"}


def cpu_baseline_refine(probs):
    """The FP64 CPU reference of one energy + gradient evaluation of every
    mode (closest points of the last GPU step held fixed), one thread; the
    GPU's FP32 gradients at the same inputs are compared with it."""
    import numpy as np
    from threadpoolctl import threadpool_limits
    from oracle import refine_ref as rr
    evals, worst, t = 0, 0.0, 0.0
    with threadpool_limits(1):
        for p in probs:
            E, g, _ = p.evaluate(p.theta, want_hessian=False)
            g = g.double().cpu().numpy()
            th = p.theta.double().cpu().numpy()
            for j in range(p.modes):
                P, Q, N = (x[j].double().cpu().numpy().T for x in (p.pts, p.q, p.nrm))
                R0 = t0 = None
                if p.base is not None:
                    b = p.base[j].double().cpu().numpy()
                    R0, t0 = b[:9].reshape(3, 3), b[9:]
                a = time.perf_counter()
                rr.energy(p.kind, th[j], P, Q, N, R0, t0)
                g64 = rr.gradient(p.kind, th[j], P, Q, N, R0, t0)
                t += time.perf_counter() - a
                evals += 1
                if np.linalg.norm(g64) >= 1.0:
                    worst = max(worst, float(np.linalg.norm(g[j] - g64) / np.linalg.norm(g64)))
    return {"value": evals / t, "unit": "mode evaluations/s (energy + gradient, 20,000 points)",
            "cores": 1, "kind": "port",
            "sample": f"{evals} modes at their refined parameters, FP64 numpy",
            "gpu_fp32_gradient_rel_err_max": worst}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=("cfg2", "batch", "refine", "precision", "verify",
                                           "fixtures"),
                    default="cfg2",
                    help="cfg2: the 2.1M-triangle mesh (default); batch: 512 meshes; "
                         "refine: Gauss-Newton refinement of the top-64 modes of the 96x96 mesh; "
                         "precision: FP32-guarded against FP64 on the 96x96 mesh; "
                         "verify: verification of 20 clusters on the 2.1M-triangle mesh; "
                         "fixtures: the reference's pairing and clustering benchmark inputs")
    ap.add_argument("--meshes", type=int, default=512)
    ap.add_argument("--check-meshes", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB copy kernel (torch.neg) three times, the "
                         "reference for rocprofv3 FETCH_SIZE / WRITE_SIZE scaling")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the voting core runs on a single MI355X "
              "(the multi-GPU sharding is a design note, not run)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "refine":
        bench_refine(args)
        return
    if args.workload == "precision":
        bench_precision(args)
        return
    if args.workload == "verify":
        bench_verify(args)
        return
    if args.workload == "fixtures":
        bench_fixtures(args)
        return
    import torch
    from mitra_algo_amd import capi, workload

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    t_setup = time.perf_counter()
    wl = workload.cfg2(n=args.grid)
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s S={wl.S}", file=sys.stderr,
          flush=True)

    def run(precision, steps, warmup):
        for _ in range(warmup):
            wl.step(precision, SEEDS, ITERS)
        torch.cuda.synchronize()
        # timed region: K full steps back to back, a device synchronisation
        # on either side; the roofline kernels carry HIP events
        wl.sess.enable_timing(True, only=TIMED)
        t0 = time.perf_counter()
        for _ in range(steps):
            pv = wl.pairvote(precision)
            ms = wl.meanshift(SEEDS, ITERS)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        kernels = {}
        for k in TIMED:
            ms_, n = wl.sess.timing(k)
            if n:
                kernels[k] = {"ms_total_per_step": ms_ / steps, "ms_avg": ms_ / n,
                              "launches_per_step": n / steps}
        wl.sess.enable_timing(False)
        # the split into the two calls, untimed: a synchronisation after each
        t_pv = t_ms = 0.0
        ns = max(3, steps // 4)
        for _ in range(ns):
            a = time.perf_counter()
            pv = wl.pairvote(precision)
            b = time.perf_counter()
            ms = wl.meanshift(SEEDS, ITERS)
            torch.cuda.synchronize()
            c = time.perf_counter()
            t_pv += (b - a) / ns
            t_ms += (c - b) / ns
        # untimed breakdown of every kernel (events around each launch)
        wl.sess.enable_timing(True)
        nb = 2
        for _ in range(nb):
            wl.step(precision, SEEDS, ITERS)
        breakdown = {}
        for k in KERNELS:
            ms_, n = wl.sess.timing(k)
            if n:
                breakdown[k] = {"ms_per_step": ms_ / nb, "ms_avg": ms_ / n,
                                "launches_per_step": n / nb}
        wl.sess.enable_timing(False)
        stats = wl.meanshift(SEEDS, ITERS, want_stats=True)
        return pv, stats, dt, t_pv, t_ms, kernels, breakdown

    pv32, ms32, dt32, pv32_s, ms32_s, k32, b32 = run(capi.FP32_GUARDED, args.steps, args.warmup)
    pair32 = (wl.sess.fetch(pv32.a, (pv32.pairs,), torch.int32).cpu(),
              wl.sess.fetch(pv32.b, (pv32.pairs,), torch.int32).cpu(),
              wl.sess.fetch(pv32.refl_pair, (pv32.reflections,), torch.int32).cpu())
    st32 = wl.sess.dev_pair_stats()
    pv64, ms64, dt64, pv64_s, ms64_s, k64, b64 = run(capi.FP64, max(3, args.steps // 2), 1)
    pair64 = (wl.sess.fetch(pv64.a, (pv64.pairs,), torch.int32).cpu(),
              wl.sess.fetch(pv64.b, (pv64.pairs,), torch.int32).cpu(),
              wl.sess.fetch(pv64.refl_pair, (pv64.reflections,), torch.int32).cpu())
    same = (pv32.pairs == pv64.pairs and pv32.reflections == pv64.reflections
            and all(torch.equal(x, y) for x, y in zip(pair32, pair64)))
    P, R = pv32.pairs, pv32.reflections
    V = 2 * P + R
    print(f"[bench] fp32g {dt32 * 1e3:.3f} ms (pairvote {pv32_s * 1e3:.3f}, meanshift "
          f"{ms32_s * 1e3:.3f}) fp64 {dt64 * 1e3:.3f} ms (pairvote {pv64_s * 1e3:.3f}, "
          f"meanshift {ms64_s * 1e3:.3f}) P={P} R={R} pairs_equal={same}", file=sys.stderr,
          flush=True)

    emit_ms = k32["vote_emit"]["ms_avg"]
    emit_gbs = emission_bytes(P, R, wl.S, 4) / (emit_ms * 1e-3) / 1e9
    q_ms = k32["ms_query"]["ms_total_per_step"]
    q_flops = meanshift_flops_kinds(ms32)
    q_tflops = q_flops / (q_ms * 1e-3) / 1e12
    q64_ms = k64["ms_query"]["ms_total_per_step"]
    q64_tflops = meanshift_flops_kinds(ms64) / (q64_ms * 1e-3) / 1e12
    traffic = load_traffic()
    counters = load_counters()

    def valu(*names):
        return {n: {k: counters[n][k] for k in ("valu_insts", "valu_issue_frac", "us")
                    if k in counters[n]} for n in names if n in counters}
    rooflines = {
        "vote_emit": {"bound": "hbm", "achieved": emit_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "valu_counters": valu("k_vote_emit<true, float>"),
                      "frac": emit_gbs / HBM_PEAK_GBS,
                      "traffic": traffic.get("vote_emit", {}).get("bytes"),
                      "algorithmic_bytes": emission_bytes(P, R, wl.S, 4),
                      "ms_avg": emit_ms},
        "ms_query": {"bound": "valu",
                     "pipe": "VALU FP32, issue-bound (packed FP32 off); MFMA not used: "
                             "the support balls of a seed and its 15 nearest seeds cover "
                             "2.1-2.9x (mirror) / 14-15x (rigid) the points of one ball",
                     "valu_counters": valu("k_ms_query<float, 4, 4>",
                                           "k_ms_query_small<float, 6, 32, false>"),
                     "achieved": q_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": q_tflops / FP32_PEAK_TFLOPS,
                     "traffic": traffic.get("ms_query", {}).get("bytes"),
                     "algorithmic_flops": q_flops,
                     "ms_per_step": q_ms},
    }
    dominant = max(("pair_walk", "vote_emit", "ms_query"),
                   key=lambda k: k32.get(k, {}).get("ms_total_per_step", 0))
    roof = dict(rooflines.get(dominant, rooflines["vote_emit"]))
    roof["kernel"] = dominant if dominant in rooflines else "vote_emit"
    out = {
        "metric": "pair votes/s + mean-shift s on 2.1M-tri mesh: FP32-guarded vs FP64 GPU vs CPU",
        "value": V / dt32,
        "unit": "votes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP64 recheck in guard bands)",
        "data": "synthetic",
        "config": {"workload": f"height field {args.grid}x{args.grid} "
                               f"({2 * (args.grid - 1) ** 2} triangles), 64 motif copies, "
                               f"100000 uniform samples -> {wl.S} after prune, t=0.02; "
                               f"mean-shift {SEEDS} seeds x {ITERS} iterations, h={H}",
                   "pairs": P, "votes": V, "reflection_votes": R,
                   "parallelism": "single GPU"},
        "pairvote_votes_per_s": V / pv32_s,
        "pairvote_ms": pv32_s * 1e3,
        "meanshift_s": ms32_s,
        "meanshift_iterations_per_s": ms32.updates / ms32_s,
        "meanshift_updates": ms32.updates,
        "meanshift_visits": {"rigid_candidates": ms32.rigid_candidates,
                             "rigid_in_support": ms32.rigid_in_support,
                             "mirror_candidates": ms32.refl_candidates,
                             "mirror_in_support": ms32.refl_in_support},
        "fp64": {"value": V / dt64, "ms_per_step": dt64 * 1e3, "pairvote_ms": pv64_s * 1e3,
                 "meanshift_s": ms64_s, "ms_query_tflops": q64_tflops,
                 "speedup_fp32g_over_fp64": dt64 / dt32, "pair_set_identical": bool(same)},
        "pair_stats_fp32g": st32,
        "kernels_fp32g": k32,
        "kernels_fp64": k64,
        "breakdown_fp32g": b32,
        "breakdown_fp64": b64,
        "roofline": roof,
        "rooflines": rooflines,
    }
    out["curvature"] = bench_curvature(wl, max(3, args.steps // 2))
    e2e = bench_end_to_end(wl.mesh, capi.FP32_GUARDED)
    e2e["pairs_equal_resident_step"] = (e2e["pairs"] == P and e2e["reflection_votes"] == R)
    out["end_to_end_fp32g"] = e2e
    print(f"[bench] curvature {out['curvature']['ms_per_call']:.3f} ms, end-to-end "
          f"{e2e['ms']:.2f} ms (host prepare alone {e2e['host_prepare_ms']:.1f} ms)",
          file=sys.stderr, flush=True)
    if not args.no_cpu:
        # the FP64 mean-shift points of the step (the CPU leg's vote set)
        f = wl.sess.fetch
        q4 = f(pv64.rigid_q4, (4, 2 * P), torch.float64, 2 * P)
        t3 = f(pv64.rigid_t3, (3, 2 * P), torch.float64, 2 * P)
        n3 = f(pv64.refl_n3, (3, R), torch.float64, pv64.refl_stride)
        d1 = f(pv64.refl_d, (R,), torch.float64)
        rp = torch.empty((6, 2 * P), dtype=torch.float64, device="cuda")
        fp_ = torch.empty((4, R), dtype=torch.float64, device="cuda")
        wl.sess.dev_ms_points(capi.FP64, 2 * P, q4, t3, R, n3, d1, rp, fp_)
        torch.cuda.synchronize()
        out["cpu_baseline"] = cpu_baseline(
            wl, {"rigid": (rp.cpu().numpy(), ms32.rigid_seeds),
                 "mirror": (fp_.cpu().numpy(), ms32.refl_seeds)}, votes=V)
    # the same FP32-guarded step captured as one HIP graph and replayed over
    # the resident inputs (a serving loop's view); `value` stays the eager step
    gpv, _ = wl.capture(capi.FP32_GUARDED, SEEDS, ITERS)
    for _ in range(args.warmup):
        wl.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        wl.replay()
    torch.cuda.synchronize()
    dtg = (time.perf_counter() - t0) / args.steps
    wl.sess.graph_check()
    out["graph_replay_fp32g"] = {"ms_per_step": dtg * 1e3, "votes_per_s": V / dtg,
                                 "speedup_over_eager": dt32 / dtg,
                                 "same_vote_count": gpv.pairs == P and gpv.reflections == R}
    print(f"[bench] graph replay {dtg * 1e3:.3f} ms per step", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
