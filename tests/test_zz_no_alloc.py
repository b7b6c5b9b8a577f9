"""The allocation-free batched mean-shift (mvs_dev_ms_batched_noalloc)
through the C-ABI: on a fresh session it allocates no device memory (the
library's own count of live device bytes does not move), and its modes and
weights equal mvs_dev_ms_batched's bit for bit (same kernels over the same
grid, in both arithmetics, both query shapes), on two calls in a row, the
second with a smaller input; sizes beyond its limits are refused."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi  # noqa: E402

pytestmark = pytest.mark.gpu


def _clusters(D, seed, n, ns):
    g = np.random.default_rng(seed)
    centers = g.uniform(-1, 1, (4, D))
    per = (n - n // 5) // 4
    pts = [c + g.normal(0, 0.03, (per, D)) for c in centers]
    pts.append(g.uniform(-1.2, 1.2, (n - 4 * per, D)))
    P = np.concatenate(pts).T.copy()
    seeds = P[:, g.choice(P.shape[1], ns, replace=False)].copy()
    return P, seeds


@pytest.mark.parametrize("lanes", [256, 32])
def test_no_alloc_meanshift_equals_the_allocating_call(lanes):
    s = capi.Session(0)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    s.set_tuning(0, 2048, 1 << 40 if lanes == 256 else 0, 64 if lanes == 256 else lanes)
    live0 = capi.thread_metrics()["live_bytes"]
    cases = [(20000, 300), (5000, 100)]  # the second call smaller than the first
    for D in (6, 4):
        for prec, dt in ((capi.FP32_GUARDED, torch.float32), (capi.FP64, torch.float64)):
            for k, (n, ns) in enumerate(cases):
                print("no-alloc", lanes, D, prec, n, ns, flush=True)
                P, seeds = _clusters(D, 10 * D + k, n, ns)
                pts = torch.as_tensor(P, dtype=dt, device=dev).contiguous()
                sd = torch.as_tensor(seeds, dtype=dt, device=dev).contiguous()
                out_na, w_na = torch.empty_like(sd), torch.empty(ns, dtype=dt, device=dev)
                out, w = torch.empty_like(sd), torch.empty(ns, dtype=dt, device=dev)
                torch.cuda.synchronize()
                s.dev_ms_batched_noalloc(prec, D, n, pts, ns, sd, 20, 0.05, 3.0, out_na, w_na)
                torch.cuda.synchronize()
                assert capi.thread_metrics()["live_bytes"] == live0
                s.dev_ms_batched(prec, D, n, pts, ns, sd, 20, 0.05, 3.0, out, w)
                torch.cuda.synchronize()
                assert torch.equal(out_na, out), (D, prec, n)
                assert torch.equal(w_na, w), (D, prec, n)
                assert torch.isfinite(out).all() and (w > 0).any()
                live0 = capi.thread_metrics()["live_bytes"]  # the allocating call's slots
    s.close()


def test_no_alloc_meanshift_refuses_sizes_beyond_its_limits():
    s = capi.Session(0)
    dev = torch.device("cuda:0")
    s.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    n = 65536 + 1
    pts = torch.zeros((4, n), dtype=torch.float32, device=dev)
    sd = torch.zeros((4, 2), dtype=torch.float32, device=dev)
    out, w = torch.empty_like(sd), torch.empty(2, dtype=torch.float32, device=dev)
    with pytest.raises(capi.MvsInvalidArgument):
        s.dev_ms_batched_noalloc(capi.FP32_GUARDED, 4, n, pts, 2, sd, 5, 0.05, 3.0, out, w)
    ns = 8192 + 1
    sd = torch.zeros((4, ns), dtype=torch.float32, device=dev)
    out, w = torch.empty_like(sd), torch.empty(ns, dtype=torch.float32, device=dev)
    with pytest.raises(capi.MvsInvalidArgument):
        s.dev_ms_batched_noalloc(capi.FP32_GUARDED, 4, 100, pts, ns, sd, 5, 0.05, 3.0, out, w)
    s.close()
