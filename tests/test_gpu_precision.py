"""Config 3 on the GPU: the 96x96 height field (2,000 samples, t = .05) in
both arithmetics through the C-ABI.  The FP32-guarded pair and mirror sets
equal the FP64 ones, and the FP32 mean-shift modes equal the FP64 GPU modes
within 1e-6 length units and 1e-5 rad: every seed (every k-th FP64 vote) runs
to its fixed point in both arithmetics, the endpoints are merged into modes
by the reference's h/2 rule, and the bar applies to the largest error over
the modes."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi, workload  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def both():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    wl = workload.cfg1()
    out = {}
    for prec in (capi.FP64, capi.FP32_GUARDED):
        dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
        pv = wl.pairvote(prec)
        P, R = pv.pairs, pv.reflections
        f = wl.sess.fetch
        pairs = (f(pv.a, (P,), torch.int32).cpu(), f(pv.b, (P,), torch.int32).cpu(),
                 f(pv.refl_pair, (R,), torch.int32).cpu())
        q4 = f(pv.rigid_q4, (4, 2 * P), dt, 2 * P)
        t3 = f(pv.rigid_t3, (3, 2 * P), dt, 2 * P)
        n3 = f(pv.refl_n3, (3, R), dt, pv.refl_stride)
        d1 = f(pv.refl_d, (R,), dt)
        rp = torch.empty((6, 2 * P), dtype=dt, device="cuda")
        fp = torch.empty((4, R), dtype=dt, device="cuda")
        wl.sess.dev_ms_points(prec, 2 * P, q4, t3, R, n3, d1, rp, fp)
        torch.cuda.synchronize()
        out[prec] = (pairs, {6: rp, 4: fp}, wl.sess.dev_pair_stats())
    return wl, out


def test_pair_and_mirror_sets_identical(both):
    wl, out = both
    p64, p32 = out[capi.FP64][0], out[capi.FP32_GUARDED][0]
    assert len(p64[0]) > 1000 and len(p64[2]) > 100
    for x, y in zip(p64, p32):
        assert torch.equal(x, y)


@pytest.mark.parametrize("D", [6, 4])
def test_converged_modes_within_the_bar(both, D):
    wl, out = both
    p64, p32 = out[capi.FP64][1][D], out[capi.FP32_GUARDED][1][D]
    n = p64.shape[1]
    idx = torch.arange(0, n, max(1, n // 500), device="cuda")[:500]
    r = workload.converged_mode_errors(wl.sess, D, p64, p32, idx)
    print(D, r)
    # nearly every seed reaches its fixed point within the update cap
    assert r["converged_fp64"] >= 0.98 * r["seeds"] and r["converged_fp32"] >= 0.98 * r["seeds"]
    assert r["modes"] >= 3
    # the bar: the largest error over the modes
    assert r["units_max"] <= 1e-6, r
    assert r["rad_max"] <= 1e-5, r
