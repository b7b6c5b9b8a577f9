"""The one-mesh step (pairs + votes, then mean-shift over them) captured as a
HIP graph: the recorded run and every replay give the eager step's pairs,
votes, seeds and modes bit for bit, in both arithmetics.  Replays alternate
with eager work of the program and host synchronisations, and each replay is
checked on its own; mvs_graph_check finds no count that differs from the
captured ones."""
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi, workload  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cfg1():
    return workload.cfg1()


def _grab(wl, pv, ms, prec):
    f = wl.sess.fetch
    dt = torch.float32 if prec == capi.FP32_GUARDED else torch.float64
    P, R = pv.pairs, pv.reflections
    out = [f(pv.a, (P,), torch.int32), f(pv.b, (P,), torch.int32),
           f(pv.rigid_q4, (4 * 2 * P,), dt), f(pv.rigid_t3, (3 * 2 * P,), dt),
           f(pv.refl_d, (R,), dt), f(pv.refl_pair, (R,), torch.int32),
           f(ms.seed_votes, (ms.rigid_seeds + ms.refl_seeds,), torch.int32),
           f(ms.rigid_modes, (6 * ms.rigid_seeds,), dt), f(ms.rigid_weight, (ms.rigid_seeds,), dt),
           f(ms.refl_modes, (4 * ms.refl_seeds,), dt), f(ms.refl_weight, (ms.refl_seeds,), dt)]
    torch.cuda.synchronize()
    return [x.cpu() for x in out]


@pytest.mark.parametrize("prec", [capi.FP32_GUARDED, capi.FP64])
def test_step_graph_replays_equal_the_eager_step(cfg1, prec):
    wl = cfg1
    wl.threshold = 0.05
    seeds, iters = 600, 15
    pv, ms = wl.step(prec, seeds, iters)
    want = _grab(wl, pv, ms, prec)
    assert pv.pairs > 1000 and ms.refl_seeds > 0
    pv, ms = wl.capture(prec, seeds, iters)
    for x, y in zip(want, _grab(wl, pv, ms, prec)):
        assert torch.equal(x, y)
    scratch = torch.zeros(1 << 20, device="cuda:0")
    for r in range(4):
        scratch.add_(1.0)  # eager work of the program between the replays
        wl.replay()
        wl.sess.graph_check()
        got = _grab(wl, pv, ms, prec)
        for k, (x, y) in enumerate(zip(want, got)):
            assert torch.equal(x, y), (r, k)
    assert float(scratch[0]) == 4.0
