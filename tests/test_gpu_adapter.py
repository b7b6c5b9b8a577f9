"""The reference's own pipeline (detect(), src/core/pipeline.cpp:5-52) driven
through the C++ HipBackend adapter, compared stage by stage with the
reference's CPU backend by oracle/_ref/adapter_check.

Bars (the reference's own test bars): features 1e-10, sample / pruned / pairs
/ vote order / cluster members exact, votes 2e-12 relative, cluster records
1e-10, identical mt19937_64 state after select and clustering.  The example
counts are the reference-held end-to-end results (benchmarks/results/
rtx4070_wsl_cli.json) at sample ratio 1, pair threshold .08, seed 42.
"""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXPECTED = {"reflection": (1898, 4452), "rigid": (1898, 4328), "partial": (2378, 5630),
            "single": (458, 1122)}


def _run(*args):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/adapter_check is built only where the reference sources are")
    print("adapter_check", *args, flush=True)
    p = subprocess.run([CHECK, *args], capture_output=True, text=True, timeout=300)
    out = json.loads(p.stdout.strip().splitlines()[-1])
    assert p.returncode == 0 and out["ok"], out
    return out


@pytest.mark.parametrize("name", sorted(EXPECTED))
def test_detect_through_adapter_matches_cpu(name):
    out = _run(os.path.join(GOLDEN, name + ".off"), "0.08", "1.0", "42")
    assert (out["pairs"], out["votes"]) == EXPECTED[name]
    assert out["clusters"] == 20
    assert out["verification_equal"] and out["side_a"] > 0
    stages = {m["stage"] for m in out["metrics"]}
    assert {"curvature", "sample_prune", "pairing", "transformation", "mean_shift",
            "verification"} <= stages
    # measured profile (src/cuda/profile.cpp:61-93, tests/cuda_profile.cpp:36-42):
    # every stage copies both ways, holds device buffers and releases them all
    for m in out["metrics"]:
        assert m["released"] and m["peak_buffer_bytes"] > 0, m
        assert m["h2d_bytes"] > 0 and m["d2h_bytes"] > 0 and m["h2d_ms"] >= 0, m
    # exact bytes of the select stage: kmin, kmax (FP64) and valid per
    # vertex plus the sample ids in; the range flag, the kept count and the
    # kept ids out
    sp = next(m for m in out["metrics"] if m["stage"] == "sample_prune")
    assert sp["h2d_bytes"] == 17 * out["vertices"] + 4 * out["samples"], (sp, out["vertices"])
    assert sp["d2h_bytes"] == 4 + 8 + 4 * out["pruned"], sp


def test_detect_through_adapter_default_options():
    # reference defaults: sample ratio .5, pair threshold .15, seed 42
    out = _run(os.path.join(GOLDEN, "reflection.off"), "0.15", "0.5", "42")
    assert out["pairs"] > 0 and out["votes"] >= 2 * out["pairs"]
