"""CPU-side checks of libmvs: it loads, exports every entry point declared in
include/mvs.h, and its host stages (prepare, sampling) equal the oracle."""
import os
import re

import numpy as np
import pytest

from mitra_algo_amd import capi, meshgen
from oracle.pyoracle import read_off

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mvs.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|uint64_t|const char \*)\s*(mvs_\w+)\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(capi.EXPORTS) <= set(names)


@pytest.mark.parametrize("name", ["reflection", "rigid", "partial", "single"])
def test_prepare_matches_oracle_examples(oracle, golden_dir, name):
    v, f = read_off(os.path.join(golden_dir, f"{name}.off"))
    _compare_prepare(oracle, v, f)


def test_prepare_matches_oracle_heightfield(oracle):
    m = meshgen.cfg1_mesh()
    _compare_prepare(oracle, m.vertices, m.faces)


def _compare_prepare(oracle, v, f):
    a = capi.prepare(v, f)
    b = oracle.prepare(v, f)
    for k in ("faces", "normals", "face_normals", "weights", "nbr_off", "nbr_idx", "inc_off",
              "inc_idx", "edges"):
        assert np.array_equal(a[k], getattr(b, k)), k
    assert a["box_size"] == b.box_size
    assert np.array_equal(a["center"], b.center)
    assert a["rejected_faces"] == b.rejected_faces


def test_prepare_errors():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    with pytest.raises(capi.MvsError, match="empty"):
        capi.prepare(v, np.zeros((0, 3), np.int32))
    with pytest.raises(capi.MvsError, match="out of range"):
        capi.prepare(v, np.array([[0, 1, 9]], np.int32))
    with pytest.raises(capi.MvsError, match="winding"):
        capi.prepare(v, np.array([[0, 1, 2], [0, 1, 3]], np.int32))
    with pytest.raises(capi.MvsError, match="manifold"):
        capi.prepare(np.vstack([v, [[1, 1, 1]]]),
                     np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.int32))
    bad = v.copy()
    bad[1, 0] = np.nan
    with pytest.raises(capi.MvsError, match="Non-finite"):
        capi.prepare(bad, np.array([[0, 1, 2]], np.int32))
    # degenerate face rejected, the rest kept
    t = capi.prepare(v, np.array([[0, 1, 2], [0, 1, 1]], np.int32))
    assert t["rejected_faces"] == 1 and len(t["faces"]) == 1


def test_sampling_matches_oracle(oracle):
    from tests import fixtures as fx
    w, kmin, kmax, valid = fx.select_fixture(1028)
    for ratio in (1.0, 0.5, 0.001):
        g1, g2 = capi.rng(42), oracle.rng(42)
        a = capi.sample_vertices(w, ratio, g1)
        b = oracle.sample_vertices(w, ratio, g2)
        assert np.array_equal(a, b)
        assert bytes(g1) == bytes(g2)
    g = capi.rng(5489)
    for _ in range(9999):
        capi.mt_next(g)
    assert capi.mt_next(g) == 9981545732273789042


def test_adapter_check_usage():
    """The reference-driven adapter check builds where the reference is and
    rejects a missing argument without touching a GPU."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
    if not os.path.exists(exe):
        pytest.skip("built only where the reference sources are")
    assert subprocess.run([exe], capture_output=True).returncode == 2


def test_batch_mesh_specs():
    """Batch meshes: sides log-uniform in [32, 512], 1-16 copies inside the
    domain, deterministic per (seed, index); samples clamp(0.05 N, 200, 20000)."""
    sides = [meshgen.batch_spec(m).side for m in range(512)]
    assert min(sides) >= 32 and max(sides) <= 512
    assert np.median(np.log(sides)) == pytest.approx(np.log(128), abs=0.25)
    for m in range(20):
        a, b = meshgen.batch_spec(m, 4), meshgen.batch_spec(m, 4)
        assert a == b and 1 <= len(a.copies) <= 16
        assert all(abs(c.cx) < a.half and abs(c.cy) < a.half for c in a.copies)
    assert len(meshgen.batch_samples(32 * 32, 0)) == 200
    assert len(meshgen.batch_samples(100 * 100, 0)) == 500
    assert len(meshgen.batch_samples(1024 * 1024, 0)) == 20000
