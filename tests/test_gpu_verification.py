"""Verification on the GPU (mvs_verification through the C ABI) against the
compiled reference's Backend::verification (oracle/_ref) and the oracle's
per-row restatement.

Bar: sides, patches and coverage identical to the reference, rms identical
(the reference's own test allows 1e-12 box, tests/cuda_verification.cpp:
71-72), every searched row's squared distance bit-identical to the oracle's
brute-force minimum, and accepted rows consistent with the oracle's states.
"""
import math
import os

import numpy as np
import pytest

from mitra_algo_amd import capi, meshgen
from oracle.pyoracle import Oracle, read_off
from tests.test_verification_oracle import (HAVE_REF, crease_mesh, example_clusters,
                                            noisy_clusters)

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref.so not built")]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def sess():
    s = capi.Session(0)
    yield s
    s.close()


@pytest.fixture(scope="module")
def ref():
    return Oracle("ref")


def topo_dict(topo):
    return dict(normals=topo.normals, weights=topo.weights, nbr_off=topo.nbr_off,
                nbr_idx=topo.nbr_idx)


def same(got, want):
    for key in ("side_a", "side_b"):
        for a, b in zip(got[key], want[key]):
            assert np.array_equal(a, b), key
    for key in ("patches_a", "patches_b"):
        for a, b in zip(got[key], want[key]):
            assert np.array_equal(a, b), key
    assert np.array_equal(got["coverage"], want["coverage"])
    assert np.array_equal(got["rms"], want["rms"])


def run(sess, ref, v, faces, topo, cl, verify):
    print(f"verification nv={len(v)} nf={len(faces)} clusters={len(cl['reflection'])} "
          f"verify={verify}", flush=True)
    got = sess.verification(v, faces, topo_dict(topo), cl, verify, topo.box_size)
    want = ref.verification(v, faces, topo, cl, verify, topo.box_size)
    same(got, want)
    return got


def rows_match_oracle(sess, v, faces, topo, cl, verify):
    nc = len(cl["reflection"])
    acc, d2 = sess.verification_rows(nc, len(v))
    state, od2 = Oracle("oracle").verification_rows(v, faces, topo.normals, cl, verify,
                                                    topo.box_size)
    searched = state != 0
    assert np.array_equal(np.isnan(d2), ~searched)
    assert np.array_equal(d2[searched], od2[searched])
    assert np.all(acc[state == 3] == 1)
    assert np.all(acc[(state < 3)] == 0)


@pytest.mark.parametrize("name", ["reflection", "rigid", "partial", "single"])
def test_examples_match_reference(sess, ref, name):
    v, f = read_off(os.path.join(GOLDEN, name + ".off"))
    topo = ref.prepare(v, f)
    cl = example_clusters(name)
    got = run(sess, ref, v, topo.faces, topo, cl, .01)
    rows_match_oracle(sess, v, topo.faces, topo, cl, .01)
    again = sess.verification(v, topo.faces, topo_dict(topo), cl, .01, topo.box_size)
    same(again, got)
    # tests/cuda_verification.cpp:65-67
    if name != "single":
        assert got["side_a"][1][0] == 0
    off_a, off_b = got["side_a"][0], got["side_b"][0]
    assert off_a[2] == off_a[1] and off_b[2] == off_b[1]


@pytest.mark.parametrize("seed", [0, 1])
def test_noisy_motions_match_reference(sess, ref, seed):
    g = np.random.default_rng(seed)
    v, f = crease_mesh(g)
    topo = ref.prepare(v, f)
    cl = noisy_clusters(g, 4)
    for verify in (.01, .05):
        run(sess, ref, v, topo.faces, topo, cl, verify)
        rows_match_oracle(sess, v, topo.faces, topo, cl, verify)


def test_double_sided_ties_match_reference(sess, ref):
    # faces at exactly equal distance with opposite normals: the reference's
    # tree and visiting order decide, and so must ours
    g = np.random.default_rng(3)
    v, f = crease_mesh(g)
    topo = ref.prepare(v, f)
    faces = np.concatenate([topo.faces, topo.faces[:, [0, 2, 1]]])
    cl = noisy_clusters(g, 4)
    run(sess, ref, v, faces, topo, cl, .05)
    rows_match_oracle(sess, v, faces, topo, cl, .05)


def cfg1_clusters(g):
    """The motif copies' motions of the 96 x 96 workload mesh (quarter turns
    about the centre and the x mirror), slightly perturbed, plus noisy
    variants."""
    cl = noisy_clusters(g, 4)
    for k in range(1, 4):
        a = k * math.pi / 2 + g.normal(0, .002)
        R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        cl["reflection"].append(0)
        cl["rotation"].append(R.ravel())
        cl["translation"].append(g.normal(0, .002, 3))
        cl["plane_normal"].append(np.zeros(3))
        cl["plane_offset"].append(0.0)
    return cl


def test_workload_mesh_matches_reference(sess, ref):
    g = np.random.default_rng(7)
    m = meshgen.cfg1_mesh()
    topo = ref.prepare(m.vertices, m.faces)
    cl = cfg1_clusters(g)
    for verify in (.01, .025):
        got = run(sess, ref, m.vertices, topo.faces, topo, cl, verify)
        assert int(got["side_a"][0][-1]) > 0


def test_large_grid_matches_reference(sess, ref):
    m = meshgen.height_field(257, 2.0, [meshgen.Copy(.5, .5, 0.0, False),
                                        meshgen.Copy(-.5, -.5, math.pi, False),
                                        meshgen.Copy(-.5, .5, 0.0, True)],
                             meshgen.Motif.make(0), 1e-4, 2)
    topo = ref.prepare(m.vertices, m.faces)
    cl = cfg1_clusters(np.random.default_rng(11))
    got = run(sess, ref, m.vertices, topo.faces, topo, cl, .025)
    assert int(got["side_a"][0][-1]) > 0 and len(got["patches_a"][1]) > 1


def test_no_clusters_and_no_faces(sess, ref):
    v, f = read_off(os.path.join(GOLDEN, "reflection.off"))
    topo = ref.prepare(v, f)
    none = dict(reflection=[], rotation=np.zeros((0, 9)), translation=np.zeros((0, 3)),
                plane_normal=np.zeros((0, 3)), plane_offset=np.zeros(0))
    got = sess.verification(v, topo.faces, topo_dict(topo), none, .01, topo.box_size)
    assert list(got["side_a"][0]) == [0] and len(got["coverage"]) == 0
    assert list(got["patches_b"][0]) == [0]
    cl = example_clusters("reflection")
    run(sess, ref, v, np.zeros((0, 3), np.int32), topo, cl, .01)


def test_invalid_inputs_raise(sess, ref):
    v, f = read_off(os.path.join(GOLDEN, "reflection.off"))
    topo = ref.prepare(v, f)
    bad = topo.faces.copy()
    bad[3, 1] = len(v)
    with pytest.raises(capi.MvsError):
        sess.verification(v, bad, topo_dict(topo), example_clusters("reflection"), .01,
                          topo.box_size)


def test_device_layer_matches_host_layer(sess, ref):
    # surface once, then two cluster sets against device-resident inputs
    import torch
    m = meshgen.height_field(129, 2.0, [meshgen.Copy(.5, .5, 0.0, False),
                                        meshgen.Copy(-.5, -.5, math.pi, False)],
                             meshgen.Motif.make(0), 1e-4, 2)
    topo = ref.prepare(m.vertices, m.faces)
    g = np.random.default_rng(13)
    sets = [cfg1_clusters(g), noisy_clusters(g, 3)]
    want = [sess.verification(m.vertices, topo.faces, topo_dict(topo), cl, .025, topo.box_size)
            for cl in sets]
    nv = len(m.vertices)
    pts = torch.from_numpy(np.ascontiguousarray(
        np.concatenate([m.vertices.T, topo.normals]))).cuda()
    off = torch.from_numpy(topo.nbr_off.astype(np.int64)).cuda()
    idx = torch.from_numpy(topo.nbr_idx).cuda()
    sess.verification_surface(m.vertices, topo.faces)
    torch.cuda.synchronize()
    for cl, w in zip(sets, want):
        print("dev_verification clusters", len(cl["reflection"]), flush=True)
        got = sess.dev_verification(nv, pts, off, idx, topo.weights, cl, .025, topo.box_size)
        same(got, w)
    with pytest.raises(capi.MvsError):
        sess.dev_verification(nv - 1, pts, off, idx, topo.weights, sets[0], .025, topo.box_size)


def test_full_mesh_matches_reference(sess, ref):
    # the 2.1M-triangle workload mesh: one copy-to-copy rigid motion and one
    # reflection through the centre, against the reference's CPU code
    from types import SimpleNamespace

    from mitra_algo_amd import capi as _capi
    from mitra_algo_amd import workload
    m = meshgen.cfg2_mesh()
    t = _capi.prepare(m.vertices, m.faces)
    topo = SimpleNamespace(normals=t["normals"], weights=t["weights"], nbr_off=t["nbr_off"],
                           nbr_idx=t["nbr_idx"], box_size=t["box_size"])
    cl = workload.copy_motions(m)
    pick = [0, len(cl["reflection"]) - 1]
    two = {k: np.asarray(v)[pick] for k, v in cl.items()}
    got = run(sess, ref, m.vertices, t["faces"], topo, two, .025)
    assert int(got["side_a"][0][-1]) > 100000 and int(got["side_b"][0][-1]) > 100000
