"""prepare() on the device (mvs_dev_prepare) against the host mvs_prepare,
which equals the oracle byte for byte (tests/test_host.py): every array --
kept faces, face and vertex normals, weights, ring and incident-face CSRs,
the (lo, hi)-ordered edge records -- and every scalar must be identical,
on the reference's example meshes, on height fields up to the 2.1M-triangle
mesh, and on meshes with rejected faces and unused vertices; errors carry
the host's messages.  Curvature over the device topology equals curvature
over the host topology bit for bit."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mitra_algo_amd import capi, meshgen  # noqa: E402
from oracle.pyoracle import read_off  # noqa: E402

pytestmark = pytest.mark.gpu
EXAMPLES = ["reflection", "rigid", "partial", "single"]


@pytest.fixture(scope="module")
def sess():
    s = capi.Session(0)
    s.set_stream(torch.cuda.current_stream(torch.device("cuda:0")).cuda_stream)
    yield s
    s.close()


def _same(sess, v, f):
    print("dev_prepare", len(v), len(f), flush=True)
    h = capi.prepare(v, f)
    d = sess.dev_prepare(v, f)
    torch.cuda.synchronize()
    F, nv = d["face_count"], d["nv"]
    assert F == len(h["faces"])
    assert d["rejected_faces"] == h["rejected_faces"]
    assert d["box_size"] == h["box_size"]
    assert np.array_equal(d["center"], h["center"])
    assert np.array_equal(d["xyz"].cpu().numpy(), np.ascontiguousarray(v.T))
    assert np.array_equal(d["faces3"][:, :F].cpu().numpy().T, h["faces"])
    assert np.array_equal(d["face_normals"][:, :F].cpu().numpy(), h["face_normals"])
    assert np.array_equal(d["normals"].cpu().numpy(), h["normals"])
    assert np.array_equal(d["weights"].cpu().numpy(), h["weights"])
    assert np.array_equal(d["nbr_off"].cpu().numpy().view(np.uint64), h["nbr_off"])
    assert d["nbr_nnz"] == len(h["nbr_idx"])
    assert np.array_equal(d["nbr_idx"][:d["nbr_nnz"]].cpu().numpy(), h["nbr_idx"])
    assert np.array_equal(d["inc_off"].cpu().numpy().view(np.uint64), h["inc_off"])
    assert d["inc_nnz"] == len(h["inc_idx"])
    assert np.array_equal(d["inc_idx"][:d["inc_nnz"]].cpu().numpy(), h["inc_idx"])
    assert d["edge_count"] == len(h["edges"])
    assert np.array_equal(d["edges"][:d["edge_count"]].cpu().numpy(), h["edges"])
    assert nv == len(v)
    return h, d


@pytest.mark.parametrize("name", EXAMPLES)
def test_examples(sess, golden_dir, name):
    v, f = read_off(os.path.join(golden_dir, f"{name}.off"))
    _same(sess, v, f)


@pytest.mark.parametrize("side", [2, 17, 96, 300])
def test_heightfields(sess, side):
    m = meshgen.cfg1_mesh() if side == 96 else meshgen.cfg2_mesh(side)
    _same(sess, m.vertices, m.faces)


def test_full_size_mesh(sess):
    m = meshgen.cfg2_mesh(1024)
    _same(sess, m.vertices, m.faces)


def test_rejected_faces_and_unused_vertices(sess):
    m = meshgen.cfg1_mesh()
    v = np.vstack([m.vertices, [[5.0, 5.0, 5.0], [6.0, 6.0, 6.0]]])  # two unused vertices
    f = m.faces.copy()
    g = np.random.default_rng(3)
    bad = g.choice(len(f), 40, replace=False)
    f[bad, 2] = f[bad, 1]  # degenerate (repeated vertex): rejected
    f = np.vstack([f, [[0, 1, 1]]]).astype(np.int32)
    h, d = _same(sess, v, f)
    assert h["rejected_faces"] == 41


def test_shuffled_faces_and_vertices(sess):
    m = meshgen.cfg2_mesh(64)
    g = np.random.default_rng(5)
    perm = g.permutation(len(m.vertices))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    v = m.vertices[perm]
    f = inv[m.faces][g.permutation(len(m.faces))].astype(np.int32)
    # rotate each face's corners (same winding)
    r = g.integers(0, 3, len(f))
    f = np.stack([f[np.arange(len(f)), (r + k) % 3] for k in range(3)], 1).astype(np.int32)
    _same(sess, v, f)


def test_errors(sess):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    with pytest.raises(capi.MvsError, match="empty"):
        sess.dev_prepare(v, np.zeros((0, 3), np.int32))
    with pytest.raises(capi.MvsError, match="out of range"):
        sess.dev_prepare(v, np.array([[0, 1, 9]], np.int32))
    with pytest.raises(capi.MvsError, match="out of range"):
        sess.dev_prepare(v, np.array([[0, 1, 2], [-1, 1, 2]], np.int32))
    with pytest.raises(capi.MvsError, match="winding"):
        sess.dev_prepare(v, np.array([[0, 1, 2], [0, 1, 3]], np.int32))
    with pytest.raises(capi.MvsError, match="manifold"):
        sess.dev_prepare(np.vstack([v, [[1, 1, 1]]]),
                         np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], np.int32))
    bad = v.copy()
    bad[1, 0] = np.nan
    with pytest.raises(capi.MvsError, match="Non-finite"):
        sess.dev_prepare(bad, np.array([[0, 1, 2]], np.int32))
    with pytest.raises(capi.MvsError, match="zero extent"):
        sess.dev_prepare(np.zeros((3, 3)), np.array([[0, 1, 2]], np.int32))
    with pytest.raises(capi.MvsError, match="nondegenerate"):
        sess.dev_prepare(v, np.array([[0, 1, 1]], np.int32))
    # the session still works after the errors
    m = meshgen.cfg2_mesh(8)
    _same(sess, m.vertices, m.faces)


def test_curvature_over_device_topology(sess):
    m = meshgen.cfg2_mesh(128)
    dev = torch.device("cuda:0")
    h = capi.prepare(m.vertices, m.faces)
    d = sess.dev_prepare(m.vertices, m.faces)
    n, nf = len(m.vertices), len(m.faces)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a)).to(dev)
    fa, va = torch.empty((11, n), dtype=torch.float64, device=dev), torch.empty(
        n, dtype=torch.uint8, device=dev)
    fb, vb = torch.empty_like(fa), torch.empty_like(va)
    x, y, z = (t(m.vertices[:, k]) for k in range(3))
    sess.dev_curvature(n, x, y, z, t(h["normals"]), len(h["faces"]), t(h["face_normals"]),
                       t(h["nbr_off"].view(np.int64)), t(h["nbr_idx"]), len(h["edges"]),
                       t(h["edges"]), 3, fa, va)
    xyz = d["xyz"]
    sess.dev_curvature(n, xyz[0], xyz[1], xyz[2], d["normals"], nf, d["face_normals"],
                       d["nbr_off"], d["nbr_idx"], d["edge_count"], d["edges"], 3, fb, vb)
    torch.cuda.synchronize()
    assert torch.equal(fa, fb) and torch.equal(va, vb)
    assert int(va.sum()) > n // 4
