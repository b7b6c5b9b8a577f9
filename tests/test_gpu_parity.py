"""Parity of the HIP path (through the C ABI) with the oracle on the
reference's own fixtures and on generated height fields.

Bars (the reference's own, tests/cuda_*.cpp): pairs and prune lists
byte-equal including order; vote lanes within 2e-12 (1 + |a|); curvature
scalars within 1e-10 (1 + |a|) with exact validity and sign-invariant
directions; cluster memberships exact, centres within 1e-10 (1 + |a|), RNG
state equal afterwards.
"""
import math
import os

import numpy as np
import pytest

from mitra_algo_amd import capi, meshgen
from oracle.pyoracle import read_off
from tests import fixtures as fx

pytestmark = pytest.mark.gpu
EXAMPLES = ["reflection", "rigid", "partial", "single"]
PRECISIONS = [capi.FP64, capi.FP32_GUARDED]


@pytest.fixture(scope="module")
def sess():
    s = capi.Session(0)
    yield s
    s.close()


def close(a, b, tol):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    assert a.shape == b.shape
    assert np.all(np.isfinite(b))
    err = np.abs(a - b) / (1 + np.abs(a))
    assert err.max(initial=0) <= tol, float(err.max())


# ------------------------------------------------------------- curvature
@pytest.mark.parametrize("name", EXAMPLES)
def test_curvature_examples(sess, oracle, golden_dir, name):
    v, f = read_off(os.path.join(golden_dir, f"{name}.off"))
    topo = capi.prepare(v, f)
    otopo = oracle.prepare(v, f)
    for smooth in (0, 3, 7):
        a_feat, a_valid = oracle.curvature(v, otopo, smooth)
        b_feat, b_valid = sess.curvature(v, topo, smooth)
        _check_curvature(a_feat, a_valid, b_feat, b_valid)
        again, _ = sess.curvature(v, topo, smooth)
        assert np.array_equal(again, b_feat)


def test_curvature_heightfield(sess, oracle):
    m = meshgen.cfg1_mesh()
    topo = capi.prepare(m.vertices, m.faces)
    a_feat, a_valid = oracle.curvature(m.vertices, oracle.prepare(m.vertices, m.faces), 3)
    b_feat, b_valid = sess.curvature(m.vertices, topo, 3)
    _check_curvature(a_feat, a_valid, b_feat, b_valid)


def test_curvature_planar(sess):
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], float)
    topo = capi.prepare(v, np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    feat, valid = sess.curvature(v, topo, 3)
    assert not valid.any() and np.all(feat[9:11] == 0)


def _check_curvature(a, av, b, bv):
    assert np.array_equal(av, bv)
    close(a[9], b[9], 1e-10)
    close(a[10], b[10], 1e-10)
    close(a[0:3], b[0:3], 1e-10)
    ok = av.astype(bool)
    for lo in (3, 6):
        d = np.abs((a[lo:lo + 3, ok] * b[lo:lo + 3, ok]).sum(0))
        assert np.all(d > 1 - 1e-10)
        assert np.allclose(np.linalg.norm(b[lo:lo + 3, ok], axis=0), 1, atol=1e-10)
    assert np.all(b[9, ok] <= b[10, ok])


# ----------------------------------------------------------------- prune
def test_prune_fixtures(sess, oracle):
    w, kmin, kmax, valid = fx.select_fixture(1028)
    feat = np.zeros((11, 1028))
    feat[9], feat[10] = kmin, kmax
    for ratio in (1.0, 0.5, 0.001):
        g1, g2 = capi.rng(42), oracle.rng(42)
        s1 = capi.sample_vertices(w, ratio, g1)
        s2, p2 = oracle.select(w, kmin, kmax, valid, ratio, 0.85, g2)
        p1 = sess.prune(feat, valid, s1, 0.85)
        assert np.array_equal(s1, s2) and np.array_equal(p1, p2) and bytes(g1) == bytes(g2)
    w, kmin, kmax, valid = fx.strict_prune_fixture(1028)
    feat[9], feat[10] = kmin, kmax
    p = sess.prune(feat, valid, np.arange(1028, dtype=np.int32), 0.75)
    assert np.array_equal(p, oracle.prune(kmin, kmax, valid, np.arange(1028), 0.75))
    assert len(p) == 1028 * 2 // 3
    assert len(sess.prune(feat, np.zeros(1028, np.uint8), np.arange(1028), 0.75)) == 0
    with pytest.raises(capi.MvsInvalidArgument):
        sess.prune(feat, valid, np.array([5000], np.int32), 0.75)


# --------------------------------------------------------------- pairing
def _gpu_pairs(sess, fix, t, max_pairs=2000000, precision=capi.FP64):
    return sess.pairing(fix.vertices, fix.feat, fix.selected, fix.box, t, max_pairs, precision)


def _cpu_pairs(oracle, fix, t, max_pairs=2000000):
    return oracle.pairing(fix.vertices, fix.kmin, fix.kmax, fix.selected, fix.box, t, max_pairs)


def _same(a, b):
    assert len(a[0]) == len(b[0])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pairing_fixtures(sess, oracle, precision):
    assert len(_gpu_pairs(sess, fx.PairFixture(0), 0.15, precision=precision)[0]) == 0
    assert len(_gpu_pairs(sess, fx.PairFixture(1), 0.15, precision=precision)[0]) == 0
    dense = fx.PairFixture(513)
    d = _gpu_pairs(sess, dense, 0.15, precision=precision)
    _same(d, _cpu_pairs(oracle, dense, 0.15))
    assert len(d[0]) == 513 * 512 // 2
    _same(_gpu_pairs(sess, dense, 0.15, len(d[0]), precision), d)
    with pytest.raises(capi.MvsBudgetError, match="budget"):
        _gpu_pairs(sess, dense, 0.15, len(d[0]) - 1, precision)
    rnd = fx.pairing_random()
    for t in fx.RANDOM_THRESHOLDS:
        _same(_gpu_pairs(sess, rnd, t, precision=precision), _cpu_pairs(oracle, rnd, t))
    dup = fx.pairing_random()
    dup.selected = np.append(dup.selected, dup.selected[0]).astype(np.int32)
    _same(_gpu_pairs(sess, dup, 0.7, precision=precision), _cpu_pairs(oracle, dup, 0.7))
    r1 = _gpu_pairs(sess, rnd, 0.15, precision=precision)
    r2 = _gpu_pairs(sess, rnd, 0.15, precision=precision)
    _same(r1, r2)
    bnd = fx.pairing_boundary()
    for t in (0.25, 0.5, 1.0):
        _same(_gpu_pairs(sess, bnd, t, precision=precision), _cpu_pairs(oracle, bnd, t))
    for t in fx.EDGE_THRESHOLDS:
        e = fx.pairing_edge(t)
        _same(_gpu_pairs(sess, e, t, precision=precision), _cpu_pairs(oracle, e, t))
    rd = _gpu_pairs(sess, fx.pairing_rounded(), math.nextafter(1.0, 0.0), precision=precision)
    assert len(rd[0]) == 1
    z = _gpu_pairs(sess, fx.pairing_zeros(), 0.15, precision=precision)
    assert len(z[0]) == 190


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pairing_rejects_invalid_input(sess, precision):
    dense = fx.PairFixture(513)
    for box, t in ((0.0, 0.15), (math.nan, 0.15), (1.0, 0.0), (1.0, math.nan), (1.0, 1.1)):
        with pytest.raises(capi.MvsError):
            sess.pairing(dense.vertices, dense.feat, dense.selected, box, t, 2000000, precision)
    for bad in (-1, 513):
        f = fx.PairFixture(513)
        f.selected[0] = bad
        with pytest.raises(capi.MvsInvalidArgument):
            _gpu_pairs(sess, f, 0.15, precision=precision)
    f = fx.PairFixture(513)
    f.feat[9][0] = math.nan
    with pytest.raises(capi.MvsInvalidArgument):
        _gpu_pairs(sess, f, 0.15, precision=precision)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pairing_large_fixtures(sess, oracle, precision):
    sp = fx.pairing_sparse(65536)
    assert len(_gpu_pairs(sess, sp, 1e-10, max_pairs=1, precision=precision)[0]) == 0
    b = fx.pairing_benchmark()
    g = _gpu_pairs(sess, b, 1e-6, 600000, precision)
    assert len(g[0]) == 495000
    _same(g, _cpu_pairs(oracle, b, 1e-6, 600000))


@pytest.mark.parametrize("precision", PRECISIONS)
def test_pairing_heightfield(sess, oracle, precision):
    m = meshgen.cfg1_mesh()
    topo = oracle.prepare(m.vertices, m.faces)
    feat, valid = oracle.curvature(m.vertices, topo, 3)
    s = meshgen.uniform_samples(2000, len(m.vertices), 3)
    sel = oracle.prune(feat[9], feat[10], valid, s, 0.85)
    for t in (0.05, 0.15):
        g = sess.pairing(m.vertices, feat, sel, topo.box_size, t, 10**8, precision)
        c = oracle.pairing(m.vertices, feat[9], feat[10], sel, topo.box_size, t, 10**8)
        _same(g, c)


# ------------------------------------------------------- transformations
def _cmp_votes(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    close(a[2], b[2], 2e-12)


def test_transformation_fixtures(sess, oracle):
    v, feat, pa, pb, box, ctr, n, d, x = fx.mirror_fixture()
    a = oracle.transformations(v, feat, pa, pb, box, ctr)
    b = sess.transformations(v, feat, pa, pb, box, ctr)
    assert len(b[0]) == 3 * len(pa)
    _cmp_votes(a, b)
    again = sess.transformations(v, feat, pa, pb, box, ctr)
    assert np.array_equal(again[2], b[2])
    mixed = fx.mixed_features(feat, n)
    a = oracle.transformations(v, mixed, pa, pb, box, ctr)
    b = sess.transformations(v, mixed, pa, pb, box, ctr)
    _cmp_votes(a, b)
    assert 2 * len(pa) < len(b[0]) < 3 * len(pa)
    assert len(sess.transformations(v, mixed, pa, pb, box, ctr, len(b[0]))[0]) == len(b[0])
    with pytest.raises(capi.MvsBudgetError):
        sess.transformations(v, mixed, pa, pb, box, ctr, 1)
    par = fx.parallel_features(len(v), n, d, x)
    b = sess.transformations(v, par, pa, pb, box, ctr)
    assert len(b[0]) == 2 * len(pa)
    _cmp_votes(oracle.transformations(v, par, pa, pb, box, ctr), b)
    same = np.zeros(1, np.int32)
    for f in (par, np.zeros_like(par)):
        _cmp_votes(oracle.transformations(v, f, same, same, box, ctr),
                   sess.transformations(v, f, same, same, box, ctr))
    for c, p in (((0.0, [0, 0, 0])), ((math.nan, [0, 0, 0]))):
        with pytest.raises(capi.MvsInvalidArgument):
            sess.transformations(v, par, pa, pb, c, p)
    with pytest.raises(capi.MvsInvalidArgument):
        sess.transformations(v, par, np.array([-1], np.int32), same, box, ctr)
    with pytest.raises(capi.MvsInvalidArgument):
        sess.transformations(v, par, np.array([len(v)], np.int32), same, box, ctr)


# ------------------------------------------------------------ clustering
def _cmp_clusters(a, b):
    assert a.limited_iterations == b["limited_iterations"]
    assert np.array_equal(a.member_off, b["member_off"])
    assert np.array_equal(a.members, b["members"])
    assert np.array_equal(a.reflection, b["reflection"])
    for k in ("rotation", "translation", "plane_normal", "plane_offset"):
        close(getattr(a, k), b[k], 1e-10)


@pytest.mark.parametrize("count", [0, 1, 64, 2048])
def test_clustering_fixtures(sess, oracle, count):
    kind, desc, rot = fx.cluster_votes(count)
    for iters in (1, 100):
        g1, g2, g3 = oracle.rng(42), capi.rng(42), capi.rng(42)
        a = oracle.clustering(kind, desc, rot, 10.0, [0, 0, 0], 0.03, iters, 4, g1)
        b = sess.clustering(kind, desc, rot, 10.0, [0, 0, 0], 0.03, iters, 4, g2)
        c = sess.clustering(kind, desc, rot, 10.0, [0, 0, 0], 0.03, iters, 4, g3)
        assert bytes(g1) == bytes(g2) == bytes(g3)
        _cmp_clusters(a, b)
        assert np.array_equal(b["members"], c["members"])
        assert np.array_equal(b["rotation"], c["rotation"])


def _ramp_votes(count: int):
    """Votes whose density rises linearly along one descriptor axis (rigid:
    component 9, mirror: component 3, values sqrt(u) on [0, 1]) with the
    other components fixed, so mean-shift trajectories climb the ramp by a
    few percent of the bandwidth per iteration and end far from their seeds:
    every trajectory's candidate list is rebuilt several times."""
    u = (np.arange(count) * 0.6180339887498949) % 1.0
    x = np.sqrt(u)
    kind = (np.arange(count) % 3 == 2).astype(np.uint8)
    desc = np.zeros((12, count))
    rot = np.zeros((9, count))
    ident = np.eye(3).reshape(9)
    for i in range(count):
        if kind[i]:
            rot[:, i] = np.diag([-1.0, 1.0, 1.0]).reshape(9)
            desc[0:4, i] = (1.0, 0.0, 0.0, x[i])
        else:
            rot[:, i] = ident
            desc[0:9, i] = ident / 2
            desc[9, i] = x[i]
            desc[10, i] = 1e-3 * np.sin(i)
    return kind, desc, rot


def test_clustering_drifting_trajectories(sess, oracle):
    kind, desc, rot = _ramp_votes(3000)
    h = 0.05
    # the fixture does what it is for: from the lowest rigid vote a
    # trajectory under the same weights moves more than 2h in 100 iterations
    rig = desc[:, kind == 0]
    m = rig[:, np.argmin(rig[9])].copy()
    start = m.copy()
    for _ in range(100):
        d2 = ((rig - m[:, None]) ** 2).sum(0)
        w = np.where(d2 < h * h, 1 - d2 / (h * h), 0.0)
        m = (rig * w).sum(1) / w.sum()
    assert np.sqrt(((m - start) ** 2).sum()) > 2 * h
    for iters in (100, 300):
        g1, g2 = oracle.rng(7), capi.rng(7)
        a = oracle.clustering(kind, desc, rot, 4.0, [0.5, 0, 0], h, iters, 20, g1)
        b = sess.clustering(kind, desc, rot, 4.0, [0.5, 0, 0], h, iters, 20, g2)
        assert bytes(g1) == bytes(g2)
        _cmp_clusters(a, b)


# ------------------------------------------------------------ end to end
@pytest.mark.parametrize("name", EXAMPLES)
def test_end_to_end_examples(sess, oracle, golden_dir, name):
    """Full chain on the GPU (prepare on the host) against the oracle chain;
    pairing runs on the oracle's features, as the reference's own pairing
    tests do, so the pair set is compared bit for bit."""
    v, f = read_off(os.path.join(golden_dir, f"{name}.off"))
    topo = capi.prepare(v, f)
    ot = oracle.prepare(v, f)
    feat, valid = oracle.curvature(v, ot, 3)
    gfeat, gvalid = sess.curvature(v, topo, 3)
    _check_curvature(feat, valid, gfeat, gvalid)
    g1, g2 = oracle.rng(42), capi.rng(42)
    s_o, p_o = oracle.select(ot.weights, feat[9], feat[10], valid, 1.0, 0.85, g1)
    s_g = capi.sample_vertices(topo["weights"], 1.0, g2)
    p_g = sess.prune(feat, valid, s_g, 0.85)
    assert np.array_equal(p_o, p_g)
    pa_o, pb_o = oracle.pairing(v, feat[9], feat[10], p_o, ot.box_size, 0.08)
    pa_g, pb_g = sess.pairing(v, feat, p_g, topo["box_size"], 0.08)
    _same((pa_o, pb_o), (pa_g, pb_g))
    vo = oracle.transformations(v, feat, pa_o, pb_o, ot.box_size, ot.center)
    vg = sess.transformations(v, feat, pa_g, pb_g, topo["box_size"], topo["center"])
    _cmp_votes(vo, vg)
    co = oracle.clustering(vo[1], vo[2][16:28], vo[2][0:9], ot.box_size, ot.center, 0.08, 100,
                           20, g1)
    cg = sess.clustering(vo[1], vo[2][16:28], vo[2][0:9], ot.box_size, ot.center, 0.08, 100, 20,
                         g2)
    _cmp_clusters(co, cg)
    assert bytes(g1) == bytes(g2)
