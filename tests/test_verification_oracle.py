"""Verification oracle on the CPU: the C restatement (oracle_verification_rows,
oracle_components, oracle_point_triangle_distance2) pinned against the
geometry known answers of tests/geometry.cpp, the compiled reference
(oracle/_ref) and closed forms on exactly mirror-symmetric grids."""
import math
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle, read_off

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
HAVE_REF = os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                       "libref.so"))
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref/libref.so not built")

ORC = Oracle("oracle")


def example_clusters(name):
    """The clusters of tests/cuda_verification.cpp:23-45: the example's own
    motion, then a stationary one."""
    if name == "rigid":
        a = .65
        R = [math.cos(a), -math.sin(a), 0, math.sin(a), math.cos(a), 0, 0, 0, 1]
        return dict(reflection=[0, 0], rotation=[R, np.eye(3).ravel()],
                    translation=[[3.2, .3, .2], [0, 0, 0]], plane_normal=[[0, 0, 0]] * 2,
                    plane_offset=[0, 0])
    return dict(reflection=[1, 0], rotation=[np.diag([-1.0, 1, 1]).ravel(), np.eye(3).ravel()],
                translation=[[0, 0, 0]] * 2, plane_normal=[[1, 0, 0], [0, 0, 0]],
                plane_offset=[0, 0])


def check_rows_against(ref_result, state, d2, topo):
    """The reference's sets agree with the oracle's row states; its rms and
    coverage follow from the oracle's distances in the reference's order."""
    nc = state.shape[0]
    # sequential sums in vertex order, as the reference accumulates them
    total = 0.0
    for w in topo.weights:
        total += float(w)
    for c in range(nc):
        sides = []
        for s, key in ((0, "side_a"), (1, "side_b")):
            off, idx = ref_result[key]
            got = idx[off[c]:off[c + 1]]
            sure = np.flatnonzero(state[c, s] == 3)
            maybe = np.flatnonzero(state[c, s] == 4)
            assert np.all(np.isin(sure, got)), (c, s)
            assert np.all(np.isin(got, np.concatenate([sure, maybe]))), (c, s)
            sides.append(got)
        err = 0.0
        for i in sides[0]:
            err += float(d2[c, 0, i])
        rms = 0.0 if len(sides[0]) == 0 else math.sqrt(err / len(sides[0]))
        assert ref_result["rms"][c] == rms
        area = 0.0
        for i in np.concatenate(sides):
            area += float(topo.weights[i])
        cov = min(1.0, area / total) if total > 0 else 0.0
        assert ref_result["coverage"][c] == cov
        for s, key in ((0, "patches_a"), (1, "patches_b")):
            groups, roff, mem = ref_result[key]
            want = ORC.components(topo.nbr_off, topo.nbr_idx, sides[s])
            got = [mem[roff[g]:roff[g + 1]] for g in range(int(groups[c]), int(groups[c + 1]))]
            assert len(got) == len(want)
            for a, b in zip(got, want):
                assert np.array_equal(a, b)


def test_triangle_distance_known_answers():
    # tests/geometry.cpp:13-19
    a, b, c = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    d = ORC.point_triangle_distance2
    assert abs(d([.2, .3, 2], a, b, c) - 4) < 1e-10
    assert abs(d([2, 0, 0], a, b, c) - 1) < 1e-10
    assert abs(d([-1, -1, 0], a, b, c) - 2) < 1e-10
    assert abs(d([1, 1, 0], a, b, c) - .5) < 1e-10
    assert abs(d([.5, 1, 0], a, b, b) - 1) < 1e-10
    assert abs(d([1, 0, 0], a, a, a) - 1) < 1e-10


def test_triangle_distance_brute_force():
    # against a dense sampling of the triangle: the exact distance is a lower
    # bound that the samples approach
    g = np.random.default_rng(5)
    for _ in range(40):
        p, a, b, c = g.uniform(-2, 2, (4, 3))
        u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
        keep = u + v <= 1
        pts = a + u[keep, None] * (b - a) + v[keep, None] * (c - a)
        brute = np.min(np.sum((pts - p) ** 2, axis=1))
        exact = ORC.point_triangle_distance2(p, a, b, c)
        assert exact <= brute + 1e-12
        assert brute - exact < 0.05 * (1 + brute)


@needs_ref
def test_triangle_distance_matches_reference_bitwise():
    # the query set of tests/cuda_verification.cpp:77-99 (seeded numpy draws
    # in [-3, 3] with the same degenerate patterns)
    ref = Oracle("ref")
    g = np.random.default_rng(12)
    for i in range(1024):
        p, a, b, c = g.uniform(-3, 3, (4, 3))
        if i % 5 == 0:
            c = a
        if i % 7 == 0:
            b = c = a
        if i % 11 == 0:
            p = a
        assert ORC.point_triangle_distance2(p, a, b, c) == ref.point_triangle_distance2(p, a, b, c)


@needs_ref
@pytest.mark.parametrize("name", ["reflection", "rigid", "partial", "single"])
def test_rows_match_reference_on_examples(name):
    ref = Oracle("ref")
    v, f = read_off(os.path.join(GOLDEN, name + ".off"))
    topo = ref.prepare(v, f)
    cl = example_clusters(name)
    r = ref.verification(v, topo.faces, topo, cl, .01, topo.box_size)
    state, d2 = ORC.verification_rows(v, topo.faces, topo.normals, cl, .01, topo.box_size)
    check_rows_against(r, state, d2, topo)
    # tests/cuda_verification.cpp:65-67: vertex 0 kept, stationary cluster empty
    if name != "single":
        assert r["side_a"][1][0] == 0
    assert r["side_a"][0][2] == r["side_a"][0][1] and r["side_b"][0][2] == r["side_b"][0][1]


def noisy_clusters(g, count):
    """Mirror planes and rigid motions near the mesh's true symmetries."""
    refl, rot, tr, pn, po = [], [], [], [], []
    for k in range(count):
        if k % 2 == 0:
            n = np.array([1.0, g.normal(0, .01), g.normal(0, .01)])
            n /= np.linalg.norm(n)
            refl.append(1)
            rot.append((np.eye(3) - 2 * np.outer(n, n)).ravel())
            tr.append(-2 * (d := g.normal(0, .003)) * n)
            pn.append(n)
            po.append(d)
        else:
            a = math.pi + g.normal(0, .01)
            R = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0],
                          [0, 0, 1]])
            refl.append(0)
            rot.append(R.ravel())
            tr.append(g.normal(0, .005, 3))
            pn.append(np.zeros(3))
            po.append(0.0)
    return dict(reflection=refl, rotation=rot, translation=tr, plane_normal=pn,
                plane_offset=po)


def crease_mesh(g, n=9):
    """A height field with a sharp ridge along x = 0 (creases make faces at
    the same distance carry very different normals)."""
    xs = np.linspace(-1, 1, n)
    X, Y = np.meshgrid(xs, xs)
    Z = 1.5 * (1 - np.abs(X)) + g.normal(0, 1e-3, X.shape)
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    F = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i + 1, (j + 1) * n + i
            F += [[a, b, c], [a, c, d]] if (i + j) % 2 == 0 else [[a, b, d], [b, c, d]]
    return V, np.array(F, np.int32)


@needs_ref
@pytest.mark.parametrize("seed", [0, 1])
def test_rows_match_reference_on_noisy_motions(seed):
    ref = Oracle("ref")
    g = np.random.default_rng(seed)
    v, f = crease_mesh(g)
    topo = ref.prepare(v, f)
    cl = noisy_clusters(g, 4)
    for verify in (.01, .05):
        r = ref.verification(v, topo.faces, topo, cl, verify, topo.box_size)
        state, d2 = ORC.verification_rows(v, topo.faces, topo.normals, cl, verify, topo.box_size)
        check_rows_against(r, state, d2, topo)


@needs_ref
def test_double_sided_mesh_ties_follow_the_reference():
    # every face twice, the copy wound the other way: rows whose nearest point
    # is a vertex see two faces at exactly the same distance with opposite
    # normals, so only the reference's visiting order names the result
    ref = Oracle("ref")
    g = np.random.default_rng(3)
    v, f = crease_mesh(g)
    topo = ref.prepare(v, f)
    faces = np.concatenate([topo.faces, topo.faces[:, [0, 2, 1]]])
    cl = noisy_clusters(g, 4)
    r = ref.verification(v, faces, topo, cl, .05, topo.box_size)
    state, d2 = ORC.verification_rows(v, faces, topo.normals, cl, .05, topo.box_size)
    assert np.count_nonzero(state == 4) > 0
    check_rows_against(r, state, d2, topo)


def test_mirror_closed_form_on_symmetric_grid():
    # z = f(|x|, y) on a grid symmetric in x, mirror x -> -x: side A is every
    # vertex with x < -tau, side B every vertex with x > tau, all at distance 0
    n = 21
    right = np.linspace(-1, 1, n)[n // 2 + 1:]
    xs = np.concatenate([-right[::-1], [0.0], right])
    X, Y = np.meshgrid(xs, xs)
    Z = .3 * np.exp(-((np.abs(X) - .4) ** 2 + Y ** 2) / .1)
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    F = []
    for j in range(n - 1):
        for i in range(n - 1):
            a, b, c, d = j * n + i, j * n + i + 1, (j + 1) * n + i + 1, (j + 1) * n + i
            # diagonal mirrored with x so the triangulation is symmetric
            F += [[a, b, c], [a, c, d]] if i < n // 2 else [[a, b, d], [b, c, d]]
    F = np.array(F, np.int32)
    topo = ORC.prepare(V, F)
    cl = dict(reflection=[1], rotation=[np.diag([-1.0, 1, 1]).ravel()], translation=[[0, 0, 0]],
              plane_normal=[[1, 0, 0]], plane_offset=[0])
    verify = .02
    tau = verify * topo.box_size
    state, d2 = ORC.verification_rows(V, topo.faces, topo.normals, cl, verify, topo.box_size)
    x = V[:, 0]
    assert np.array_equal(np.flatnonzero(state[0, 0] == 3), np.flatnonzero(x < -tau))
    assert np.array_equal(np.flatnonzero(state[0, 1] == 3), np.flatnonzero(x > tau))
    assert np.all(d2[0, 0, x < -tau] == 0) and np.all(d2[0, 1, x > tau] == 0)
    assert np.all(state[0, 0, x >= -tau] == 0) and np.all(state[0, 1, x <= tau] == 0)


def test_components_closed_form():
    # a path 0-1-2-3-4-5 with 2 removed: patches {0, 1} and {3, 4, 5}; seeds
    # in the given order, each patch sorted
    off = np.array([0, 1, 3, 5, 7, 9, 10], np.uint64)
    idx = np.array([1, 0, 2, 1, 3, 2, 4, 3, 5, 4], np.int32)
    got = ORC.components(off, idx, np.array([5, 0, 1, 3, 4], np.int32))
    assert [list(p) for p in got] == [[3, 4, 5], [0, 1]]
    assert ORC.components(off, idx, np.array([], np.int32)) == []
