"""Pins the FP64 oracle's geometry against the reference's own tests.

Each case regenerates the reference's inputs bit-exactly (the suite's LCG,
test/bvh.test.ts:7-13, is exact in uint32) and checks the assertion the
reference makes, or a stronger one (exact equality where it uses 5 decimals).
"""
import math

import numpy as np
import pytest

import oracle as O


def normalize(v):
    v = np.asarray(v, np.float64)
    n = math.sqrt(float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
    return v / n if n else v


# ---- ray.test.ts:4-37 --------------------------------------------------------
V0, V1, V2 = [-1, -1, 0], [1, -1, 0], [0, 1, 0]


def test_triangle_hits_both_sides_with_barycentrics():
    front = O.intersect_triangle([0, 0, -3], [0, 0, 1], V0, V1, V2)
    back = O.intersect_triangle([0, 0, 3], [0, 0, -1], V0, V1, V2)
    assert front[0] == pytest.approx(3) and back[0] == pytest.approx(3)
    t, u, v = O.intersect_triangle([0, -1 / 3, -1], [0, 0, 1], V0, V1, V2)
    assert u == pytest.approx(1 / 3) and v == pytest.approx(1 / 3)
    # Closed form for a point inside: origin (x, y) on z = 0 plane gives
    # u, v from the affine map of the triangle; (0.25, -0.5): v0 + u e1 + v e2.
    t, u, v = O.intersect_triangle([0.25, -0.5, -2], [0, 0, 1], V0, V1, V2)
    assert (u, v, t) == pytest.approx((0.5, 0.25, 2.0))


def test_triangle_misses():
    assert O.intersect_triangle([2, 2, -1], [0, 0, 1], V0, V1, V2) is None
    assert O.intersect_triangle([0, 0, -1], [0, 0, -1], V0, V1, V2) is None
    assert O.intersect_triangle([0, 0, -3], [0, 0, 1], V0, V1, V2, 1e-5, 2) is None
    assert O.intersect_triangle([0, 0, 0], [1, 0, 0], V0, V1, V2) is None  # parallel
    # tMin is exclusive, tMax exclusive: t == 3 exactly at both bounds.
    assert O.intersect_triangle([0, 0, -3], [0, 0, 1], V0, V1, V2, 3.0, 10) is None
    assert O.intersect_triangle([0, 0, -3], [0, 0, 1], V0, V1, V2, 0, 3.0) is None


def test_sphere_roots():
    assert O.intersect_sphere([0, 0, -5], [0, 0, 1], [0, 0, 0], 1) == pytest.approx(4)
    assert O.intersect_sphere([0, 0, 0], [0, 0, 1], [0, 0, 0], 1) == pytest.approx(1)
    assert O.intersect_sphere([0, 5, -5], [0, 0, 1], [0, 0, 0], 1) is None
    assert O.intersect_sphere([0, 0, 5], [0, 0, 1], [0, 0, 0], 1) is None


# ---- aabb.test.ts:21-33 -------------------------------------------------------
def _inv(d):
    with np.errstate(divide="ignore"):
        return 1.0 / np.asarray(d, np.float64)


def test_slab_known_answers():
    lo, hi = [-1, -1, -1], [1, 1, 1]
    assert O.aabb_intersect(lo, hi, [0, 0, -5], _inv([0, 0, 1])) == pytest.approx(4)
    assert O.aabb_intersect(lo, hi, [0, 0, -5], _inv([0, 0, -1])) == math.inf
    assert O.aabb_intersect(lo, hi, [5, 5, -5], _inv([0, 0, 1])) == math.inf
    assert O.aabb_intersect(lo, hi, [0.5, 0.5, -5], _inv([0, 0, 1])) == pytest.approx(4)
    assert O.aabb_intersect(lo, hi, [0, 0, -5], _inv([0, 0, 1]), 2) == math.inf
    assert O.aabb_intersect(lo, hi, [0, 0, 0], _inv([1, 0, 0])) == 0
    # Origin exactly on a slab plane with a zero direction component: 0*inf = NaN
    # must be ignored, not reject (aabb.ts:84-98 comparisons are false for NaN).
    assert O.aabb_intersect(lo, hi, [1, 0, -5], _inv([0, 0, 1])) == pytest.approx(4)


# ---- bvh.test.ts:30-87 ------------------------------------------------------
def test_empty_and_single_leaf():
    b = O.build_bvh(np.zeros(0, np.float32))
    assert b.node_count == 1
    b = O.build_bvh(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32))
    assert b.node_count == 1 and b.nodes[7] == 1 and b.nodes[3] == 0


def _leaf_coverage(b, n):
    seen = np.zeros(n, np.int64)
    nodes = b.nodes.reshape(-1, 8)
    for k in range(b.node_count):
        cnt, first = int(nodes[k, 7]), int(nodes[k, 3])
        for i in range(cnt):
            seen[b.tri_indices[first + i]] += 1
    return seen


def test_every_triangle_exactly_once():
    soup = O.random_soup(500, O.Lcg(7))
    b = O.build_bvh(soup)
    assert (_leaf_coverage(b, 500) == 1).all()


def test_children_adjacent_and_nested():
    soup = O.random_soup(300, O.Lcg(3))
    b = O.build_bvh(soup)
    nodes = b.nodes.reshape(-1, 8)
    for k in range(b.node_count):
        if nodes[k, 7] != 0:
            continue
        left = int(nodes[k, 3])
        for c in (left, left + 1):
            assert (nodes[c, :3] >= nodes[k, :3] - 1e-6).all()
            assert (nodes[c, 4:7] <= nodes[k, 4:7] + 1e-6).all()
    # Tighter than the reference: each node's box is exactly the union of its children's.
    for k in range(b.node_count):
        if nodes[k, 7] == 0 and b.node_count > 1:
            left = int(nodes[k, 3])
            assert (np.minimum(nodes[left, :3], nodes[left + 1, :3]) == nodes[k, :3]).all()
            assert (np.maximum(nodes[left, 4:7], nodes[left + 1, 4:7]) == nodes[k, 4:7]).all()


def test_coincident_centroids_terminate():
    one = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    soup = np.array(one * 64, np.float32)
    b = O.build_bvh(soup, 4)
    assert b.node_count > 1
    t, tri, u, v = O.intersect_bvh_batch(b, soup, np.array([[0.2, 0.2, -1.0]]), np.array([[0.0, 0.0, 1.0]]))
    assert t[0] == pytest.approx(1)
    # 64 identical triangles tie at t = 1: the left-first walk keeps the lowest slot.
    assert tri[0] == b.tri_indices[0]


# ---- bvh.test.ts:89-132 -----------------------------------------------------
def test_bvh_matches_brute_force_random_soup():
    rand = O.Lcg(42)
    soup = O.random_soup(2000, rand)
    b = O.build_bvh(soup)
    org = np.zeros((400, 3))
    dirs = np.zeros((400, 3))
    for i in range(400):
        org[i] = [rand() * 16 - 8, rand() * 16 - 8, rand() * 16 - 8]
        dirs[i] = normalize([rand() - 0.5, rand() - 0.5, rand() - 0.5])
    a = O.intersect_bvh_batch(b, soup, org, dirs)
    r = O.intersect_all_batch(soup, org, dirs)
    hits = int((r[1] != 0xFFFFFFFF).sum())
    assert hits > 50
    # Exact equality of every field (the reference checks t to 5 decimals).
    for x, y in zip(a, r):
        assert np.array_equal(x, y)


def test_bvh_matches_brute_force_torus_knot():
    pos, nrm = O.torus_knot(tubular=80, radial=10)
    assert pos.size // 9 == 1600
    b = O.build_bvh(pos)
    rand = O.Lcg(99)
    org = np.zeros((200, 3))
    dirs = np.zeros((200, 3))
    for i in range(200):
        org[i] = [rand() * 6 - 3, rand() * 6 - 3, 4]
        dirs[i] = normalize([rand() * 0.4 - 0.2, rand() * 0.4 - 0.2, -1])
    a = O.intersect_bvh_batch(b, pos, org, dirs)
    r = O.intersect_all_batch(pos, org, dirs)
    assert np.array_equal(a[0], r[0]) and np.array_equal(a[1], r[1])
    assert (a[1] != 0xFFFFFFFF).sum() > 20


def test_bvh_respects_tmax():
    soup = O.random_soup(100, O.Lcg(5))
    b = O.build_bvh(soup)
    o = np.array([[0.0, 0.0, -20.0]])
    d = np.array([[0.0, 0.0, 1.0]])
    ref = O.intersect_all_batch(soup, o, d)
    if ref[1][0] != 0xFFFFFFFF:  # the reference's own case (it passes vacuously on a miss)
        a = O.intersect_bvh_batch(b, soup, o, d, 1e-5, ref[0][0] * 0.5)
        assert a[1][0] == 0xFFFFFFFF and a[0][0] == math.inf
    # A case that always hits: aim at the centroid of triangle 17.
    c = soup.reshape(-1, 3, 3)[17].astype(np.float64).mean(axis=0)
    o = np.array([[0.0, 0.0, -20.0]])
    d = normalize(c - o[0])[None, :]
    ref = O.intersect_all_batch(soup, o, d)
    assert ref[1][0] != 0xFFFFFFFF
    assert O.intersect_bvh_batch(b, soup, o, d)[0][0] == ref[0][0]
    a = O.intersect_bvh_batch(b, soup, o, d, 1e-5, ref[0][0] * 0.5)
    assert a[1][0] == 0xFFFFFFFF and a[0][0] == math.inf
    a = O.intersect_bvh_batch(b, soup, o, d, 1e-5, ref[0][0])  # tMax is exclusive
    assert a[1][0] != ref[1][0] or a[0][0] > ref[0][0]


def test_bvh_matches_brute_force_many_random():
    rng = np.random.default_rng(3)
    soup = (rng.random(9 * 3000) * 10 - 5).astype(np.float32)
    b = O.build_bvh(soup)
    org = rng.random((3000, 3)) * 16 - 8
    dirs = rng.standard_normal((3000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    a = O.intersect_bvh_batch(b, soup, org, dirs)
    r = O.intersect_all_batch(soup, org, dirs)
    for x, y in zip(a, r):
        assert np.array_equal(x, y)


# ---- mesh.test.ts:47-55 -----------------------------------------------------
def test_torus_knot_closed_with_unit_normals():
    pos, nrm = O.torus_knot(tubular=40, radial=8)
    assert pos.size // 9 == 40 * 8 * 2
    keys = [tuple(np.round(pos[i:i + 3].astype(np.float64), 5)) for i in range(0, pos.size, 3)]
    edges = {}
    for t in range(0, len(keys), 3):
        for e in range(3):
            k = (keys[t + e], keys[t + (e + 1) % 3])
            edges[k] = edges.get(k, 0) + 1
    for (a, b), n in edges.items():
        assert n == 1 and edges.get((b, a)) == 1
    lens = np.linalg.norm(nrm.reshape(-1, 3).astype(np.float64), axis=1)
    assert np.allclose(lens, 1, atol=1e-4)
