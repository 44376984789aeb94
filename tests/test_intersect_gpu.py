"""GPU parity of batched intersectBVH against the FP64 oracle.

Exact mode must reproduce the CPU traversal's hit records bit for bit
(t, u, v as float64, original triangle id); at an exact tie the id may only
differ where |Δt| <= 1e-6·t (none are expected on these inputs).
"""
import math

import numpy as np
import pytest

import oracle as O
from lucent_amd import workloads
from lucent_amd.bvh import build_bvh

pytestmark = pytest.mark.gpu
MISS = 0xFFFFFFFF


def normalize(v):
    v = np.asarray(v, np.float64)
    return v / math.sqrt(float(v @ v))


def assert_identical(gpu, ref, label):
    t, tri, u, v = ref
    same_id = gpu.tri == tri
    bad = ~(same_id & (gpu.t.view(np.uint64) == t.view(np.uint64)) & (gpu.u.view(np.uint64) == u.view(np.uint64))
            & (gpu.v.view(np.uint64) == v.view(np.uint64)))
    # Ties: a different id is tolerated only at equal t within 1e-6 relative.
    both = (gpu.tri != MISS) & (tri != MISS)
    dt = np.abs(np.where(both, gpu.t, 0.0) - np.where(both, t, 0.0))
    tie = (~same_id) & both & (dt <= 1e-6 * np.abs(np.where(both, t, 0.0)))
    hard = bad & ~tie
    assert not hard.any(), (f"{label}: {int(hard.sum())} of {len(t)} hit records differ; first at "
                            f"{int(np.flatnonzero(hard)[0])}: gpu=({gpu.t[hard][0]}, {gpu.tri[hard][0]}) "
                            f"cpu=({t[hard][0]}, {tri[hard][0]})")
    return int(tie.sum())


def reference_soup_case():
    rand = O.Lcg(42)
    soup = O.random_soup(2000, rand)
    org = np.zeros((400, 3))
    dirs = np.zeros((400, 3))
    for i in range(400):
        org[i] = [rand() * 16 - 8, rand() * 16 - 8, rand() * 16 - 8]
        dirs[i] = normalize([rand() - 0.5, rand() - 0.5, rand() - 0.5])
    return soup, org, dirs


def test_reference_soup_double_rays(ctx):
    soup, org, dirs = reference_soup_case()
    b = build_bvh(soup)
    sc = ctx.scene_from_bvh(soup, b)
    gpu = sc.intersect(org, dirs)
    ref = O.intersect_bvh_batch(b, soup, org, dirs)
    assert_identical(gpu, ref, "soup/f64")
    assert (gpu.tri != MISS).sum() > 50
    # The brute force agrees too (bvh.test.ts:90-110).
    bf = O.intersect_all_batch(soup, org, dirs)
    assert np.array_equal(gpu.tri, bf[1])


def test_reference_soup_float_rays(ctx):
    soup, org, dirs = reference_soup_case()
    org32, dirs32 = org.astype(np.float32), dirs.astype(np.float32)
    b = build_bvh(soup)
    sc = ctx.scene_from_bvh(soup, b)
    assert_identical(sc.intersect(org32, dirs32), O.intersect_bvh_batch(b, soup, org32, dirs32), "soup/f32")


def test_torus_knot_exact_t(ctx):
    pos, _ = O.torus_knot(tubular=80, radial=10)
    rand = O.Lcg(99)
    org = np.zeros((200, 3))
    dirs = np.zeros((200, 3))
    for i in range(200):
        org[i] = [rand() * 6 - 3, rand() * 6 - 3, 4]
        dirs[i] = normalize([rand() * 0.4 - 0.2, rand() * 0.4 - 0.2, -1])
    b = build_bvh(pos)
    gpu = ctx.scene_from_bvh(pos, b).intersect(org, dirs)
    assert_identical(gpu, O.intersect_bvh_batch(b, pos, org, dirs), "knot")


def test_degenerate_and_tiny_trees(ctx):
    one = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0] * 64, np.float32)
    b = build_bvh(one)
    gpu = ctx.scene_from_bvh(one, b).intersect(np.array([[0.2, 0.2, -1.0]]), np.array([[0.0, 0.0, 1.0]]))
    assert gpu.t[0] == 1.0 and gpu.tri[0] == b.tri_indices[0]  # lowest slot wins the 64-way tie
    single = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    b = build_bvh(single)
    gpu = ctx.scene_from_bvh(single, b).intersect(np.array([[0.2, 0.2, -1.0], [2, 2, -1]]),
                                                  np.array([[0.0, 0.0, 1.0], [0, 0, 1]]))
    assert gpu.tri.tolist() == [0, MISS] and gpu.t[1] == math.inf


def test_rebuilt_device_tree_keeps_hit_records(ctx):
    """The device tree rebuilt at upload (default) and the caller's own tree
    give bit-identical exact hit records, ties included (64 coincident
    triangles: the caller's lowest slot wins either way)."""
    soup, org, dirs = reference_soup_case()
    b = build_bvh(soup)
    one = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0] * 64, np.float32)
    b1 = build_bvh(one)
    res = {}
    try:
        for leaf in (0, 2, 1):
            ctx.set_tree_build(leaf, 32)
            res[leaf] = (ctx.scene_from_bvh(soup, b).intersect(org, dirs),
                         ctx.scene_from_bvh(one, b1).intersect(np.array([[0.2, 0.2, -1.0], [0.1, 0.3, 2.0]]),
                                                               np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])))
    finally:
        ctx.set_tree_build(2, 32)
    for leaf in (2, 1):
        for got, ref in zip(res[leaf], res[0]):
            for f in ("t", "tri", "u", "v"):
                assert np.array_equal(getattr(got, f), getattr(ref, f)), (leaf, f)
    assert res[2][1].tri.tolist() == [b1.tri_indices[0]] * 2


def test_empty_scene_misses(ctx):
    empty = np.zeros(0, np.float32)
    b = build_bvh(empty)
    gpu = ctx.scene_from_bvh(empty, b).intersect(np.array([[0.0, 0, 0]]), np.array([[0.0, 0, 1]]))
    assert gpu.tri[0] == MISS and gpu.t[0] == math.inf


def test_tmin_tmax_bounds(ctx):
    soup, org, dirs = reference_soup_case()
    b = build_bvh(soup)
    sc = ctx.scene_from_bvh(soup, b)
    full = O.intersect_bvh_batch(b, soup, org, dirs)
    hit = full[1] != MISS
    # tmax is one scalar per batch: use the median hit distance for all rays.
    t_lim = float(np.median(full[0][hit]))
    assert_identical(sc.intersect(org, dirs, 1e-5, t_lim), O.intersect_bvh_batch(b, soup, org, dirs, 1e-5, t_lim),
                     "tmax")
    assert_identical(sc.intersect(org, dirs, 2.0, math.inf), O.intersect_bvh_batch(b, soup, org, dirs, 2.0, math.inf),
                     "tmin")


def test_spill_path_gives_identical_results(ctx):
    from lucent_amd import _native
    import ctypes as C
    pos, _ = O.torus_knot()
    b = build_bvh(pos)
    sc = ctx.scene_from_bvh(pos, b)
    o, d = workloads.sphere_rays(20000, 7)
    o = o * np.float32(1.5)
    ref = O.intersect_bvh_batch(b, pos, o, d)
    n = o.shape[0]
    t = np.empty(n)
    u = np.empty(n)
    v = np.empty(n)
    tri = np.empty(n, np.uint32)
    o = np.ascontiguousarray(o)
    d = np.ascontiguousarray(d)
    _native.check(_native.lib().lucent_intersect(ctx.handle, sc.handle, o.ctypes.data, d.ctypes.data, n, 1e-5,
                                                 math.inf, 0x100, t.ctypes.data, tri.ctypes.data, u.ctypes.data,
                                                 v.ctypes.data))

    class H:
        pass
    h = H()
    h.t, h.tri, h.u, h.v = t, tri, u, v
    assert_identical(h, ref, "spill")


def test_any_hit_agrees_on_occlusion(ctx):
    w = workloads.config1(1 << 14, seed=3)
    b = build_bvh(w["positions"])
    sc = ctx.scene_from_bvh(w["positions"], b)
    closest = sc.intersect(w["origins"], w["dirs"])
    anyh = sc.intersect(w["origins"], w["dirs"], any_hit=True)
    assert np.array_equal(closest.tri != MISS, anyh.tri != MISS)
    # Any returned hit is a real hit no nearer than the closest one.
    m = anyh.tri != MISS
    assert (anyh.t[m] >= closest.t[m]).all()


def test_glsl_mode_close_to_exact(ctx):
    w = workloads.config1(1 << 16, seed=4)
    b = build_bvh(w["positions"])
    sc = ctx.scene_from_bvh(w["positions"], b)
    ex = sc.intersect(w["origins"], w["dirs"])
    gl = sc.intersect(w["origins"], w["dirs"], precision="glsl")
    both = (ex.tri != MISS) & (gl.tri != MISS)
    assert both.mean() > 0.99
    assert (ex.tri[both] == gl.tri[both]).mean() > 0.999
    assert np.allclose(gl.t[both], ex.t[both], rtol=1e-5)


def test_config1_full_batch_bit_identical(ctx):
    """2^20 rays against the 81,920-triangle displaced icosphere (the bench workload)."""
    w = workloads.config1()
    b = build_bvh(w["positions"])
    sc = ctx.scene_from_bvh(w["positions"], b)
    gpu = sc.intersect(w["origins"], w["dirs"])
    ref = O.intersect_bvh_batch(b, w["positions"], w["origins"], w["dirs"])
    ties = assert_identical(gpu, ref, "config1")
    assert ties == 0
    assert (gpu.tri != MISS).mean() > 0.99
