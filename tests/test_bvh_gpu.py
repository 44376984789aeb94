"""GPU BVH builders (lucent_build_bvh_gpu, the linear BVH, and
lucent_build_bvh_gpu_clustered, clusters merged by surface area): the reference's node-layout
invariants (bvh.test.ts:31-86: every triangle in exactly one leaf, children
adjacent and nested, boxes containing their triangles); with the LBVH kept as
the device tree (lucent_set_tree_build(0)), exact closest hits through it
equal the FP64 oracle's and the path-traced film agrees with the SAH tree's
and the oracle's; the LBVH as the upload-time device tree; build time plus
traversal-cost ratio against the SAH builder on 1.5M triangles."""
import time

import numpy as np
import pytest

from lucent_amd import scene as S
from lucent_amd import workloads
from lucent_amd.bvh import build_bvh, build_bvh_gpu

pytestmark = pytest.mark.gpu


def soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    return (c + rng.normal(0, 0.2, (n, 3, 3))).astype(np.float32).reshape(-1)


def check_layout(bvh, pos, max_leaf):
    nodes = bvh.nodes.reshape(-1, 8)
    tris = pos.reshape(-1, 3, 3)
    n = tris.shape[0]
    assert sorted(bvh.tri_indices.tolist()) == list(range(n))
    seen = np.zeros(n, np.int64)
    stack = [0]
    reached = 0
    while stack:
        k = stack.pop()
        reached += 1
        lo, hi = nodes[k, :3], nodes[k, 4:7]
        first, count = int(nodes[k, 3]), int(nodes[k, 7])
        if count > 0:
            assert count <= max_leaf
            ids = bvh.tri_indices[first:first + count]
            seen[ids] += 1
            t = tris[ids]
            assert np.all(t.min(axis=(0, 1)) >= lo) and np.all(t.max(axis=(0, 1)) <= hi)
        else:
            assert first > k and first + 1 < bvh.node_count
            for c in (first, first + 1):
                assert np.all(nodes[c, :3] >= lo) and np.all(nodes[c, 4:7] <= hi)
                stack.append(c)
    assert np.all(seen == 1) and reached == bvh.node_count


RADII = [0, 16]  # the linear BVH; the clustered build


@pytest.mark.parametrize("radius", RADII)
@pytest.mark.parametrize("which", ["soup", "knot"])
def test_layout_invariants(ctx, which, radius):
    pos = soup(20000, 3) if which == "soup" else S.torus_knot(2, 3, 1.0, 0.3, 400, 24).positions
    for max_leaf in (1, 4, 8):
        check_layout(build_bvh_gpu(ctx, pos, max_leaf, radius), np.asarray(pos, np.float32), max_leaf)


@pytest.mark.parametrize("radius", RADII + [1, 64])
def test_tiny_inputs(ctx, radius):
    for n in (1, 2, 3, 5, 300):
        pos = soup(n, n)
        b = build_bvh_gpu(ctx, pos, 4, radius)
        check_layout(b, pos, 4)
    # Coincident triangles: equal boxes everywhere, ties at every step.
    same = np.tile(soup(1, 7), 257)
    check_layout(build_bvh_gpu(ctx, same, 2, radius), same, 2)
    b = build_bvh_gpu(ctx, np.zeros(0, np.float32), 4, radius)
    assert b.node_count == 1
    with pytest.raises(ValueError):
        build_bvh_gpu(ctx, np.zeros(10, np.float32), 4, radius)


class caller_tree:
    """lucent_set_tree_build(0): the caller's tree is the device tree, so a
    scene built from a GPU LBVH is traversed through the LBVH itself."""

    def __init__(self, ctx):
        self.ctx = ctx

    def __enter__(self):
        self.ctx.set_tree_build(0, 32)

    def __exit__(self, *exc):
        self.ctx.set_tree_build(2, 32)


@pytest.mark.parametrize("radius", RADII)
def test_lbvh_traversal_matches_the_oracle(ctx, radius):
    """Exact closest hits through the GPU-built tree equal the FP64 oracle's
    (intersectBVH over the reference builder's tree), ties included: the
    exact mode breaks ties by the lowest slot of the caller's tree, and the
    caller's tree is the LBVH here, so ids may differ only at equal t."""
    import oracle as O
    cfg = workloads.config1(1 << 16)
    pos = cfg["positions"]
    lb = build_bvh_gpu(ctx, pos, 4, radius)
    with caller_tree(ctx):
        sc = ctx.scene_from_bvh(pos, lb)
        st = sc.stats()
        hb = sc.intersect(cfg["origins"], cfg["dirs"])
    assert st["node_count"] == lb.node_count
    ref = O.build_bvh(pos)
    t, tri, u, v = O.intersect_bvh_batch(ref, pos, cfg["origins"], cfg["dirs"])
    assert np.array_equal(hb.t.view(np.uint64), t.view(np.uint64))
    same = hb.tri == tri
    tie = ~same & (np.abs(hb.t - t) <= 1e-6 * np.abs(t))
    assert np.all(same | tie)
    assert np.array_equal(hb.u[same], u[same]) and np.array_equal(hb.v[same], v[same])
    assert (tri != 0xFFFFFFFF).mean() > 0.99
    # The LBVH's own leaf order decides ties among its slots: the oracle
    # traversing the LBVH agrees bit for bit.
    t2, tri2, u2, v2 = O.intersect_bvh_batch(lb, pos, cfg["origins"], cfg["dirs"])
    assert np.array_equal(hb.tri, tri2) and np.array_equal(hb.u, u2) and np.array_equal(hb.v, v2)
    sc.destroy()


def test_device_built_upload_tree(ctx):
    """lucent_set_tree_build(leaf, 0): the upload-time device tree is built on
    the GPU; exact hits keep their meaning (ties go to the caller's slots)."""
    import oracle as O
    cfg = workloads.config1(1 << 15, seed=5)
    pos = cfg["positions"]
    ref = O.build_bvh(pos)
    ctx.set_tree_build(2, 0)
    try:
        sc = ctx.scene_from_bvh(pos, build_bvh(pos))
        hb = sc.intersect(cfg["origins"], cfg["dirs"])
    finally:
        ctx.set_tree_build(2, 32)
    t, tri, u, v = O.intersect_bvh_batch(ref, pos, cfg["origins"], cfg["dirs"])
    for a, b in ((hb.t, t), (hb.tri, tri), (hb.u, u), (hb.v, v)):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    sc.destroy()


@pytest.mark.parametrize("radius", RADII)
def test_path_traced_film_through_the_lbvh(ctx, radius):
    """The knot scene's film traced through the LBVH itself (no rebuild)
    equals the one traced through the SAH tree up to equal-distance ties,
    and agrees with the oracle."""
    from lucent_amd.renderer import Renderer, RenderSettings
    sc = workloads.knot_scene()
    w, h = 144, 90
    films = []
    for builder in ("sah", "gpu"):
        r = Renderer(ctx, w, h, RenderSettings(8, 0.0))
        pos = S.merge_meshes([m.mesh for m in sc.meshes]).positions
        if builder == "gpu":
            with caller_tree(ctx):
                lb = build_bvh_gpu(ctx, pos, 4, radius)
                r.load_scene(sc, lb)
                assert r.device_scene.stats()["node_count"] == lb.node_count
        else:
            r.load_scene(sc, build_bvh(pos))
        r.film.set_seed(0)
        r.reset()
        r.trace_samples(4)
        films.append(r.film.read())
        r.close()
    a, b = films[0]["color"], films[1]["color"]
    same = np.all(a == b, axis=-1).mean()
    rel = abs(a[..., :3].mean() - b[..., :3].mean()) / a[..., :3].mean()
    print(f"pixels identical: {same:.5f}, image mean rel diff {rel:.2e}")
    # Different trees visit equal-distance triangles in a different order; nothing else differs.
    assert same > 0.999 and rel < 1e-5
    from test_pathtrace_gpu import compare, oracle_pixels
    yy, xx = np.mgrid[30:54, 50:90]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)
    ref, per = oracle_pixels(None, sc, w, h, 4, xy, seed=0)
    compare(films[1], ref, per, xy, 4, "knots through the LBVH")


def sah_cost(bvh):
    nodes = bvh.nodes.reshape(-1, 8)
    ext = nodes[:, 4:7] - nodes[:, :3]
    area = ext[:, 0] * ext[:, 1] + ext[:, 1] * ext[:, 2] + ext[:, 2] * ext[:, 0]
    leaf = nodes[:, 7] > 0
    return float((area[~leaf] * 1.0).sum() + (area[leaf] * nodes[leaf, 7]).sum()) / float(area[0])


def test_build_time_and_cost_on_1_5m_triangles(ctx):
    mesh = S.uv_sphere(866, 866, 1.0, displace=lambda p, n: 0.05 * np.sin(7 * p[:, 0]) * np.cos(5 * p[:, 2]))
    pos = mesh.positions
    build_bvh_gpu(ctx, pos[: 9 * 1000])  # loads the code objects
    t0 = time.perf_counter()
    g = build_bvh_gpu(ctx, pos)
    t_gpu = time.perf_counter() - t0
    t0 = time.perf_counter()
    c = build_bvh(pos)
    t_cpu = time.perf_counter() - t0
    ratio = sah_cost(g) / sah_cost(c)
    print(f"{pos.size // 9} tris: GPU LBVH {t_gpu * 1e3:.0f} ms (incl. upload and layout), CPU SAH {t_cpu * 1e3:.0f} ms; "
          f"SAH cost ratio LBVH/SAH {ratio:.3f}")
    assert g.node_count > 1 and ratio < 2.0


def test_clustered_build_cost_on_1_5m_triangles(ctx):
    """The clustered GPU build on the 1.5M-triangle sphere: its tree's SAH
    cost against the host SAH builder's and the linear BVH's, and its time."""
    mesh = S.uv_sphere(866, 866, 1.0, displace=lambda p, n: 0.05 * np.sin(7 * p[:, 0]) * np.cos(5 * p[:, 2]))
    pos = mesh.positions
    build_bvh_gpu(ctx, pos[: 9 * 1000], 4, 16)
    t0 = time.perf_counter()
    g = build_bvh_gpu(ctx, pos, 4, 16)
    t_gpu = time.perf_counter() - t0
    lb = build_bvh_gpu(ctx, pos)
    c = build_bvh(pos)
    check_layout(g, np.asarray(pos, np.float32), 4)
    ratio, ratio_lbvh = sah_cost(g) / sah_cost(c), sah_cost(lb) / sah_cost(c)
    print(f"{pos.size // 9} tris: clustered GPU build {t_gpu * 1e3:.0f} ms (incl. upload and layout); SAH cost ratio "
          f"clustered/SAH {ratio:.3f}, LBVH/SAH {ratio_lbvh:.3f}")
    assert ratio < ratio_lbvh and ratio < 1.2
    # The path-tracing benchmark's knots: the clustered tree beats the
    # reference builder's (12 bins) by the SAH measure.
    knots = S.merge_meshes([m.mesh for m in workloads.knot_scene().meshes]).positions
    kc = sah_cost(build_bvh(knots))
    k_ratio, k_lbvh = sah_cost(build_bvh_gpu(ctx, knots, 4, 16)) / kc, sah_cost(build_bvh_gpu(ctx, knots)) / kc
    print(f"knots: SAH cost ratio clustered/SAH {k_ratio:.3f}, LBVH/SAH {k_lbvh:.3f}")
    assert k_ratio < 1.0 < k_lbvh
