"""GPU linear-BVH builder (lucent_build_bvh_gpu): the reference's node-layout
invariants (bvh.test.ts:31-86: every triangle in exactly one leaf, children
adjacent and nested, boxes containing their triangles), exact closest hits
bit-identical to those over the SAH tree (exact mode does not depend on the
tree), path-traced films equal to the SAH tree's, and build time plus
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


@pytest.mark.parametrize("which", ["soup", "knot"])
def test_layout_invariants(ctx, which):
    pos = soup(20000, 3) if which == "soup" else S.torus_knot(2, 3, 1.0, 0.3, 400, 24).positions
    for max_leaf in (1, 4, 8):
        check_layout(build_bvh_gpu(ctx, pos, max_leaf), np.asarray(pos, np.float32), max_leaf)


def test_tiny_inputs(ctx):
    for n in (1, 2, 3, 5):
        pos = soup(n, n)
        b = build_bvh_gpu(ctx, pos, 4)
        check_layout(b, pos, 4)
    b = build_bvh_gpu(ctx, np.zeros(0, np.float32))
    assert b.node_count == 1
    with pytest.raises(ValueError):
        build_bvh_gpu(ctx, np.zeros(10, np.float32))


def test_exact_hits_identical_to_the_sah_tree(ctx):
    cfg = workloads.config1(1 << 16)
    pos = cfg["positions"]
    a = ctx.scene_from_bvh(pos, build_bvh(pos))
    b = ctx.scene_from_bvh(pos, build_bvh_gpu(ctx, pos))
    ha = a.intersect(cfg["origins"], cfg["dirs"])
    hb = b.intersect(cfg["origins"], cfg["dirs"])
    for f in ("t", "tri", "u", "v"):
        assert np.array_equal(getattr(ha, f), getattr(hb, f)), f
    a.destroy()
    b.destroy()


def test_path_traced_film_matches_the_sah_tree(ctx):
    from lucent_amd.renderer import Renderer, RenderSettings
    sc = workloads.knot_scene()
    w, h = 144, 90
    films = []
    for builder in ("sah", "gpu"):
        r = Renderer(ctx, w, h, RenderSettings(8, 0.0))
        pos = S.merge_meshes([m.mesh for m in sc.meshes]).positions
        r.load_scene(sc, build_bvh(pos) if builder == "sah" else build_bvh_gpu(ctx, pos))
        r.film.set_seed(0)
        r.reset()
        r.trace_samples(4)
        films.append(r.film.read()["color"])
        r.close()
    same = np.all(films[0] == films[1], axis=-1).mean()
    rel = abs(films[0][..., :3].mean() - films[1][..., :3].mean()) / films[0][..., :3].mean()
    print(f"pixels identical: {same:.5f}, image mean rel diff {rel:.2e}")
    # Different trees visit equal-distance triangles in a different order; nothing else differs.
    assert same > 0.999 and rel < 1e-5


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
