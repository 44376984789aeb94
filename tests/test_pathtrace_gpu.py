"""GPU parity of the progressive path tracer against the FP64 oracle.

Both sides draw identical random numbers (same PCG stream per pixel and
sample), so most paths are the same path; they part only where float32 and
float64 round a decision differently (a Russian-roulette draw, a lobe pick,
a grazing hit).  The checks:
  * per-pixel z-score |mean_gpu - mean_oracle| / (s / sqrt(n)) below 6
    (s = the oracle's per-sample standard deviation, floored at 1e-3 of the
    pixel mean plus 1e-6), for every pixel and channel;
  * at least 90% of pixel channels agree to a relative 1e-3;
  * the image mean agrees to a relative 2e-3;
  * guides (albedo, normal, depth stored FP16) within 2e-3 plus the
    rounding the FP16 running mean accumulates (spp * 2^-12 relative) for
    99% of pixels; history equal to the sample count.
"""
import math

import numpy as np
import pytest

import oracle as O
from lucent_amd import scene as S
from lucent_amd import workloads
from lucent_amd.renderer import Renderer, RenderSettings

pytestmark = pytest.mark.gpu


def feature_scene():
    """Every integrator feature the benchmark scenes use, at a small size:
    glass with Beer-Lambert absorption, clearcoat, sheen, metal, a sphere, a
    sphere light, a quad light, the analytic ground and an HDRI with a sun lobe."""
    k1 = S.transform(S.torus_knot(2, 3, 1.0, 0.3, 100, 12), translate=(-1.2, 1.1, 0), rotate_y=0.3)
    k2 = S.transform(S.torus_knot(3, 4, 1.0, 0.3, 100, 12), translate=(1.3, 1.1, 0.2), rotate_y=1.1)
    glass_m = S.Material(albedo=(0.95, 0.9, 0.85), transmission=1.0, roughness=0.05, ior=1.5,
                         attenuation_color=(0.7, 0.9, 0.8), attenuation_distance=1.0)
    coat_m = S.Material(albedo=(0.6, 0.15, 0.1), roughness=0.6, clearcoat=1.0, clearcoat_roughness=0.1,
                        sheen_color=(0.5, 0.5, 0.5), sheen_roughness=0.5)
    lamp = S.quad((-0.5, 3.5, -0.5), (0.5, 3.5, -0.5), (0.5, 3.5, 0.5), (-0.5, 3.5, 0.5))
    w, h = 128, 64
    env = S.Environment(map=(w, h, workloads.synthetic_hdri(w, h, sigma_deg=4.0, peak=300.0)), rotation=20.0,
                        intensity=0.8, background=True)
    return S.Scene(
        camera=S.Camera(position=(0.0, 2.5, 6.0), target=(0.0, 1.0, 0.0), fov=45.0, aperture=0.0, focus_distance=6.0),
        meshes=[S.MeshInstance(k1, glass_m), S.MeshInstance(k2, coat_m),
                S.MeshInstance(lamp, S.light((1.0, 0.9, 0.8), 8.0))],
        spheres=[S.Sphere((0.0, 0.5, 1.5), 0.5, S.metal((0.9, 0.8, 0.6), 0.25)),
                 S.Sphere((2.5, 2.5, -1.5), 0.3, S.light((1.0, 1.0, 1.0), 5.0))],
        env=env, ground=S.Ground(True, 0.0, S.matte((0.5, 0.55, 0.6), 0.9)))


def render(ctx, scene, width, height, spp, seed=0, max_bounces=8):
    r = Renderer(ctx, width, height, RenderSettings(max_bounces=max_bounces, firefly_clamp=0.0))
    r.load_scene(scene)
    r.film.set_seed(seed)
    r.reset()
    r.trace_samples(spp)
    out = r.film.read()
    packed = r.packed
    r.close()
    return out, packed


def oracle_pixels(packed, scene, width, height, spp, xy, seed=0, max_bounces=8):
    """The oracle builds its own packing, tree, textures and lighting from
    `scene`; `packed` (the product's) is not used."""
    o = O.PtOracle(scene)
    f = o.frame(scene.camera, width, height, max_bounces, 0.0)
    return o.trace_pixels(f, seed, spp, xy, per_sample=True)


def compare(gpu, ref, per_sample, xy, spp, label):
    g = gpu["color"][xy[:, 1], xy[:, 0]].astype(np.float64)
    o = ref
    mean_o = o[:, :3]
    mean_g = g[:, :3]
    s = per_sample.std(axis=1, ddof=1) if spp > 1 else np.abs(mean_o)
    sigma = s / math.sqrt(spp) + 1e-3 * np.abs(mean_o) + 1e-6
    z = np.abs(mean_g - mean_o) / sigma
    rel = np.abs(mean_g - mean_o) / (np.abs(mean_o) + 1e-3)
    frac_close = float((rel < 1e-3).mean())
    img_rel = abs(mean_g.mean() - mean_o.mean()) / max(1e-9, mean_o.mean())
    worst = np.unravel_index(np.argmax(z), z.shape)
    msg = (f"{label}: max z {z.max():.2f} at pixel {xy[worst[0]].tolist()} ch {worst[1]} "
           f"(gpu {mean_g[worst]:.6g}, oracle {mean_o[worst]:.6g}); {frac_close:.4f} of channels within 1e-3; "
           f"image mean rel diff {img_rel:.2e}")
    print(msg)
    assert z.max() < 6.0, msg
    assert frac_close > 0.90, msg
    assert img_rel < 2e-3, msg
    # Guides: FP16 storage of running means.  Each of the spp updates rounds
    # the stored mean to half precision (relative 2^-11, half an ulp each),
    # so a pixel whose samples disagree drifts by up to spp * 2^-12 relative.
    ga = gpu["albedo"][xy[:, 1], xy[:, 0]].astype(np.float64)
    gn = gpu["normal"][xy[:, 1], xy[:, 0]].astype(np.float64)
    r16 = 2e-3 + spp * 2.0 ** -12
    ok_a = np.all(np.abs(ga[:, :3] - o[:, 4:7]) <= r16 * np.abs(o[:, 4:7]) + 2e-3, axis=1)
    ok_c = np.abs(ga[:, 3] - o[:, 7]) <= r16
    ok_n = np.all(np.abs(gn[:, :3] - o[:, 8:11]) <= r16, axis=1)
    ok_d = np.abs(gn[:, 3] - o[:, 11]) <= r16 * np.abs(o[:, 11]) + 1e-3
    for name, ok in (("albedo", ok_a), ("coverage", ok_c), ("normal", ok_n), ("depth", ok_d)):
        assert ok.mean() > 0.99, f"{label}: guide {name} agrees on only {ok.mean():.4f} of pixels"
    assert np.all(gpu["history"][xy[:, 1], xy[:, 0]] == spp)


def all_pixels(w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)


def test_feature_scene_matches_oracle(ctx):
    sc = feature_scene()
    W, H, SPP = 96, 64, 8
    gpu, packed = render(ctx, sc, W, H, SPP, seed=11)
    xy = all_pixels(W, H)
    ref, per = oracle_pixels(packed, sc, W, H, SPP, xy, seed=11)
    compare(gpu, ref, per, xy, SPP, "feature scene")


def test_benchmark_scene_crop_matches_oracle(ctx):
    """The 1440x900 benchmark scene: full-film GPU render at 4 spp, checked on
    a 48x32 crop through the knots plus 1500 random pixels."""
    sc = workloads.knot_scene()
    W, H, SPP = workloads.PT_WIDTH, workloads.PT_HEIGHT, 4
    gpu, packed = render(ctx, sc, W, H, SPP, seed=5)
    yy, xx = np.mgrid[400:432, 700:748]
    crop = np.stack([xx.reshape(-1), yy.reshape(-1)], -1)
    rnd = np.random.default_rng(0).integers(0, [W, H], size=(1500, 2))
    xy = np.concatenate([crop, rnd]).astype(np.uint32)
    ref, per = oracle_pixels(packed, sc, W, H, SPP, xy, seed=5)
    compare(gpu, ref, per, xy, SPP, "benchmark crop")


def test_determinism_and_running_mean(ctx):
    """Same seeds give bit-identical films; 4 + 4 samples equal 8 in one call."""
    sc = feature_scene()
    W, H = 64, 48
    a, _ = render(ctx, sc, W, H, 8, seed=3)
    b, _ = render(ctx, sc, W, H, 8, seed=3)
    for k in a:
        assert np.array_equal(a[k], b[k])
    r = Renderer(ctx, W, H, RenderSettings(8, 0.0))
    r.load_scene(sc)
    r.film.set_seed(3)
    r.reset()
    r.trace_samples(4)
    r.trace_samples(4)
    c = r.film.read()
    assert r.film.info()["seed"] == 11 and r.frame == 8
    r.close()
    for k in a:
        assert np.array_equal(a[k], c[k])


def test_reset_keeps_seed_advancing(ctx):
    sc = feature_scene()
    r = Renderer(ctx, 32, 32, RenderSettings(4, 0.0))
    r.load_scene(sc)
    r.film.set_seed(0)
    r.trace_samples(2)
    r.reset()
    assert r.frame == 0 and r.film.info()["seed"] == 2
    r.trace_samples(1)
    h = r.film.read()["history"]
    assert np.all(h == 1)
    r.close()


def _film_after(ctx, scene, w, h, spp, seed, settings):
    r = Renderer(ctx, w, h, settings)
    r.load_scene(scene)
    r.film.set_seed(seed)
    r.reset()
    r.trace_samples(spp)
    out = r.film.read()
    info = r.film.info()
    r.close()
    return out, info


@pytest.mark.parametrize("clamp,occupancy", [(0.0, 2), (8.0, 2), (0.0, 4), (0.0, 6)])
def test_persistent_tiles_bit_identical_to_per_sample_launches(ctx, clamp, occupancy):
    """The persistent tile kernel (all samples of a pixel in one lane, refilled
    as paths end), in each register/occupancy variant, and the reference's
    structure (one launch per sample) give the same film bit for bit,
    including ragged edge tiles (70x45 is not a multiple of 8)."""
    sc = feature_scene()
    a, ia = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, clamp, per_sample_launch=True))
    b, ib = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, clamp, occupancy=occupancy))
    assert ia == ib
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_autotune_picks_a_variant_and_leaves_films_alone(ctx):
    from lucent_amd.renderer import autotune_occupancy as autotune
    r = Renderer(ctx, 48, 40, RenderSettings(8, 0.0))
    r.load_scene(feature_scene())
    r.trace_samples(2)
    before = r.film.read()
    info = r.film.info()
    st = RenderSettings(8, 0.0)
    res = autotune(ctx, [(r, 8)], st, trial_samples=2)
    assert st.occupancy == res["occupancy"] and 2 <= res["occupancy"] <= 6
    assert all(v > 0 for v in res["ms"].values())
    # The scan from 2 waves up keeps a variant unless a higher one is > 1 % faster.
    best = None
    for o in sorted(res["ms"], key=int):
        if best is None or np.float32(res["ms"][o]) < np.float32(res["ms"][best]) * np.float32(0.99):
            best = o
    assert int(best) == res["occupancy"]
    after = r.film.read()
    assert r.film.info() == info
    for k in before:
        assert np.array_equal(before[k], after[k])
    r.close()


def test_launch_autotuner_searches_the_shape_and_leaves_films_alone(ctx):
    """lucent_autotune_launch: occupancy first, then tile, shade batch and
    traversal turns, each candidate kept only when > 1 % faster; the chosen
    shape is one of the candidates and the caller's film is untouched."""
    from lucent_amd.renderer import autotune
    r = Renderer(ctx, 64, 48, RenderSettings(8, 0.0))
    r.load_scene(feature_scene())
    r.trace_samples(2)
    before, info = r.film.read(), r.film.info()
    st = RenderSettings(8, 0.0)
    res = autotune(ctx, [(r, 8)], st, trial_samples=2)
    trials = res["trials"]
    # occupancy 2..6; tiles: sized for 128/512/1024 paths (256 is the start)
    # and five uniform shapes; shade batches; traversal turns.
    assert len(trials) == 5 + 8 + 4 + 3
    assert [t["occupancy"] for t in trials[:5]] == [2, 3, 4, 5, 6]
    assert all(t["tile"] == [0, 0] and t["job_paths"] == 256 for t in trials[:5])
    assert [(*t["tile"], t["job_paths"]) for t in trials[5:13]] == [
        (0, 0, 128), (0, 0, 512), (0, 0, 1024), (8, 8, 0), (8, 4, 0), (4, 4, 0), (4, 2, 0), (2, 2, 0)]
    assert all(t["ms"] > 0 for t in trials)
    chosen = (st.occupancy, st.tile_w, st.tile_h, st.job_paths, st.shade_batch, st.trace_turns)
    assert chosen in [(t["occupancy"], *t["tile"], t["job_paths"], t["shade_batch"], t["trace_turns"])
                      for t in trials]
    print("autotuned", res["occupancy"], res["tile"], res["job_paths"], res["shade_batch"], res["trace_turns"])
    after = r.film.read()
    assert r.film.info() == info
    for k in before:
        assert np.array_equal(before[k], after[k])
    r.close()


# The last four size each film's tile to its sample count (5 and 3 samples
# in the batch below: job_paths 64 gives 4x2 and 4x4 tiles, 1 single
# pixels, 4096 8x8 tiles, the default 256 paths 8x4 and 8x8 tiles).
SHAPES = [dict(tile_w=4, tile_h=4), dict(tile_w=2, tile_h=2), dict(tile_w=8, tile_h=2, shade_batch=16),
          dict(shade_batch=48, trace_turns=3), dict(tile_w=1, tile_h=1, shade_batch=1, trace_turns=1),
          dict(tile_w=0, tile_h=4), dict(job_paths=64), dict(job_paths=1, shade_batch=8), dict(job_paths=4096),
          dict()]


@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_launch_shape_keeps_films_bit_identical(ctx, shape):
    """Every launch shape (job tile, shade batch, traversal turns) gives the
    film of the per-sample launches, for one film and for a batch (whose
    record ring takes the same shape)."""
    from lucent_amd.renderer import trace_batch
    kw = SHAPES[shape]
    sc = feature_scene()
    want, info = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, 0.0, per_sample_launch=True))
    got, gi = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, 0.0, **kw))
    assert gi == info
    for k in want:
        assert np.array_equal(got[k], want[k]), k
    settings = RenderSettings(8, 0.0, **kw)
    rs = []
    for (w, h) in ((64, 40), (33, 17)):
        r = Renderer(ctx, w, h, settings)
        r.load_scene(sc)
        r.film.set_seed(11)
        r.reset()
        rs.append(r)
    trace_batch(ctx, [(rs[0], 5), (rs[1], 3)], settings)
    for r, n in zip(rs, (5, 3)):
        w_, _ = _film_after(ctx, sc, r.film.width, r.film.height, n, 11, RenderSettings(8, 0.0, per_sample_launch=True))
        g_ = r.film.read()
        for k in w_:
            assert np.array_equal(g_[k], w_[k]), (r.film.width, k)
        r.close()


def test_launch_shape_rejects_oversized_tiles(ctx):
    from lucent_amd._native import LucentError
    r = Renderer(ctx, 16, 16, RenderSettings(8, 0.0, tile_w=16, tile_h=8))
    r.load_scene(feature_scene())
    with pytest.raises(LucentError, match="64 pixels"):
        r.trace_samples(1)
    r.close()


@pytest.mark.parametrize("budget", [0, 1])
def test_batch_equals_individual_films(ctx, budget):
    """lucent_trace_batch over several scenes, sizes and sample counts: every
    film equals what lucent_trace_samples makes of it alone; a second batch
    continues each film's seeds and history.  With the default record budget
    the batch hands out pixel samples (per-wave record rings); with a budget
    of 1 byte it hands out whole pixels."""
    from lucent_amd.renderer import trace_batch
    ctx.set_record_budget(budget)
    try:
        _batch_equals_individual_films(ctx)
    finally:
        ctx.set_record_budget(0)


def _batch_equals_individual_films(ctx):
    from lucent_amd.renderer import trace_batch

    scenes = [feature_scene(), workloads.spheres_scene(), feature_scene()]
    sizes = [(64, 40), (48, 48), (33, 17)]
    spps = [(3, 2), (1, 4), (5, 0)]
    settings = RenderSettings(8, 0.0)
    rs = []
    for sc, (w, h) in zip(scenes, sizes):
        r = Renderer(ctx, w, h, settings)
        r.load_scene(sc)
        r.film.set_seed(11)
        r.reset()
        rs.append(r)
    trace_batch(ctx, [(r, n[0]) for r, n in zip(rs, spps)], settings)
    trace_batch(ctx, [(r, n[1]) for r, n in zip(rs, spps)], settings)
    for r, sc, (w, h), n in zip(rs, scenes, sizes, spps):
        got = r.film.read()
        want, info = _film_after(ctx, sc, w, h, n[0] + n[1], 11, RenderSettings(8, 0.0, per_sample_launch=True))
        assert r.film.info() == info
        for k in want:
            assert np.array_equal(got[k], want[k]), (w, h, k)
    from lucent_amd._native import LucentError
    with pytest.raises(LucentError, match="twice"):
        trace_batch(ctx, [(rs[0], 1), (rs[0], 1)], settings)
    for r in rs:
        r.close()


def alpha_scene():
    """Alpha-tested geometry as the batched workload has it: a Worley-masked
    (mask mode, cutoff 0.5) knot in front of a textured-alpha quad and a lamp,
    so that closest hits and shadow rays both pass through the holes."""
    tex = workloads.worley_alpha(64, cells=4, seed=5)
    knot = S.transform(S.torus_knot(2, 3, 1.0, 0.35, 120, 16, uvs=True), translate=(0.0, 1.2, 0.0))
    masked = S.Material(albedo=(0.8, 0.5, 0.3), roughness=0.5, alpha_mode="mask", alpha_cutoff=0.5, base_color_map=0)
    wall = S.quad((-3.0, 0.0, -2.0), (3.0, 0.0, -2.0), (3.0, 3.5, -2.0), (-3.0, 3.5, -2.0))
    wall.uvs = np.tile(np.array([0, 0, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1], np.float32) * 2.0, 1)
    cut = S.Material(albedo=(0.3, 0.6, 0.8), roughness=0.8, alpha_mode="mask", alpha_cutoff=0.5, base_color_map=0)
    lamp = S.quad((-0.5, 3.8, -0.5), (0.5, 3.8, -0.5), (0.5, 3.8, 0.5), (-0.5, 3.8, 0.5))
    w, h = 128, 64
    env = S.Environment(map=(w, h, workloads.synthetic_hdri(w, h, sigma_deg=4.0, peak=300.0)), rotation=10.0,
                        intensity=1.0, background=True)
    return S.Scene(
        camera=S.Camera(position=(0.5, 2.0, 5.5), target=(0.0, 1.2, 0.0), fov=45.0, aperture=0.0, focus_distance=5.5),
        meshes=[S.MeshInstance(knot, masked), S.MeshInstance(wall, cut),
                S.MeshInstance(lamp, S.light((1.0, 0.9, 0.8), 8.0))],
        env=env, ground=S.Ground(True, 0.0, S.matte((0.5, 0.55, 0.6), 0.9)), textures=[tex])


def test_alpha_masked_scene_matches_oracle(ctx):
    sc = alpha_scene()
    W, H, spp = 48, 32, 16
    gpu, packed = render(ctx, sc, W, H, spp)
    assert packed.color_layers == [0]
    assert np.count_nonzero(packed.tri_geometry.reshape(-1, 12)[:, 11]) == packed.tri_count - 2
    xy = all_pixels(W, H)
    ref, per = oracle_pixels(packed, sc, W, H, spp, xy)
    compare(gpu, ref, per, xy, spp, "alpha-masked scene")


def test_batch_with_alpha_scene_bit_identical(ctx):
    from lucent_amd.renderer import trace_batch

    scenes = [alpha_scene(), workloads.batch_scene(1)[0]]
    sizes = [(40, 24), (64, 48)]
    settings = RenderSettings(8, 0.0)
    rs = []
    for sc, (w, h) in zip(scenes, sizes):
        r = Renderer(ctx, w, h, settings)
        r.load_scene(sc)
        rs.append(r)
    trace_batch(ctx, [(rs[0], 6), (rs[1], 3)], settings)
    for r, sc, (w, h), n in zip(rs, scenes, sizes, (6, 3)):
        want, _ = _film_after(ctx, sc, w, h, n, 0, RenderSettings(8, 0.0, per_sample_launch=True))
        got = r.film.read()
        for k in want:
            assert np.array_equal(got[k], want[k]), k
        r.close()


def blend_scene():
    """alpha_scene with the knot alpha-blended (blend mode draws a random
    number per candidate hit inside traversal)."""
    import dataclasses
    sc = alpha_scene()
    m = sc.meshes[0]
    sc.meshes[0] = S.MeshInstance(m.mesh, dataclasses.replace(m.material, alpha_mode="blend", alpha=0.6))
    return sc


@pytest.mark.parametrize("which,occupancy", [("feature", 3), ("feature", 4), ("knots", 3), ("alpha", 3), ("blend", 3)])
def test_fused_traversal_loop_bit_identical_to_inline_shadows(ctx, which, occupancy):
    """The tile kernel's default structure (a lane's closest-hit and shadow
    rays through one traversal loop, shading in batches of lanes) and the
    inline structure (each shadow ray traced where the integrator asks for
    it) both give the per-sample launches' film bit for bit.  A scene with an
    alpha-blended material takes the inline structure."""
    sc = {"feature": feature_scene, "knots": workloads.knot_scene, "alpha": alpha_scene, "blend": blend_scene}[which]()
    w, h, spp = (144, 90, 3) if which == "knots" else (70, 45, 6)
    a, ia = _film_after(ctx, sc, w, h, spp, 4, RenderSettings(8, 0.0, per_sample_launch=True))
    b, ib = _film_after(ctx, sc, w, h, spp, 4, RenderSettings(8, 0.0, occupancy=occupancy))
    c, ic = _film_after(ctx, sc, w, h, spp, 4, RenderSettings(8, 0.0, occupancy=occupancy, inline_shadows=True))
    assert ia == ib == ic
    for k in a:
        assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a[k], c[k]), k


@pytest.mark.parametrize("w,h,spp", [(1, 1, 9), (3, 2, 5), (9, 1, 3), (2, 2, 130)])
def test_tiny_films_in_every_structure(ctx, w, h, spp):
    """Films smaller than a tile (one pixel, a partial tile, many samples on
    four pixels): the per-sample hand-out, the pixel hand-out and the inline
    structure give the per-sample launches' film bit for bit."""
    sc = feature_scene()
    a, ia = _film_after(ctx, sc, w, h, spp, 1, RenderSettings(8, 0.0, per_sample_launch=True))
    for st in (RenderSettings(8, 0.0), RenderSettings(8, 0.0, inline_shadows=True)):
        b, ib = _film_after(ctx, sc, w, h, spp, 1, st)
        assert ia == ib
        for k in a:
            assert np.array_equal(a[k], b[k]), k


def test_record_budget_splits_the_launch(ctx):
    """A single-film launch whose path records exceed the record budget runs
    as several launches, each continuing the film (here one sample per
    launch, then three): the film is still bit-identical to the per-sample
    launch structure's."""
    sc = feature_scene()
    w, h, spp = 70, 45, 7
    per_sample = ((w + 7) // 8) * ((h + 7) // 8) * 64 * 48
    a, ia = _film_after(ctx, sc, w, h, spp, 5, RenderSettings(8, 0.0, per_sample_launch=True))
    for budget in (1, 3 * per_sample):
        ctx.set_record_budget(budget)
        try:
            b, ib = _film_after(ctx, sc, w, h, spp, 5, RenderSettings(8, 0.0))
        finally:
            ctx.set_record_budget(0)
        assert ia == ib
        for k in a:
            assert np.array_equal(a[k], b[k]), (budget, k)


def test_scene_without_triangles_in_every_structure(ctx):
    """Spheres and the ground only: rays start with an empty traversal stack
    (nothing to pop) in the per-sample, per-sample-hand-out, tile and
    wavefront structures, which all give the same film."""
    sc = S.Scene(camera=S.Camera(position=(0.0, 1.5, 6.0), target=(0.0, 1.0, 0.0), fov=45.0, aperture=0.0,
                                 focus_distance=6.0),
                 meshes=[], spheres=[S.Sphere((0.0, 1.0, 0.0), 1.0, S.matte((0.7, 0.3, 0.2), 0.4)),
                                     S.Sphere((1.8, 0.6, 0.5), 0.6, S.light((1.0, 0.9, 0.8), 6.0))],
                 ground=S.Ground(True, 0.0, S.matte((0.5, 0.55, 0.6), 0.9)))
    a, ia = _film_after(ctx, sc, 40, 24, 4, 3, RenderSettings(8, 0.0, per_sample_launch=True))
    for st in (RenderSettings(8, 0.0), RenderSettings(8, 0.0, inline_shadows=True), RenderSettings(8, 0.0, wavefront=True)):
        b, ib = _film_after(ctx, sc, 40, 24, 4, 3, st)
        assert ia == ib
        for k in a:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("clamp", [0.0, 8.0])
def test_wavefront_bit_identical_to_per_sample_launches(ctx, clamp):
    """The wavefront structure (separate closest-hit, shadow and shading
    passes over a path queue, persistent traversal with lane refill) gives
    the per-sample launches' film bit for bit: same traversal order, same
    shading code, same random numbers."""
    sc = feature_scene()
    a, ia = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, clamp, per_sample_launch=True))
    b, ib = _film_after(ctx, sc, 70, 45, 7, 5, RenderSettings(8, clamp, wavefront=True))
    assert ia == ib
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("which", ["alpha", "knots"])
def test_wavefront_bit_identical_on_alpha_and_benchmark_scenes(ctx, which):
    sc = alpha_scene() if which == "alpha" else workloads.knot_scene()
    w, h, spp = (40, 24, 6) if which == "alpha" else (144, 90, 3)
    a, _ = _film_after(ctx, sc, w, h, spp, 2, RenderSettings(8, 0.0, per_sample_launch=True))
    b, _ = _film_after(ctx, sc, w, h, spp, 2, RenderSettings(8, 0.0, wavefront=True))
    for k in a:
        assert np.array_equal(a[k], b[k]), k


# ---------------------------------------------------------------------------
# Lighting and surface features against the oracle: the procedural sky with
# the analytic sun (renderer.ts:355-385, environment.ts:106-132,
# lights.glsl:81-97), the firefly clamp at the reference default
# (trace.frag:65-69), the texture-driven surface (trace.frag:78-110), and
# the wavefront structure compared with the oracle directly.
# ---------------------------------------------------------------------------
def sky_scene():
    sc = feature_scene()
    sc.env = S.Environment(map=None, background=True)
    sc.sky = S.Sky(True, (0.25, 0.45, 0.9), (0.9, 0.85, 0.8), (0.25, 0.22, 0.2), 1.0)
    sc.sun = S.Sun(elevation=38.0, azimuth=25.0, color=(1.0, 0.95, 0.85), intensity=3.0, angular_radius=2.0)
    return sc


def textured_feature_scene():
    """Every texture-driven branch of surfaceAt: a checker ground, a knot
    with base-colour, metal-rough, normal and emissive maps, a knot with
    only a normal map (sRGB and linear layers both in use)."""
    n = 64
    yy, xx = np.mgrid[0:n, 0:n] / n
    base = np.zeros((n, n, 4), np.uint8)
    base[..., 0] = (128 + 100 * np.sin(6.283 * xx)).astype(np.uint8)
    base[..., 1] = (128 + 100 * np.cos(6.283 * 2 * yy)).astype(np.uint8)
    base[..., 2] = 90
    base[..., 3] = 255
    mr = np.zeros((n, n, 4), np.uint8)
    mr[..., 1] = (60 + 190 * xx).astype(np.uint8)   # roughness factor
    mr[..., 2] = (255 * (yy > 0.5)).astype(np.uint8)  # metallic factor
    mr[..., 3] = 255
    bump = 0.35 * np.sin(6.283 * 3 * xx)
    nx, nz = -bump, np.ones_like(bump)
    ln = np.sqrt(nx * nx + nz * nz)
    nmap = np.zeros((n, n, 4), np.uint8)
    nmap[..., 0] = np.rint((nx / ln * 0.5 + 0.5) * 255).astype(np.uint8)
    nmap[..., 1] = 128
    nmap[..., 2] = np.rint((nz / ln * 0.5 + 0.5) * 255).astype(np.uint8)
    nmap[..., 3] = 255
    emis = np.zeros((n, n, 4), np.uint8)
    emis[..., 0] = (255 * ((xx * 8).astype(int) % 2)).astype(np.uint8)
    emis[..., 1] = 40
    emis[..., 3] = 255
    k1 = S.transform(S.torus_knot(2, 3, 1.0, 0.3, 100, 12, uvs=True), translate=(-1.2, 1.1, 0), rotate_y=0.3)
    k2 = S.transform(S.torus_knot(3, 4, 1.0, 0.3, 100, 12, uvs=True), translate=(1.3, 1.1, 0.2), rotate_y=1.1)
    m1 = S.Material(albedo=(1.0, 1.0, 1.0), roughness=0.7, metallic=1.0, base_color_map=0, metal_rough_map=1,
                    normal_map=2, emissive_map=3, emission=(0.6, 0.6, 0.6), normal_scale=0.8,
                    uv_transform=(2.0, 0.0, 0.1, 0.0, 1.0, 0.0))
    m2 = S.Material(albedo=(0.3, 0.5, 0.7), roughness=0.35, normal_map=2, clearcoat=0.5, clearcoat_roughness=0.1)
    ground = S.Material(albedo=(0.8, 0.8, 0.8), albedo2=(0.1, 0.15, 0.2), checker_scale=2.0, roughness=0.9)
    sc = feature_scene()
    sc.meshes = [S.MeshInstance(k1, m1), S.MeshInstance(k2, m2), sc.meshes[2]]
    sc.ground = S.Ground(True, 0.0, ground)
    sc.textures = [base, mr, nmap, emis]
    return sc


def test_sky_and_sun_scene_matches_oracle(ctx):
    sc = sky_scene()
    W, H, SPP = 96, 64, 8
    gpu, _ = render(ctx, sc, W, H, SPP, seed=13)
    xy = all_pixels(W, H)
    ref, per = oracle_pixels(None, sc, W, H, SPP, xy, seed=13)
    compare(gpu, ref, per, xy, SPP, "sky + sun")


def test_firefly_clamp_default_matches_oracle(ctx):
    sc = feature_scene()
    W, H, SPP = 96, 64, 8
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=8, firefly_clamp=8.0))
    r.load_scene(sc)
    r.film.set_seed(17)
    r.reset()
    r.trace_samples(SPP)
    gpu = r.film.read()
    r.close()
    xy = all_pixels(W, H)
    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, 8, 8.0)
    ref, per = o.trace_pixels(f, 17, SPP, xy, per_sample=True)
    compare(gpu, ref, per, xy, SPP, "clamp 8")
    # The clamp bites: the unclamped film differs.
    f0 = o.frame(sc.camera, W, H, 8, 0.0)
    ref0, _ = o.trace_pixels(f0, 17, SPP, xy)
    assert np.abs(ref0[:, :3] - ref[:, :3]).max() > 1e-2


def test_textured_surface_matches_oracle(ctx):
    sc = textured_feature_scene()
    W, H, SPP = 96, 64, 8
    gpu, _ = render(ctx, sc, W, H, SPP, seed=19)
    xy = all_pixels(W, H)
    ref, per = oracle_pixels(None, sc, W, H, SPP, xy, seed=19)
    compare(gpu, ref, per, xy, SPP, "textured surface")


@pytest.mark.parametrize("which", ["feature", "alpha"])
def test_wavefront_matches_oracle(ctx, which):
    sc = feature_scene() if which == "feature" else alpha_scene()
    W, H, SPP = 96, 64, 8
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=8, firefly_clamp=0.0, wavefront=True))
    r.load_scene(sc)
    r.film.set_seed(23)
    r.reset()
    r.trace_samples(SPP)
    gpu = r.film.read()
    r.close()
    xy = all_pixels(W, H)
    ref, per = oracle_pixels(None, sc, W, H, SPP, xy, seed=23)
    compare(gpu, ref, per, xy, SPP, f"wavefront {which}")
