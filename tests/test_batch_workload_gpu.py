"""The batched multi-scene workload (64 seeded scenes of 1k-1.5M triangles,
mixed resolutions and sample counts, alpha masks on some meshes) under test
on a slice that holds its largest scene (#52, 1,077,448 triangles), a
1920x1080 alpha-masked film (#34), a 640x480 alpha-masked film (#46) and
small and large sample counts: one lucent_trace_batch launch gives every
film bit for bit what one lucent_trace_samples launch per scene gives it
(the plain one-scene-per-launch path, renderer.ts:702-737 being the
reference's structure), with the default record budget (per-sample hand-out
through per-wave record rings) and with one small enough to force the
whole-pixel hand-out; a crop of every film agrees with the oracle."""
import numpy as np
import pytest

from lucent_amd import workloads
from lucent_amd.renderer import Renderer, RenderSettings, trace_batch

from test_pathtrace_gpu import compare, oracle_pixels

pytestmark = pytest.mark.gpu
SLICE = (52, 34, 46, 22, 62, 1, 35)
_cache = {}


def scenes():
    if not _cache:
        for i in SLICE:
            _cache[i] = workloads.batch_scene(i)
    return _cache


def films_batched(ctx, budget):
    settings = RenderSettings(workloads.BATCH_DEPTH, 0.0)
    ctx.set_record_budget(budget)
    try:
        rs = []
        for i, (sc, w, h, spp) in scenes().items():
            r = Renderer(ctx, w, h, settings)
            r.load_scene(sc)
            r.film.set_seed(100 + i)
            r.reset()
            rs.append((r, spp))
        print(f"batch of {len(rs)} films, record budget {budget}", flush=True)
        trace_batch(ctx, rs, settings)
        out = [r.film.read() for r, _ in rs]
        for r, _ in rs:
            r.close()
        return out
    finally:
        ctx.set_record_budget(0)


def films_plain(ctx):
    out = []
    for i, (sc, w, h, spp) in scenes().items():
        print(f"plain launch: scene {i} {w}x{h} spp {spp}", flush=True)
        r = Renderer(ctx, w, h, RenderSettings(workloads.BATCH_DEPTH, 0.0))
        r.load_scene(sc)
        r.film.set_seed(100 + i)
        r.reset()
        r.trace_samples(spp)
        out.append(r.film.read())
        r.close()
    return out


@pytest.fixture(scope="module")
def plain(ctx):
    return films_plain(ctx)


@pytest.mark.parametrize("budget", [0, 1 << 20])
def test_batch_slice_bit_identical_to_plain_launches(ctx, plain, budget):
    got = films_batched(ctx, budget)
    for i, a, b in zip(SLICE, got, plain):
        for k in b:
            assert np.array_equal(a[k], b[k]), f"scene {i}: {k} differs (budget {budget})"


def test_batch_slice_crops_match_oracle(ctx, plain):
    for i, f in zip(SLICE, plain):
        sc, w, h, spp = scenes()[i]
        x0, y0 = w // 2 - 8, h // 2 - 6
        yy, xx = np.mgrid[y0:y0 + 12, x0:x0 + 16]
        xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)
        ref, per = oracle_pixels(None, sc, w, h, spp, xy, seed=100 + i, max_bounces=workloads.BATCH_DEPTH)
        compare(f, ref, per, xy, spp, f"batch scene {i} ({w}x{h}, {spp} spp)")
