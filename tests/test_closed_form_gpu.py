"""The GPU path tracer (through the C ABI) on scenes whose converged
radiance the mathematics fixes (tests/closed_forms.py): exact values where
the estimator is deterministic, z-scores against quadratures of the BSDF
where it is not, and a per-sample brute force of the alpha test folded into
pixel means.  Each scene runs in the default persistent launch and in the
wavefront structure."""
import math

import numpy as np
import pytest

import oracle as O
import closed_forms as CF
from lucent_amd.renderer import Renderer, RenderSettings

pytestmark = pytest.mark.gpu
STRUCTURES = {"units": {}, "wavefront": {"wavefront": True}}


def render(ctx, scene, w, h, spp, max_bounces, seed=1, **kw):
    print(f"render {w}x{h} spp {spp} bounces {max_bounces} {kw}", flush=True)
    r = Renderer(ctx, w, h, RenderSettings(max_bounces=max_bounces, firefly_clamp=0.0, **kw))
    r.load_scene(scene)
    r.film.set_seed(seed)
    r.reset()
    r.trace_samples(spp)
    out = r.film.read()
    r.close()
    return out


@pytest.mark.parametrize("structure", sorted(STRUCTURES))
def test_empty_scene_exact(ctx, structure):
    f = render(ctx, CF.empty_scene(), 40, 24, 4, 8, **STRUCTURES[structure])
    L = np.array(CF.ENV_RGB, np.float32) * np.float32(CF.ENV_INTENSITY)
    assert np.all(f["color"][..., :3] == L)
    assert np.all(f["albedo"][..., 3] == 0)


@pytest.mark.parametrize("structure", sorted(STRUCTURES))
def test_emitter_exact(ctx, structure):
    f = render(ctx, CF.emitter_scene(), 40, 24, 4, 8, **STRUCTURES[structure])
    inner = f["albedo"][..., 3] == 1
    assert inner.sum() > 200
    assert np.all(f["color"][inner][:, :3] == np.array(CF.EMISSION, np.float32))


def z_vs_expected(gpu_mean, per_oracle, expected, label, zmax=5.0):
    ok = ~np.isnan(expected[:, 0])
    spp = per_oracle.shape[1]
    sigma = per_oracle[ok].std(axis=1, ddof=1) / math.sqrt(spp) + 1e-4 * np.abs(expected[ok]) + 1e-6
    z = np.abs(gpu_mean[ok] - expected[ok]) / sigma
    agg_sigma = math.sqrt(float((sigma ** 2).sum())) / ok.sum()
    agg = np.abs(gpu_mean[ok].mean(0) - expected[ok].mean(0))
    print(f"{label}: {ok.sum()} px, max z {z.max():.2f}, aggregate diff {agg} (sigma {agg_sigma:.2e})")
    assert ok.sum() > 20 and z.max() < zmax, label
    assert np.all(agg < 4 * agg_sigma + 2e-4 * np.abs(expected[ok].mean(0))), label


def oracle_per_sample(scene, w, h, spp, max_bounces, xy, seed=1):
    o = O.PtOracle(scene)
    f = o.frame(scene.camera, w, h, max_bounces, 0.0)
    return o.trace_pixels(f, seed, spp, xy, per_sample=True)[1]


@pytest.mark.parametrize("structure", sorted(STRUCTURES))
@pytest.mark.parametrize("name", sorted(CF.SPHERE_MATERIALS))
def test_sphere_directional_albedo(ctx, name, structure):
    mat = CF.SPHERE_MATERIALS[name]
    W, H, SPP = 32, 24, 256
    yy, xx = np.mgrid[4:20:2, 8:24:2]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)
    sc = CF.sphere_scene(mat)
    f = render(ctx, sc, W, H, SPP, 1, **STRUCTURES[structure])
    g = f["color"][xy[:, 1], xy[:, 0], :3].astype(np.float64)
    z_vs_expected(g, oracle_per_sample(sc, W, H, SPP, 1, xy), CF.sphere_expected(mat, W, H, xy),
                  f"sphere {name} ({structure})")


@pytest.mark.parametrize("structure", sorted(STRUCTURES))
def test_plane_under_quad_light(ctx, structure):
    W, H, SPP = 32, 24, 256
    yy, xx = np.mgrid[2:22:3, 2:30:3]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)
    sc = CF.plane_light_scene()
    f = render(ctx, sc, W, H, SPP, 1, **STRUCTURES[structure])
    g = f["color"][xy[:, 1], xy[:, 0], :3].astype(np.float64)
    z_vs_expected(g, oracle_per_sample(sc, W, H, SPP, 1, xy), CF.plane_expected(W, H, xy),
                  f"plane under quad light ({structure})")


@pytest.mark.parametrize("structure", sorted(STRUCTURES))
def test_alpha_mask_brute_force(ctx, structure):
    """Pixel means of the cut-out quad equal the brute-force count of the
    samples the mask stops (the per-sample decisions of the oracle test,
    folded into the film's float32 running mean)."""
    W, H, SPP, SEED = 24, 16, 8, 3
    yy, xx = np.mgrid[0:H, 0:W]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1)
    want = CF.mask_expected(W, H, xy, SEED, SPP)
    f = render(ctx, CF.mask_scene(), W, H, SPP, 4, seed=SEED, **STRUCTURES[structure])
    decided = np.all(want >= 0, axis=1)
    hits = (want == 1).sum(1)
    E = np.array(CF.EMISSION)
    L = np.array(CF.ENV_RGB) * CF.ENV_INTENSITY
    expect = (hits[:, None] * E + (SPP - hits)[:, None] * L) / SPP
    g = f["color"][xy[:, 1], xy[:, 0], :3].astype(np.float64)
    err = np.abs(g[decided] - expect[decided]) / expect[decided]
    assert decided.mean() > 0.8 and err.max() < 1e-5, f"max rel err {err.max():.2e}"
    mixed = (hits > 0) & (hits < SPP) & decided
    assert mixed.sum() > 10
