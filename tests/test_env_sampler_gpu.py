"""Environment importance sampler built on the device (lucent_env_sampler_gpu)
against the host build (lucent_env_sampler, itself bit-identical to the
oracle's buildEnvSampler restatement): grid, total and every table texel
bit-identical on maps that exercise box averaging, uneven boxes, tall maps,
a single texel, black maps (uniform fallback), one dominant texel (a long
Vose chain) and zero rows; the host's error on a bad texel with the output
left untouched; and a scene whose sampler is rebuilt (by the host build or
on the device) after a texel update renders bit-identically to one set up
with the new map."""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

import oracle as O
from lucent_amd import _native, workloads
from lucent_amd.diff import DiffRenderer
from lucent_amd.scene import Environment

pytestmark = pytest.mark.gpu


def host_sampler(rgb, w, h, max_w):
    table = np.zeros(min(w, max_w) * h * 4 + 4, np.float32)
    gw, gh, total = C.c_uint32(), C.c_uint32(), C.c_double()
    _native.check(_native.lib().lucent_env_sampler(np.ascontiguousarray(rgb, np.float32).ctypes.data, w, h, max_w,
                                                   table.ctypes.data, C.byref(gw), C.byref(gh), C.byref(total)))
    return table[: gw.value * gh.value * 4], gw.value, gh.value, total.value


def rgba_tensor(rgb, w, h):
    import torch
    a = np.zeros((h, w, 4), np.float32)
    a[..., :3] = np.asarray(rgb, np.float32).reshape(h, w, 3)
    return torch.from_numpy(a).to("cuda:0")


def make_map(kind):
    rng = np.random.default_rng(11)
    if kind == "hdri_1024":
        return workloads.synthetic_hdri(1024, 512, sigma_deg=3.0), 1024, 512, 1024
    if kind == "hdri_2048_boxes":
        return workloads.synthetic_hdri(2048, 1024, sigma_deg=2.0), 2048, 1024, 1024
    if kind == "odd":
        return rng.random(50 * 31 * 3).astype(np.float32), 50, 31, 20
    if kind == "tall":
        return rng.random(16 * 64 * 3).astype(np.float32) ** 4, 16, 64, 1024
    if kind == "one_texel":
        return np.float32([0.5, 2.0, 1.0]), 1, 1, 1024
    if kind == "black":
        return np.zeros(64 * 32 * 3, np.float32), 64, 32, 1024
    if kind == "one_hot":
        m = np.full((128, 256, 3), 1e-3, np.float32)
        m[40, 77] = 5e4
        return m.reshape(-1), 256, 128, 1024
    if kind == "zero_rows":
        m = rng.random((64, 128, 3)).astype(np.float32)
        m[:20] = 0
        m[50:, :, 1] = 0
        return m.reshape(-1), 128, 64, 1024
    raise ValueError(kind)


KINDS = ["hdri_1024", "hdri_2048_boxes", "odd", "tall", "one_texel", "black", "one_hot", "zero_rows"]


@pytest.mark.parametrize("kind", KINDS)
def test_device_sampler_bit_identical_to_host(ctx, kind):
    rgb, w, h, max_w = make_map(kind)
    print(f"env sampler {kind}: {w}x{h}, max width {max_w}", flush=True)
    ref, rgw, rgh, rtotal = host_sampler(rgb, w, h, max_w)
    table, gw, gh, total, ms = ctx.env_sampler_gpu(rgba_tensor(rgb, w, h), max_w)
    got = table.cpu().numpy().reshape(-1)
    assert (gw, gh) == (rgw, rgh)
    assert total == rtotal
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    if kind in ("odd", "one_hot"):
        # The oracle's restatement of buildEnvSampler agrees too.
        otable, ogw, ogh, ototal = O.env_sampler(rgb, w, h, max_w)
        assert (ogw, ogh, ototal) == (gw, gh, total)
        assert np.array_equal(np.asarray(otable, np.float32).reshape(-1).view(np.uint32), got.view(np.uint32))
    print(f"  grid {gw}x{gh}, device build {ms:.3f} ms", flush=True)


def test_bad_texel_fails_like_the_host_and_leaves_the_table(ctx):
    import torch
    rgb, w, h, _ = make_map("odd")
    bad = rgb.copy()
    bad[(3 * 50 + 7) * 3 + 1] = -1.0
    with pytest.raises(_native.LucentError) as host_err:
        host_sampler(bad, w, h, 1024)
    t = rgba_tensor(bad, w, h)
    table = torch.full((w * h, 4), 7.0, dtype=torch.float32, device="cuda:0")
    gw, gh, total, ms = C.c_uint32(), C.c_uint32(), C.c_double(), C.c_float()
    with pytest.raises(_native.LucentError) as dev_err:
        _native.check(_native.lib().lucent_env_sampler_gpu(
            ctx.handle, C.c_void_p(t.data_ptr()), w, h, 1024, C.c_void_p(table.data_ptr()), C.byref(gw), C.byref(gh),
            C.byref(total), C.byref(ms), None))
    assert str(dev_err.value) == str(host_err.value)
    assert "weight 157 is negative or not finite" in str(dev_err.value)
    assert bool((table == 7.0).all())


@pytest.mark.parametrize("on_device", [False, True])
def test_scene_rebuild_after_texel_update_matches_a_fresh_scene(ctx, on_device):
    sc = workloads.spheres_scene()
    ew, eh, env = sc.env.map
    env = np.asarray(env, np.float32).reshape(eh, ew, 3)
    # A new map with a different brightness distribution: a hot band.
    new = env.copy()
    new[eh // 4: eh // 4 + 6] *= 40.0
    w, h, spp, depth = 160, 96, 4, 4
    d = DiffRenderer(ctx, sc, w, h, depth)
    stale_before = d.forward(spp, 0, record=False)
    d.set_env_texels(new, rebuild_sampler=False)
    stale = d.forward(spp, 0, record=False)
    d.set_env_texels(new, rebuild_sampler=True, on_device=on_device)
    rebuilt = d.forward(spp, 0, record=False)
    fresh_scene = replace(sc, env=replace(sc.env, map=(ew, eh, new.reshape(-1))))
    f = DiffRenderer(ctx, fresh_scene, w, h, depth)
    fresh = f.forward(spp, 0, record=False)
    assert np.array_equal(rebuilt, fresh)
    assert not np.array_equal(stale, fresh)  # the old sampler draws other directions
    assert not np.array_equal(stale_before, fresh)
    d.close()
    f.close()


def implied_scaled(table):
    """n x the probability with which the alias table selects each cell:
    its own share plus the shares of the cells that alias it."""
    t = np.asarray(table, np.float64).reshape(-1, 4)
    prob, alias = t[:, 0], t[:, 1].astype(np.int64)
    out = prob.copy()
    np.add.at(out, alias, 1.0 - prob)
    return out


@pytest.mark.parametrize("kind", KINDS)
def test_parallel_device_sampler_has_the_host_distribution(ctx, kind):
    """lucent_env_sampler_gpu_parallel: the same grid, pmf fields and
    distribution as the host build (every cell selected with probability
    pmf up to float32 rounding), with alias pairs of its own."""
    rgb, w, h, max_w = make_map(kind)
    print(f"parallel env sampler {kind}: {w}x{h}, max width {max_w}", flush=True)
    ref, rgw, rgh, rtotal = host_sampler(rgb, w, h, max_w)
    table, gw, gh, total, ms = ctx.env_sampler_gpu(rgba_tensor(rgb, w, h), max_w, parallel=True)
    got = table.cpu().numpy().reshape(-1, 4)
    ref = ref.reshape(-1, 4)
    n = gw * gh
    assert (gw, gh) == (rgw, rgh)
    assert total == pytest.approx(rtotal, rel=1e-12)
    # pmf * n of the cell and of its alias (float32): as the host's, to the
    # last bit of the total.
    assert np.allclose(got[:, 2], ref[:, 2], rtol=2e-7, atol=0)
    alias = got[:, 1].astype(np.int64)
    assert np.array_equal(got[:, 1], alias.astype(np.float32)) and alias.min() >= 0 and alias.max() < n
    assert np.array_equal(got[:, 3], got[alias, 2])
    assert got[:, 0].min() >= 0 and got[:, 0].max() <= 1
    # The distribution: both tables select cell i with probability
    # pmf_i up to the float32 rounding of their stored probabilities.
    target = ref[:, 2].astype(np.float64)
    for t in (ref, got):
        inbound = np.bincount(t[:, 1].astype(np.int64), minlength=n).astype(np.float64)
        err = np.abs(implied_scaled(t) - target)
        assert np.all(err <= 2e-7 * (1.0 + target) * (1.0 + inbound)), float(err.max())
    print(f"  grid {gw}x{gh}, parallel device build {ms:.3f} ms", flush=True)


def test_parallel_device_sampler_fails_like_the_host_and_leaves_the_table(ctx):
    import torch
    rgb, w, h, _ = make_map("odd")
    bad = rgb.copy()
    bad[(3 * 50 + 7) * 3 + 1] = np.inf
    with pytest.raises(_native.LucentError) as host_err:
        host_sampler(bad, w, h, 1024)
    t = rgba_tensor(bad, w, h)
    table = torch.full((w * h, 4), 7.0, dtype=torch.float32, device="cuda:0")
    gw, gh, total, ms = C.c_uint32(), C.c_uint32(), C.c_double(), C.c_float()
    with pytest.raises(_native.LucentError) as dev_err:
        _native.check(_native.lib().lucent_env_sampler_gpu_parallel(
            ctx.handle, C.c_void_p(t.data_ptr()), w, h, 1024, C.c_void_p(table.data_ptr()), C.byref(gw), C.byref(gh),
            C.byref(total), C.byref(ms), None))
    assert str(dev_err.value) == str(host_err.value)
    assert bool((table == 7.0).all())


def test_scene_with_parallel_rebuilt_sampler_renders_the_same_image_in_expectation(ctx):
    """A sampler rebuilt by the parallel device build picks other cells for
    the same random numbers, so its images differ from a fresh scene's by
    noise only: the same mean brightness within its standard error, and a
    pixel difference no larger than between two seeds of the fresh scene."""
    sc = workloads.spheres_scene()
    ew, eh, env = sc.env.map
    env = np.asarray(env, np.float32).reshape(eh, ew, 3)
    new = env.copy()
    new[eh // 4: eh // 4 + 6] *= 40.0
    w, h, spp, depth = 96, 64, 64, 4
    d = DiffRenderer(ctx, sc, w, h, depth)
    d.set_env_texels(new, rebuild_sampler=True, on_device=2)
    par = d.forward(spp, 0, record=False)
    fresh_scene = replace(sc, env=replace(sc.env, map=(ew, eh, new.reshape(-1))))
    f = DiffRenderer(ctx, fresh_scene, w, h, depth)
    fresh0 = f.forward(spp, 0, record=False)
    fresh1 = f.forward(spp, 1000, record=False)
    assert not np.array_equal(par, fresh0)
    noise = float(np.sqrt(np.mean((fresh1 - fresh0) ** 2)))
    diff = float(np.sqrt(np.mean((par - fresh0) ** 2)))
    assert diff <= 1.25 * noise, (diff, noise)
    # Mean brightness per pixel: the standard error of the image mean from
    # the seed-to-seed spread of independent pixels (an upper bound for the
    # partly correlated difference at equal seeds).
    lum = [img.mean(axis=-1) for img in (par, fresh0, fresh1)]
    se = float(np.std(lum[2] - lum[1]) / np.sqrt(lum[0].size))
    assert abs(float(lum[0].mean() - lum[1].mean())) <= 5 * se, (float(lum[0].mean()), float(lum[1].mean()), se)
    d.close()
    f.close()
