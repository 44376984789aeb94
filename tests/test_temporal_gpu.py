"""GPU parity of the temporal reprojection sample (trace.frag:274-355 with
uReproject = 1) and of the presentation pass (display.frag) against the
oracle.

Reprojection: the GPU film traced at one camera is read back and handed to
the oracle as the previous view; after a camera move the GPU's reprojecting
sample and the oracle's restatement of it (FP64, same PCG stream) are
compared per pixel.  Paths that float32 and float64 resolve differently (a
Russian-roulette draw, a grazing hit) give a different new sample, so the
checks are fractions: colour within 1e-3 relative on >= 93% of pixels,
history within 1e-4 on >= 98%, FP16 guides within half precision on >= 97%.
Display: bytes within 1 everywhere and equal on >= 99% of pixels.
"""
import ctypes as C
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from lucent_amd import _native
from lucent_amd import scene as S
from lucent_amd.renderer import Film, Renderer, RenderSettings, camera_struct

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(Path(__file__).parent))
from test_pathtrace_gpu import feature_scene  # noqa: E402

W, H = 96, 64


def orbit(cam, degrees, dy=0.0):
    """The camera rotated about the vertical axis through its target."""
    px, py, pz = np.subtract(cam.position, cam.target)
    a = math.radians(degrees)
    x, z = px * math.cos(a) + pz * math.sin(a), -px * math.sin(a) + pz * math.cos(a)
    pos = (cam.target[0] + x, cam.target[1] + py + dy, cam.target[2] + z)
    return S.Camera(position=pos, target=cam.target, fov=cam.fov, aperture=cam.aperture,
                    focus_distance=cam.focus_distance)


def all_xy(w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)


def compare_reprojected(gpu, ref, ok, label):
    n = ref.shape[0]
    col = gpu["color"].reshape(n, 4).astype(np.float64)
    alb = gpu["albedo"].reshape(n, 4).astype(np.float64)
    nrm = gpu["normal"].reshape(n, 4).astype(np.float64)
    hist = gpu["history"].reshape(n).astype(np.float64)
    c_ok = np.all(np.abs(col - ref[:, 0:4]) <= 1e-3 * np.abs(ref[:, 0:4]) + 1e-4, axis=1)
    h_ok = np.abs(hist - ref[:, 12]) <= 1e-4 * np.abs(ref[:, 12]) + 1e-4
    a_ok = np.all(np.abs(alb - ref[:, 4:8]) <= 2e-3 * np.abs(ref[:, 4:8]) + 2e-3, axis=1)
    n_ok = np.all(np.abs(nrm[:, :3] - ref[:, 8:11]) <= 2e-3, axis=1) & \
        (np.abs(nrm[:, 3] - ref[:, 11]) <= 2e-3 * np.abs(ref[:, 11]) + 1e-3)
    msg = (f"{label}: reprojected {ok.mean():.3f}; colour {c_ok.mean():.4f}, history {h_ok.mean():.4f}, "
           f"albedo {a_ok.mean():.4f}, normal {n_ok.mean():.4f} of pixels agree")
    print(msg)
    assert 0.3 < ok.mean() < 1.0, msg  # some history survives, some pixels are disoccluded
    assert c_ok.mean() >= 0.93 and h_ok.mean() >= 0.98 and a_ok.mean() >= 0.97 and n_ok.mean() >= 0.97, msg


def test_reprojected_sample_matches_oracle(ctx):
    sc = feature_scene()
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=8, firefly_clamp=0.0))
    r.load_scene(sc)
    r.film.set_seed(11)
    r.reset()
    r.trace_samples(8)
    prev = r.film.read()
    c0, c1 = sc.camera, orbit(sc.camera, 4.0, 0.15)
    r.upload_camera(c1)
    r.trace_samples(1)
    gpu = r.film.read()
    info = r.film.info()
    assert info["samples"] == 9 and info["seed"] == 11 + 9
    o = O.PtOracle(sc)
    ref, ok = o.trace_reprojected(o.frame(c1, W, H, 8, 0.0), prev, o.frame(c0, W, H, 8, 0.0), 8.0, 11 + 8,
                                  all_xy(W, H))
    compare_reprojected(gpu, ref, ok, "orbit 4 deg")
    r.close()


def test_reprojection_from_a_preview_film_of_another_size(ctx):
    """The preview case: the history comes from a half-resolution film."""
    sc = feature_scene()
    r = Renderer(ctx, W // 2, H // 2, RenderSettings(max_bounces=8, firefly_clamp=0.0))
    r.load_scene(sc)
    r.film.set_seed(3)
    r.reset()
    r.trace_samples(12)
    prev = r.film.read()
    c0, c1 = sc.camera, orbit(sc.camera, -3.0)
    dst = Film(ctx, W, H)
    dst.set_seed(40)
    st = RenderSettings(max_bounces=8, firefly_clamp=0.0).struct()
    _native.check(_native.lib().lucent_trace_reprojected(ctx.handle, r.device_scene.handle, dst.handle,
                                                         C.byref(camera_struct(c1)), r.film.handle,
                                                         C.byref(camera_struct(c0)), 8.0, C.byref(st), None))
    gpu = dst.read()
    assert dst.info()["samples"] == 9 and dst.info()["seed"] == 41
    o = O.PtOracle(sc)
    ref, ok = o.trace_reprojected(o.frame(c1, W, H, 8, 0.0), prev, o.frame(c0, W // 2, H // 2, 8, 0.0), 8.0, 40,
                                  all_xy(W, H))
    compare_reprojected(gpu, ref, ok, "preview 48x32 -> 96x64")
    dst.destroy()
    r.close()


def test_renderer_camera_move_sequence_is_the_api_sequence(ctx):
    """Renderer.upload_camera + trace_samples(3) = one lucent_trace_reprojected
    and lucent_trace_samples(2) on another film, bit for bit; without
    temporal reuse the move resets the film."""
    sc = feature_scene()
    s = RenderSettings(max_bounces=8, firefly_clamp=0.0)
    r = Renderer(ctx, W, H, s)
    r.load_scene(sc)
    r.film.set_seed(5)
    r.reset()
    r.trace_samples(4)
    c0, c1 = sc.camera, orbit(sc.camera, 2.0)
    r.upload_camera(c1)
    r.trace_samples(3)
    a = r.film.read()
    assert r.frame == 7
    b = Film(ctx, W, H)
    b.set_seed(5)
    b.reset()
    lib = _native.lib()
    st = s.struct()
    _native.check(lib.lucent_trace_samples(ctx.handle, r.device_scene.handle, b.handle, C.byref(camera_struct(c0)),
                                           C.byref(st), 4, None))
    _native.check(lib.lucent_trace_reprojected(ctx.handle, r.device_scene.handle, b.handle,
                                               C.byref(camera_struct(c1)), b.handle, C.byref(camera_struct(c0)),
                                               8.0, C.byref(st), None))
    _native.check(lib.lucent_trace_samples(ctx.handle, r.device_scene.handle, b.handle, C.byref(camera_struct(c1)),
                                           C.byref(st), 2, None))
    bb = b.read()
    for k in a:
        assert np.array_equal(a[k], bb[k]), k
    r.settings.temporal = False
    r.upload_camera(c0)
    assert r.frame == 0
    b.destroy()
    r.close()


DISPLAY_CASES = [(0.0, 1.0, 0), (0.5, 1.3, 0), (1.0, 0.7, 1), (0.0, 2.0, 2), (0.8, 1.0, 0)]


@pytest.mark.parametrize("view", [(W, H), (150, 100)])
def test_display_matches_oracle(ctx, view):
    sc = feature_scene()
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=8, firefly_clamp=0.0))
    r.load_scene(sc)
    r.trace_samples(8)
    r.run_denoiser()
    film = r.film.read()
    den = r.read_denoised()["illum_var"]
    vw, vh = view
    lib = _native.lib()
    for weight, exposure, tonemap in DISPLAY_CASES:
        p = _native.DisplayParamsStruct(weight, exposure, tonemap)
        got = np.empty((vh, vw, 4), np.uint8)
        _native.check(lib.lucent_screenshot(ctx.handle, r.film.handle, C.byref(p), vw, vh, got.ctypes.data))
        want, _ = O.display(film["color"], film["albedo"], den if weight > 0 else None, vw, vh, weight, exposure,
                            tonemap)
        d = np.abs(got.astype(np.int32) - want.astype(np.int32))
        eq = float((d == 0).mean())
        print(f"view {vw}x{vh} weight {weight} exposure {exposure} tonemap {tonemap}: max |diff| {d.max()}, "
              f"equal {eq:.5f}")
        assert d.max() <= 1 and eq >= 0.99
    r.close()


def test_display_into_device_memory_equals_screenshot(ctx):
    import torch

    sc = feature_scene()
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=4, firefly_clamp=0.0))
    r.load_scene(sc)
    r.render(4)
    shot = r.screenshot(120, 80)
    buf = torch.empty((80, 120, 4), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    r.display(buf.data_ptr(), 120, 80)
    ctx.synchronize()
    assert np.array_equal(buf.cpu().numpy(), shot)
    p = _native.DisplayParamsStruct(0.5, 1.0, 0)
    r2 = Renderer(ctx, 16, 16, RenderSettings(max_bounces=2, firefly_clamp=0.0))
    r2.load_scene(sc)
    r2.trace_samples(1)
    out = np.empty((16, 16, 4), np.uint8)
    status = _native.lib().lucent_screenshot(ctx.handle, r2.film.handle, C.byref(p), 16, 16, out.ctypes.data)
    assert status != 0 and b"denoise" in _native.lib().lucent_last_error(ctx.handle)
    r2.close()
    r.close()
