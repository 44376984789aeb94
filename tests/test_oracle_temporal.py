"""The oracle's temporal reprojection (trace.frag:274-355) and presentation
(display.frag) restatements, pinned by closed forms:

  * with an unchanged camera a sky sample projects to its own jittered
    position (x + 0.5 + r0, y + 0.5 + r1, from the pixel's PCG stream), so a
    film that is linear in x and y reprojects to exactly that linear
    function; the history becomes min(n, cap) and the sample is mixed in with
    weight 1/(n+1);
  * the history cap follows the reuse factor: 1 for rough dielectrics,
    0 for smooth metal (trace.frag:187-193);
  * disoccluded pixels (no tap saw the same surface) and views turned away
    start over: the output is the plain sample;
  * display: ACES(1) = 2.54/3.16, Reinhard(1) = 1/2, Catmull-Rom reproduces
    constants and linear ramps, saturated white stays 255 under dithering.
"""
import math

import numpy as np

import oracle as O
from lucent_amd import scene as S
from lucent_amd import workloads

W, H = 24, 16


def sky_scene(camera):
    """A tiny triangle far behind the camera: every camera ray misses."""
    tri = S.from_triangles([((0.0, -50.0, 40.0), (0.1, -50.0, 40.0), (0.0, -49.9, 40.0))])
    w, h = 64, 32
    env = S.Environment(map=(w, h, workloads.synthetic_hdri(w, h, sigma_deg=6.0, peak=20.0)), background=True)
    return S.Scene(camera=camera, meshes=[S.MeshInstance(tri, S.matte((0.5, 0.5, 0.5)))], env=env,
                   ground=S.Ground(False, 0.0, S.matte((0.5, 0.5, 0.5))))


def plane_scene(material):
    """A large quad filling the view (every primary ray hits it)."""
    q = S.quad((-20.0, -20.0, 0.0), (20.0, -20.0, 0.0), (20.0, 20.0, 0.0), (-20.0, 20.0, 0.0))
    w, h = 64, 32
    env = S.Environment(map=(w, h, workloads.synthetic_hdri(w, h, sigma_deg=6.0, peak=20.0)), background=True)
    cam = S.Camera(position=(0.0, 0.0, 5.0), target=(0.0, 0.0, 0.0), fov=40.0, aperture=0.0, focus_distance=5.0)
    return S.Scene(camera=cam, meshes=[S.MeshInstance(q, material)], env=env,
                   ground=S.Ground(False, 0.0, S.matte((0.5, 0.5, 0.5))))


def cam(pos, target=(0.0, 0.0, 0.0)):
    return S.Camera(position=pos, target=target, fov=50.0, aperture=0.0, focus_distance=5.0)


def all_xy(w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.stack([xx.reshape(-1), yy.reshape(-1)], -1).astype(np.uint32)


def uniform_film(w, h, color, coverage, normal_depth, history):
    return {"color": np.broadcast_to(np.float32(color), (h, w, 4)).copy(),
            "albedo": np.broadcast_to(np.float32([0.5, 0.5, 0.5, coverage]), (h, w, 4)).copy(),
            "normal": np.broadcast_to(np.float32(normal_depth), (h, w, 4)).copy(),
            "history": np.full((h, w), history, np.float32)}


def fresh_samples(o, f, seed, xy):
    ref, per = o.trace_pixels(f, seed, 1, xy, per_sample=True)
    return ref


def test_sky_reprojection_lands_on_the_jittered_sample_position():
    c = cam((0.0, 0.0, 5.0), (0.0, 0.0, -10.0))
    sc = sky_scene(c)
    o = O.PtOracle(sc)
    f = o.frame(c, W, H, 3, 0.0)
    film = uniform_film(W, H, (0, 0, 0, 0), 0.0, (0, 0, 1, 1e4), 5.0)
    yy, xx = np.mgrid[0:H, 0:W]
    # colour r = x + 0.5 at pixel centres, g = y + 0.5: linear in the pixel coordinates
    film["color"][..., 0] = xx + 0.5
    film["color"][..., 1] = yy + 0.5
    seed = 77
    xy = all_xy(W, H)
    out, ok = o.trace_reprojected(f, film, f, 8.0, seed, xy)
    new = fresh_samples(o, f, seed, xy)
    interior = (xy[:, 0] < W - 1) & (xy[:, 1] < H - 1)
    assert ok[interior].all()
    for p in np.nonzero(interior)[0]:
        x, y = int(xy[p, 0]), int(xy[p, 1])
        u = O.pcg_stream(x + y * W, seed, 2)
        r0, r1 = [float(np.float32(v)) / 2 ** 32 for v in u]
        k = 1.0 / 6.0  # history 5 -> 1/(5+1)
        want_r = (x + 0.5 + r0) * (1 - k) + new[p, 0] * k
        want_g = (y + 0.5 + r1) * (1 - k) + new[p, 1] * k
        assert abs(out[p, 0] - want_r) < 1e-5 * (x + 1), (x, y, out[p, 0], want_r)
        assert abs(out[p, 1] - want_g) < 1e-5 * (y + 1), (x, y, out[p, 1], want_g)
        assert abs(out[p, 2] - new[p, 2] * k) < 1e-12
        assert abs(out[p, 12] - 6.0) < 1e-12


def test_history_is_capped():
    c = cam((0.0, 0.0, 5.0), (0.0, 0.0, -10.0))
    sc = sky_scene(c)
    o = O.PtOracle(sc)
    f = o.frame(c, W, H, 3, 0.0)
    film = uniform_film(W, H, (0.25, 0.5, 0.75, 0.1), 0.0, (0, 0, 1, 1e4), 100.0)
    xy = all_xy(W, H)[:40]
    out, ok = o.trace_reprojected(f, film, f, 8.0, 3, xy)
    new = fresh_samples(o, f, 3, xy)
    assert ok.all()
    np.testing.assert_allclose(out[:, 0], 0.25 * 8 / 9 + new[:, 0] / 9, rtol=1e-12)
    assert np.all(out[:, 12] == 9.0)


def test_turned_away_view_starts_over():
    c = cam((0.0, 0.0, 5.0), (0.0, 0.0, -10.0))
    back = cam((0.0, 0.0, 5.0), (0.0, 0.0, 20.0))
    sc = sky_scene(c)
    o = O.PtOracle(sc)
    f = o.frame(c, W, H, 3, 0.0)
    fb = o.frame(back, W, H, 3, 0.0)
    film = uniform_film(W, H, (9, 9, 9, 81), 0.0, (0, 0, 1, 1e4), 4.0)
    xy = all_xy(W, H)
    out, ok = o.trace_reprojected(f, film, fb, 8.0, 11, xy)
    new = fresh_samples(o, f, 11, xy)
    assert not ok.any()
    np.testing.assert_array_equal(out[:, :12], new)
    assert np.all(out[:, 12] == 1.0)


def test_surface_over_a_sky_history_is_disoccluded():
    sc = plane_scene(S.matte((0.6, 0.5, 0.4), 1.0))
    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, 2, 0.0)
    film = uniform_film(W, H, (9, 9, 9, 81), 0.0, (0, 0, 1, 1e4), 4.0)  # coverage 0 everywhere
    xy = all_xy(W, H)
    out, ok = o.trace_reprojected(f, film, f, 8.0, 5, xy)
    assert not ok.any()
    assert np.all(out[:, 7] == 1.0)  # the new sample hits the plane
    np.testing.assert_array_equal(out[:, :12], fresh_samples(o, f, 5, xy))


def surface_history(o, f, seed, spp):
    ref, _ = o.trace_pixels(f, seed, spp, all_xy(W, H))
    r = ref.reshape(H, W, 12)
    return {"color": r[..., 0:4].astype(np.float32), "albedo": r[..., 4:8].astype(np.float32),
            "normal": r[..., 8:12].astype(np.float32), "history": np.full((H, W), float(spp), np.float32)}


def test_reuse_caps_history_by_material():
    """Rough dielectric: glossy 0, reuse 1, history min(n, 8).  Metal of
    roughness r (not a mirror, so the guide stays on the surface): glossy =
    1 - smoothstep(0.1, 0.7, r), reuse = smoothstep(0.1, 0.7, r), history
    8 * reuse + 1 (trace.frag:187-193, 320)."""
    xy = all_xy(W, H)
    interior = (xy[:, 0] > 0) & (xy[:, 0] < W - 1) & (xy[:, 1] > 0) & (xy[:, 1] < H - 1)
    sc = plane_scene(S.matte((0.6, 0.5, 0.4), 1.0))
    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, 2, 0.0)
    out, ok = o.trace_reprojected(f, surface_history(o, f, 100, 4), f, 8.0, 104, xy)
    assert ok[interior].all()
    np.testing.assert_allclose(out[interior, 12], 5.0, rtol=1e-12)
    r = float(np.float32(0.15))
    t = (r - 0.1) / 0.6
    reuse = t * t * (3 - 2 * t)
    sc = plane_scene(S.metal((0.9, 0.9, 0.9), 0.15))
    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, 2, 0.0)
    out, ok = o.trace_reprojected(f, surface_history(o, f, 100, 4), f, 8.0, 104, xy)
    assert ok[interior].all()
    np.testing.assert_allclose(out[interior, 12], 8.0 * reuse + 1.0, rtol=1e-12)


def test_display_closed_forms():
    h, w = 8, 12
    alb = np.zeros((h, w, 4), np.float32)
    ones = np.ones((h, w, 4), np.float32)
    # ACES(1) = 2.54 / 3.16, Reinhard(1) = 1/2, tone mapping "none" passes 1 through.
    for tm, lin in ((0, 2.54 / 3.16), (1, 0.5), (2, 1.0)):
        _, c = O.display(ones, alb, None, w, h, 0.0, 1.0, tm)
        srgb = 1.055 * lin ** (1 / 2.4) - 0.055
        yy, xx = np.mgrid[0:h, 0:w]
        a = 0.06711056 * (xx + 0.5) + 0.00583715 * (yy + 0.5)
        b = 52.9829189 * (a - np.floor(a))
        dither = ((b - np.floor(b)) - 0.5) / 255
        np.testing.assert_allclose(c[..., 0], srgb + dither, rtol=0, atol=1e-12)
    # Saturated white stays 255 through the dither; black stays 0.
    rgba, _ = O.display(ones * 50.0, alb, None, w, h, 0.0, 1.0, 2)
    assert np.all(rgba[..., :3] == 255) and np.all(rgba[..., 3] == 255)
    rgba, _ = O.display(ones * 0.0, alb, None, w, h, 0.0, 1.0, 0)
    assert np.all(rgba[..., :3] == 0)


def test_catmull_rom_reproduces_constants_and_ramps():
    rh, rw, vh, vw = 10, 16, 20, 32
    yy, xx = np.mgrid[0:rh, 0:rw]
    ramp = np.zeros((rh, rw, 4), np.float32)
    ramp[..., 0] = 0.125 * (xx + 0.5) / 16  # exact in float32
    ramp[..., 1] = 0.3
    alb = np.zeros((rh, rw, 4), np.float32)
    # Undo the output transform: tonemap none, exposure 1, values < 1: linear
    # radiance = srgb^-1(c - dither).
    _, c = O.display(ramp, alb, None, vw, vh, 0.0, 1.0, 2)
    vy, vx = np.mgrid[0:vh, 0:vw]
    a = 0.06711056 * (vx + 0.5) + 0.00583715 * (vy + 0.5)
    b = 52.9829189 * (a - np.floor(a))
    s = c - (((b - np.floor(b)) - 0.5) / 255)[..., None]
    lin = np.where(s >= 0.0031308 * 12.92, ((s + 0.055) / 1.055) ** 2.4, s / 12.92)
    np.testing.assert_allclose(lin[..., 1], float(np.float32(0.3)), rtol=1e-9)
    inner = (vx >= 4) & (vx < vw - 4)
    src = (vx + 0.5) * (rw / vw) - 0.5  # a continuous render-pixel index
    np.testing.assert_allclose(lin[..., 0][inner], (0.125 * (src + 0.5) / 16)[inner], rtol=1e-9)
