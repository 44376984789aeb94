"""Pins the FP64 denoiser oracle with closed forms: without edge stops the
à-trous filter is a normalised B3-spline convolution; a constant image is a
fixed point; the temporal variance branch follows variance.frag:24-27."""
import numpy as np
import pytest

import oracle as O
from lucent_amd.scene import Camera


def _film(w, h, seed=0, hist=8.0):
    rng = np.random.default_rng(seed)
    color = rng.random((h, w, 4)).astype(np.float32) + 0.1
    color[..., 3] = (color[..., :3].mean(-1) ** 2 + 0.05).astype(np.float32)
    albedo = np.ones((h, w, 4), np.float32)
    normal = np.zeros((h, w, 4), np.float32)
    normal[..., 2] = 1
    normal[..., 3] = 5.0
    history = np.full((h, w), hist, np.float32)
    return color, albedo, normal, history


CAM = O.camera13(Camera((0, 0, 5), (0, 0, 0), 45.0))


def b3_convolve(img, step):
    h, w = img.shape[:2]
    k1 = np.array([0.0625, 0.25, 0.375, 0.25, 0.0625])
    out = np.zeros_like(img)
    wsum = np.zeros((h, w, 1))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            k = k1[dx + 2] * k1[dy + 2]
            ys, xs = np.mgrid[0:h, 0:w]
            qy, qx = ys + dy * step, xs + dx * step
            ok = (qy >= 0) & (qy < h) & (qx >= 0) & (qx < w)
            vals = np.zeros_like(img)
            vals[ok] = img[qy[ok], qx[ok]]
            out += vals * k * ok[..., None]
            wsum += (k * ok)[..., None]
    return out / wsum


def test_reduces_to_b3_convolution_without_edge_stops():
    color, albedo, normal, history = _film(40, 30)
    out, _ = O.denoise(color, albedo, normal, history, CAM, sigma_l=1e30, sigma_n=0.0, sigma_p=1e30, iterations=3)
    ref = color[..., :3].astype(np.float64)  # albedo 1: illumination = colour
    for it in range(3):
        ref = b3_convolve(ref, 1 << it)
    assert np.allclose(out[..., :3], ref, rtol=1e-12, atol=1e-12)


def test_constant_image_is_fixed_point():
    color, albedo, normal, history = _film(33, 21)
    color[..., :3] = 0.75
    out, remod = O.denoise(color, albedo * 0.5, normal, history, CAM, iterations=5)
    assert np.allclose(out[..., :3], 0.75 / 0.5, rtol=1e-12)
    assert np.allclose(remod, 0.75, rtol=1e-12)


def test_temporal_variance_branch():
    color, albedo, normal, history = _film(8, 8, hist=16.0)
    out, _ = O.denoise(color, albedo, normal, history, CAM, iterations=0)
    c = color.astype(np.float64)
    l = 0.2126 * c[..., 0] + 0.7152 * c[..., 1] + 0.0722 * c[..., 2]
    var = np.maximum(0, c[..., 3] - l * l) / 1.0 / 16.0
    assert np.allclose(out[..., 3], var, rtol=1e-12)
    assert np.allclose(out[..., :3], c[..., :3], rtol=1e-12)


def test_uncovered_pixels_pass_through():
    color, albedo, normal, history = _film(12, 10)
    albedo[..., 3] = 0.0
    out, _ = O.denoise(color, albedo, normal, history, CAM, iterations=3)
    assert np.array_equal(out[..., :3], color[..., :3].astype(np.float64))
    assert np.all(out[..., 3] == 0)


# ---------------------------------------------------------------------------
# Hand-computed films: each edge stop of atrous.frag:56-80 on its own, with
# the weights worked out for a film of two or three pixels.
# ---------------------------------------------------------------------------
K_CENTRE = 0.375 * 0.375  # h0 * h0
K_SIDE = 0.25 * 0.375     # h1 * h0, a horizontal neighbour at one step


def _row_film(lums, normals, depths, variances, hist=16.0):
    """A 1-row film of grey pixels (albedo 1) with the given luminance, guide
    normal, depth and per-pixel variance of the mean (temporal branch:
    second moment = l^2 + hist * var)."""
    w = len(lums)
    color = np.zeros((1, w, 4), np.float32)
    for i, l in enumerate(lums):
        color[0, i, :3] = l
        color[0, i, 3] = l * l + hist * variances[i]
    albedo = np.ones((1, w, 4), np.float32)
    normal = np.zeros((1, w, 4), np.float32)
    normal[0, :, :3] = normals
    normal[0, :, 3] = depths
    history = np.full((1, w), hist, np.float32)
    return color, albedo, normal, history


def _f32(x):
    return float(np.float32(x))


def test_normal_stop_by_hand():
    """cos = 0.8 between the two normals, sigma_n = 2: w = k * 0.64."""
    color, albedo, normal, history = _row_film([1.0, 0.5], [(0, 0, 1), (0.6, 0, 0.8)], [5.0, 5.0], [0.01, 0.04])
    out, _ = O.denoise(color, albedo, normal, history, CAM, sigma_l=1e30, sigma_n=2.0, sigma_p=1e30, iterations=1)
    v0 = (_f32(color[0, 0, 3]) - 1.0) / 16
    v1 = (_f32(color[0, 1, 3]) - 0.25) / 16
    wn = _f32(0.8) / np.linalg.norm(np.float64([_f32(0.6), 0, _f32(0.8)]))  # dot((0,0,1), n1/|n1|) = 0.8
    w = K_SIDE * wn * wn
    ws = K_CENTRE + w
    assert out[0, 0, 0] == pytest.approx((1.0 * K_CENTRE + 0.5 * w) / ws, rel=1e-12)
    assert out[0, 1, 0] == pytest.approx((0.5 * K_CENTRE + 1.0 * w) / ws, rel=1e-12)
    assert out[0, 0, 3] == pytest.approx((v0 * K_CENTRE ** 2 + v1 * w * w) / ws ** 2, rel=1e-9)


def test_luminance_stop_and_variance_blur_by_hand():
    """sigma_l = 4 on the 3x3-blurred variance: pixel 0's blur is
    (0.25 v0 + 0.125 v1) / 0.375 (only two pixels are inside), and the
    luminance step is chosen so that pixel 0's weight is exactly e^-ln2 = 1/2."""
    v0, v1 = 0.01, 0.04
    var0 = (0.25 * v0 + 0.125 * v1) / 0.375
    var1 = (0.125 * v0 + 0.25 * v1) / 0.375
    denom0 = 4 * np.sqrt(var0) + 1e-6
    l1 = 0.5
    l0 = _f32(l1 + np.log(2) * denom0)
    color, albedo, normal, history = _row_film([l0, l1], [(0, 0, 1), (0, 0, 1)], [5.0, 5.0], [v0, v1])
    # The stored second moments are float32: recompute the variances they imply.
    m = color[0, :, 3].astype(np.float64)
    lv = np.array([l0, l1], np.float64)
    v = (m - lv * lv) / 16
    var0 = (0.25 * v[0] + 0.125 * v[1]) / 0.375
    var1 = (0.125 * v[0] + 0.25 * v[1]) / 0.375
    out, _ = O.denoise(color, albedo, normal, history, CAM, sigma_l=4.0, sigma_n=0.0, sigma_p=1e30, iterations=1)
    w0 = K_SIDE * np.exp(-(l0 - l1) / (4 * np.sqrt(var0) + 1e-6))
    w1 = K_SIDE * np.exp(-(l0 - l1) / (4 * np.sqrt(var1) + 1e-6))
    assert w0 == pytest.approx(K_SIDE / 2, rel=1e-6)
    assert out[0, 0, 0] == pytest.approx((l0 * K_CENTRE + l1 * w0) / (K_CENTRE + w0), rel=1e-12)
    assert out[0, 1, 0] == pytest.approx((l1 * K_CENTRE + l0 * w1) / (K_CENTRE + w1), rel=1e-12)


def test_plane_stop_by_hand():
    """Camera at the origin looking down -z with tan(fov/2) = 0.75 over a
    3x1 film: the primary directions are (-1.5, 0, -1), (0, 0, -1),
    (1.5, 0, -1) normalised; normals +z.  For the middle pixel at depth 2,
    the left neighbour at depth 2 lies 2 - 2/sqrt(3.25) off its plane, the
    right one at depth 2*sqrt(3.25) on it; the footprint is 2 * 0.75 * 2 / 1."""
    cam = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, -1, 0.75], np.float32)
    d2 = _f32(2 * np.sqrt(3.25))
    color, albedo, normal, history = _row_film([0.25, 1.0, 0.5], [(0, 0, 1)] * 3, [2.0, 2.0, d2],
                                               [0.01, 0.01, 0.01])
    out, _ = O.denoise(color, albedo, normal, history, cam, sigma_l=1e30, sigma_n=0.0, sigma_p=1.5, iterations=1)
    reach = 1.5 * (2 * 0.75 * 2.0 / 1) * 1 * 1
    off_left = abs(2.0 - 2.0 / np.sqrt(3.25))
    off_right = abs(2.0 - d2 / np.sqrt(3.25))
    wl = K_SIDE * np.exp(-off_left / (reach + 1e-6))
    wr = K_SIDE * np.exp(-off_right / (reach + 1e-6))
    assert wr == pytest.approx(K_SIDE, rel=1e-6)  # on the plane
    want = (1.0 * K_CENTRE + 0.25 * wl + 0.5 * wr) / (K_CENTRE + wl + wr)
    assert out[0, 1, 0] == pytest.approx(want, rel=1e-9)
