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
