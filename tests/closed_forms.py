"""Scenes whose converged radiance the mathematics fixes, and those values.

Shared by the oracle's own checks (test_oracle_integrator.py, CPU) and the
GPU checks (test_closed_form_gpu.py).  Each expected value comes from an
independent computation — a closed form, or a numpy quadrature of the BSDF
the oracle evaluates — never from the integrator under test:

  * empty scene under a constant environment: every camera ray escapes at
    bounce 0, unweighted (trace.frag:146-155), so every pixel is L exactly
    (L a power of two, so bilinear filtering of the constant map is exact);
  * a directly viewed emissive quad: bounce-0 emission carries no MIS
    weight and ends the path (trace.frag:170-184): exactly its emission;
  * a lone convex sphere under a constant environment at maxBounces 1:
    environment NEE plus the BSDF-sampled escape, MIS-combined
    (trace.frag:216-261, lights.glsl:81-91), estimate L * rho(wo) with
    rho(wo) = int f(wo, wi) cos(theta_i) dw_i, the directional albedo;
  * a plane under one quad light, black environment, maxBounces 1: light
    NEE plus the BSDF-sampled hit of the light, MIS-combined, estimate
    int_A f(wo, wi) Le cos(theta_i) cos(theta_l) / d^2 dA; the area
    quadrature is itself pinned by the rectangle form factor.
"""
from __future__ import annotations

import math

import numpy as np

import oracle as O
from oracle import packing as OP
from lucent_amd import scene as S

# Powers of two: bilinear filtering and running means of them are exact.
ENV_RGB = (0.5, 0.25, 1.0)
ENV_INTENSITY = 2.0
EMISSION = (4.0, 2.0, 1.0)


def const_env(rgb=ENV_RGB, w=16, h=8, intensity=ENV_INTENSITY, background=True):
    img = np.empty((h, w, 3), np.float32)
    img[...] = rgb
    return S.Environment(map=(w, h, img.reshape(-1)), rotation=0.0, intensity=intensity, background=background)


def black_env():
    return S.Environment(map=(16, 8, np.zeros(16 * 8 * 3, np.float32)), intensity=1.0, background=True)


def empty_scene():
    return S.Scene(camera=S.Camera((0.0, 0.5, 3.0), (0.0, 0.0, 0.0), 50.0, 0.0, 3.0), meshes=[], env=const_env())


def emitter_scene():
    lamp = S.quad((-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (1.0, 1.0, 0.0), (-1.0, 1.0, 0.0))
    return S.Scene(camera=S.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 40.0, 0.0, 3.0),
                   meshes=[S.MeshInstance(lamp, S.Material(albedo=(0.0, 0.0, 0.0), emission=EMISSION))],
                   env=const_env())


SPHERE_MATERIALS = {
    "matte": S.Material(albedo=(0.5, 0.25, 0.75), roughness=1.0),
    "rough_metal": S.Material(albedo=(0.9, 0.6, 0.3), metallic=1.0, roughness=0.5),
    "coat_sheen": S.Material(albedo=(0.6, 0.2, 0.1), roughness=0.7, clearcoat=1.0, clearcoat_roughness=0.4,
                             sheen_color=(0.5, 0.5, 0.5), sheen_roughness=0.6),
}
SPHERE_CAMERA = S.Camera((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 40.0, 0.0, 4.0)


def sphere_scene(material):
    return S.Scene(camera=SPHERE_CAMERA, meshes=[], spheres=[S.Sphere((0.0, 0.0, 0.0), 1.0, material)],
                   env=const_env())


LIGHT_Y = 1.5
LIGHT_HALF = 0.5
PLANE_MATERIAL = S.Material(albedo=(0.5, 0.75, 0.25), roughness=0.8)
PLANE_CAMERA = S.Camera((0.0, 2.0, 3.0), (0.0, 0.0, 0.0), 45.0, 0.0, 3.0)


def plane_light_scene():
    h, a = LIGHT_Y, LIGHT_HALF
    # Corners ordered so the quad faces down (-y), towards the plane.
    lamp = S.quad((-a, h, -a), (a, h, -a), (a, h, a), (-a, h, a))
    return S.Scene(camera=PLANE_CAMERA, meshes=[S.MeshInstance(lamp, S.light((1.0, 1.0, 1.0), 8.0))],
                   env=black_env(), ground=S.Ground(True, 0.0, PLANE_MATERIAL))


# ---------------------------------------------------------------------------
# Camera rays (trace.frag:53-62 with a pinhole) at sub-pixel positions.
# ---------------------------------------------------------------------------
def camera_dirs(camera, width, height, xy, sub=4):
    """Directions of sub x sub stratified rays over each listed pixel's
    footprint [x + 0.5, x + 1.5) x [y + 0.5, y + 1.5): (n, sub*sub, 3)."""
    fw, r, u, t = O.camera_basis(camera.position, camera.target, camera.fov)
    fw, r, u = np.float32(fw).astype(np.float64), np.float32(r).astype(np.float64), np.float32(u).astype(np.float64)
    xy = np.asarray(xy, np.float64).reshape(-1, 2)
    o = (np.arange(sub) + 0.5) / sub
    jx, jy = np.meshgrid(o, o, indexing="xy")
    px = xy[:, 0:1] + 0.5 + jx.reshape(1, -1)
    py = xy[:, 1:2] + 0.5 + jy.reshape(1, -1)
    ux = px / width * 2 - 1
    uy = py / height * 2 - 1
    aspect = width / height
    d = fw[None, None, :] + r[None, None, :] * (ux * aspect * t)[..., None] + u[None, None, :] * (uy * t)[..., None]
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


# ---------------------------------------------------------------------------
# Directional albedo by quadrature of the oracle's BSDF evaluation.
# ---------------------------------------------------------------------------
def hemisphere_grid(n_theta=256, n_phi=128):
    th = (np.arange(n_theta) + 0.5) * (math.pi / 2 / n_theta)
    ph = (np.arange(n_phi) + 0.5) * (2 * math.pi / n_phi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    wi = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (math.pi / 2 / n_theta) * (2 * math.pi / n_phi) * np.sin(T).reshape(-1)
    return wi, dw


def albedo_table(material, cos_values, n_theta=256, n_phi=128):
    """rho(cos theta_o) per channel for each value (n, 3): midpoint rule in
    (theta, phi) over the upper hemisphere of f(wo, wi) cos(theta_i)."""
    mat = OP.pack_materials([material])
    wi, dw = hemisphere_grid(n_theta, n_phi)
    out = []
    for c in cos_values:
        s = math.sqrt(max(0.0, 1 - c * c))
        wo = np.array([s, 0.0, c])
        f, _ = O.eval_opaque_batch(mat, wo, wi)
        out.append((f * (wi[:, 2] * dw)[:, None]).sum(0))
    return np.asarray(out)


def sphere_expected(material, width, height, xy, sub=4):
    """Expected pixel radiance of sphere_scene(material) at maxBounces 1: the
    footprint average of L * rho(cos theta_o); NaN where a sub-ray misses the
    sphere or grazes it (cos theta_o < 0.15)."""
    d = camera_dirs(SPHERE_CAMERA, width, height, xy, sub)
    o = np.asarray(SPHERE_CAMERA.position, np.float64)
    b = (d * o).sum(-1)
    c = float(o @ o) - 1.0
    disc = b * b - c
    hit = disc > 0
    t = -b - np.sqrt(np.maximum(disc, 0))
    p = o + d * t[..., None]
    cos_o = np.where(hit, -(p * d).sum(-1), -1.0)
    grid = np.linspace(0.1, 1.0, 46)
    table = albedo_table(material, grid)
    rho = np.stack([np.interp(cos_o, grid, table[:, k]) for k in range(3)], -1)
    L = np.array(ENV_RGB) * ENV_INTENSITY
    ok = np.all(cos_o >= 0.15, axis=1)
    e = (rho * L).mean(1)
    e[~ok] = np.nan
    return e


def rectangle_form_factor(a, b, h):
    """Differential area to a parallel a x b rectangle at height h with one
    corner straight above (the textbook closed form)."""
    A, B = a / h, b / h
    return (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A))
            + B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)


def light_quadrature(points, f_of, n=64):
    """sum over an n x n midpoint grid of the light quad of
    f_of(point, wi) * cos(theta_i) * cos(theta_l) / d^2 * dA for plane points
    (normal +y); f_of returns (m, 3)."""
    h, a = LIGHT_Y, LIGHT_HALF
    g = (np.arange(n) + 0.5) / n * 2 * a - a
    X, Z = np.meshgrid(g, g, indexing="ij")
    q = np.stack([X.reshape(-1), np.full(n * n, h), Z.reshape(-1)], -1)
    dA = (2 * a / n) ** 2
    out = []
    for p in points:
        v = q - p
        d2 = (v * v).sum(-1)
        wi = v / np.sqrt(d2)[:, None]
        cos_i = wi[:, 1]
        cos_l = wi[:, 1]  # the light faces -y
        w = cos_i * cos_l / d2 * dA
        out.append((f_of(p, wi) * w[:, None]).sum(0))
    return np.asarray(out)


def plane_expected(width, height, xy, sub=2, n=48):
    """Expected pixel radiance of plane_light_scene() at maxBounces 1 (NaN
    where a sub-ray misses the plane)."""
    d = camera_dirs(PLANE_CAMERA, width, height, xy, sub)
    o = np.asarray(PLANE_CAMERA.position, np.float64)
    t = -o[1] / d[..., 1]
    ok = np.all((d[..., 1] < -1e-3) & (t > 0), axis=1)
    mat = OP.pack_materials([PLANE_MATERIAL])
    Le = 8.0
    res = np.full((d.shape[0], 3), np.nan)
    for i in np.flatnonzero(ok):
        pts = o + d[i] * t[i][:, None]
        acc = np.zeros(3)
        for k, p in enumerate(pts):
            wo_w = -d[i, k]
            # Local frame: the plane normal +y becomes +z.
            wo = np.array([wo_w[0], wo_w[2], wo_w[1]])

            def f_of(_, wi_w, wo=wo):
                wi = np.stack([wi_w[:, 0], wi_w[:, 2], wi_w[:, 1]], -1)
                f, _ = O.eval_opaque_batch(mat, wo, wi)
                return f

            acc += light_quadrature([p], f_of, n)[0] * Le
        res[i] = acc / len(pts)
    return res


# ---------------------------------------------------------------------------
# Alpha cut-out: brute force over a 4x4 0/1 mask (scene.glsl:176-189).
# ---------------------------------------------------------------------------
MASK_CAMERA = S.Camera((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 40.0, 0.0, 3.0)


def mask_texture(seed=4, n=4):
    rng = np.random.default_rng(seed)
    img = np.full((n, n, 4), 255, np.uint8)
    img[..., 3] = np.where(rng.random((n, n)) < 0.5, 0, 255)
    return img


def mask_scene():
    """An emissive unit-uv quad over [-1, 1]^2 at z = 0, cut out by a 4x4
    mask (mask mode, cutoff 0.5), in front of a constant environment: a
    camera ray that the alpha test lets through sees L, one it stops sees
    the emission."""
    q = S.quad((-1.0, -1.0, 0.0), (1.0, -1.0, 0.0), (1.0, 1.0, 0.0), (-1.0, 1.0, 0.0))
    uvs = np.array([0, 0, 1, 0, 1, 1, 0, 0, 1, 1, 0, 1], np.float32)
    m = S.Material(albedo=(0.0, 0.0, 0.0), emission=EMISSION, alpha_mode="mask", alpha_cutoff=0.5, base_color_map=0)
    return S.Scene(camera=MASK_CAMERA, meshes=[S.MeshInstance(S.Mesh(q.positions, q.normals, uvs), m)],
                   env=const_env(), textures=[mask_texture()])


def sample_rays(camera, width, height, xy, seed0, spp):
    """The jittered camera-ray directions of each listed pixel's samples
    (seeds seed0 .. seed0 + spp - 1), from the PCG stream (trace.frag:53-62,
    324-330): (n, spp, 3)."""
    fw, r, u, t = O.camera_basis(camera.position, camera.target, camera.fov)
    fw, r, u = np.float32(fw).astype(np.float64), np.float32(r).astype(np.float64), np.float32(u).astype(np.float64)
    xy = np.asarray(xy, np.int64).reshape(-1, 2)
    out = np.zeros((xy.shape[0], spp, 3))
    aspect = width / height
    for i, (x, y) in enumerate(xy):
        for k in range(spp):
            s = O.pcg_stream(int(x + y * width), seed0 + k, 2)
            r0, r1 = (float(np.float32(v)) * 2.0 ** -32 for v in s)
            ux = (x + 0.5 + r0) / width * 2 - 1
            uy = (y + 0.5 + r1) / height * 2 - 1
            d = fw + r * (ux * aspect * t) + u * (uy * t)
            out[i, k] = d / np.linalg.norm(d)
    return out


def mask_alpha(u, v, img):
    """Bilinear REPEAT sample of the mask's alpha at (u, v), texel centres at
    (i + 0.5) / n, rows in v order."""
    a = img[..., 3].astype(np.float64) / 255.0
    n = a.shape[0]
    x, y = u * n - 0.5, v * n - 0.5
    x0, y0 = np.floor(x), np.floor(y)
    fx, fy = x - x0, y - y0
    x0, y0 = x0.astype(np.int64), y0.astype(np.int64)

    def at(yy, xx):
        return a[np.mod(yy, n), np.mod(xx, n)]

    top = at(y0, x0) * (1 - fx) + at(y0, x0 + 1) * fx
    bot = at(y0 + 1, x0) * (1 - fx) + at(y0 + 1, x0 + 1) * fx
    return top * (1 - fy) + bot * fy


def mask_expected(width, height, xy, seed0, spp, margin=1e-4):
    """Per sample: 1 where the quad stops the ray, 0 where it passes, -1
    where the ray's alpha is within `margin` of the cutoff or it grazes the
    quad's edge (left undecided)."""
    d = sample_rays(MASK_CAMERA, width, height, xy, seed0, spp)
    o = np.asarray(MASK_CAMERA.position, np.float64)
    t = -o[2] / d[..., 2]
    p = o + d * t[..., None]
    inside = (np.abs(p[..., 0]) < 1) & (np.abs(p[..., 1]) < 1)
    edge = np.abs(np.abs(p[..., :2]) - 1).min(-1) < 1e-6
    alpha = mask_alpha((p[..., 0] + 1) / 2, (p[..., 1] + 1) / 2, mask_texture())
    res = np.where(inside & (alpha >= 0.5), 1, 0)
    res[edge | (inside & (np.abs(alpha - 0.5) < margin))] = -1
    return res
