"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

Procedural definitions fixed here (the configurations name them but leave
their functional forms open):

* ``pcg_hash``: the standard PCG hash of a uint32 key (Jarzynski & Olano
  2020: one LCG step, then the RXS-M-XS output function that the reference
  shader's pcg() also uses, common.glsl:13-18); the value-noise lattice hash.
* ``value_noise3``: trilinear interpolation, with smoothstep fade, of lattice
  values in [-1, 1) drawn from ``pcg_hash`` of the lattice coordinates, the
  octave and the seed.
* ``fbm3``: 4 octaves, base frequency 2, lacunarity 2, gain 0.5, normalised
  by the sum of the octave amplitudes so that it lies in [-1, 1].
* ``icosphere_fbm``: an icosahedron subdivided ``subdiv`` times (each edge
  midpoint projected to the unit sphere; 20*4^subdiv triangles), each vertex
  then scaled radially by ``radius * (1 + amplitude * fbm3(v))``.  Vertices
  are shared through the subdivision, so the displaced surface stays closed.
* ``sphere_rays``: origins ``3 * u`` with ``u`` uniform on the unit sphere
  (normalised Gaussian triple), targets uniform in the unit ball
  (radius U^(1/3) along a uniform direction), direction = normalise(target -
  origin) computed in float64, both stored as float32.
"""
from __future__ import annotations

import math
from dataclasses import replace

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def pcg_hash(x: np.ndarray) -> np.ndarray:
    """PCG hash of uint32 keys (state*747796405 + 2891336453, RXS-M-XS output)."""
    s = np.asarray(x, dtype=np.uint64) & _M32
    state = (s * np.uint64(747796405) + np.uint64(2891336453)) & _M32
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & _M32
    return ((word >> np.uint64(22)) ^ word) & _M32


def _lattice(ix, iy, iz, octave: int, seed: int) -> np.ndarray:
    m = np.uint64(0xFFFFFFFF)
    k = pcg_hash((ix.astype(np.int64).astype(np.uint64) & m) ^ np.uint64((seed * 0x9E3779B9) & 0xFFFFFFFF))
    k = pcg_hash(k ^ (iy.astype(np.int64).astype(np.uint64) & m))
    k = pcg_hash(k ^ (iz.astype(np.int64).astype(np.uint64) & m))
    k = pcg_hash(k ^ np.uint64(octave))
    return k.astype(np.float64) / 2147483648.0 - 1.0


def value_noise3(p: np.ndarray, octave: int, seed: int) -> np.ndarray:
    f = np.floor(p)
    i = f.astype(np.int64)
    t = p - f
    w = t * t * (3.0 - 2.0 * t)
    out = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                val = _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, octave, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += val * wx * wy * wz
    return out


def fbm3(p: np.ndarray, seed: int, octaves: int = 4, base_freq: float = 2.0, lacunarity: float = 2.0,
         gain: float = 0.5) -> np.ndarray:
    total = np.zeros(p.shape[0])
    amp, freq, norm = 1.0, base_freq, 0.0
    for o in range(octaves):
        total += amp * value_noise3(p * freq, o, seed)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return total / norm


def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit icosphere: (vertices float64 [V,3], faces int64 [F,3]) with F = 20*4^subdiv."""
    g = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0], [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(subdiv):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        edges = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        edges.sort(axis=1)
        uniq, inv = np.unique(edges, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        m = f.shape[0]
        ab, bc, ca = inv[:m] + base, inv[m:2 * m] + base, inv[2 * m:] + base
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    return v, f


def icosphere_fbm(subdiv: int = 6, radius: float = 1.0, amplitude: float = 0.1, seed: int = 1) -> np.ndarray:
    """Displaced icosphere as a float32 triangle soup (9 floats per triangle)."""
    v, f = icosphere(subdiv)
    r = radius * (1.0 + amplitude * fbm3(v, seed))
    v = v * r[:, None]
    return v[f].reshape(-1).astype(np.float32)


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    g = rng.standard_normal((n, 3))
    return g / np.linalg.norm(g, axis=1, keepdims=True)


def sphere_rays(n: int, seed: int = 1, origin_radius: float = 3.0) -> tuple[np.ndarray, np.ndarray]:
    """Rays from a sphere of radius 3 aimed at uniform points of the unit ball (float32 xyz)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    org = origin_radius * _unit_vectors(rng, n)
    tgt = _unit_vectors(rng, n) * (rng.random(n) ** (1.0 / 3.0))[:, None]
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


CFG1_SUBDIV = 6
CFG1_RAYS = 1 << 20


def config1(n_rays: int = CFG1_RAYS, seed: int = 1) -> dict:
    """Plain traversal workload: 81,920-triangle displaced icosphere and 2^20 sphere rays."""
    pos = icosphere_fbm(CFG1_SUBDIV, 1.0, 0.1, seed)
    org, d = sphere_rays(n_rays, seed + 1)
    return {"positions": pos, "origins": org, "dirs": d}


# ---------------------------------------------------------------------------
# Progressive path-trace workload (1440x900, 64 spp, depth 8)
# ---------------------------------------------------------------------------
def direction_from_uv(u, v):
    """environment.ts:88-93 (v = 0 is straight up)."""
    phi = (u - 0.5) * 2 * np.pi
    theta = v * np.pi
    s = np.sin(theta)
    return np.stack([s * np.cos(phi), np.cos(theta), s * np.sin(phi)], -1)


def synthetic_hdri(width: int = 1024, height: int = 512, sun_elevation: float = 35.0, sun_azimuth: float = 40.0,
                   sigma_deg: float = 0.5, peak: float = 5e4) -> np.ndarray:
    """Sky gradient plus a Gaussian sun lobe, float32 rgb rows top to bottom.

    Sky: L = 0.2 + 1.8 * (elevation + 90) / 180 (linear in elevation angle,
    0.2 straight down to 2.0 at the zenith), tinted (0.85, 0.92, 1.0).
    Sun: peak * exp(-a^2 / (2 sigma^2)) with a the angle to the sun direction
    (elevation 35 deg, azimuth 40 deg), sigma = 0.5 deg, tinted (1, 0.96, 0.9).
    Texels are evaluated at their centres.
    """
    xs = (np.arange(width) + 0.5) / width
    ys = (np.arange(height) + 0.5) / height
    u, v = np.meshgrid(xs, ys)
    d = direction_from_uv(u, v)
    elev = np.degrees(np.arcsin(np.clip(d[..., 1], -1, 1)))
    sky = 0.2 + 1.8 * (elev + 90.0) / 180.0
    el, az = np.radians(sun_elevation), np.radians(sun_azimuth)
    sd = np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    ang = np.degrees(np.arccos(np.clip(d @ sd, -1, 1)))
    sun = peak * np.exp(-(ang ** 2) / (2 * sigma_deg ** 2))
    rgb = sky[..., None] * np.array([0.85, 0.92, 1.0]) + sun[..., None] * np.array([1.0, 0.96, 0.9])
    return rgb.astype(np.float32).reshape(-1)


def displace_along_normals(mesh, amplitude: float, seed: int, freq_scale: float = 1.0):
    """Moves each vertex along its (ring) normal by amplitude * fbm3(position)."""
    from .scene import Mesh
    p = mesh.positions.reshape(-1, 3).astype(np.float64)
    n = mesh.normals.reshape(-1, 3).astype(np.float64)
    d = amplitude * fbm3(p * freq_scale, seed)
    q = p + n * d[:, None]
    return Mesh(q.reshape(-1).astype(np.float32), mesh.normals.copy(), mesh.uvs)


KNOT_PQ = [(2, 3), (3, 4), (2, 5), (3, 5), (3, 7), (2, 7)]


def random_material(rng: np.random.Generator):
    """Per-mesh layered BSDF draw: albedo U[0.05,0.95]^3, roughness U[0.02,1],
    metallic Bernoulli(0.3), clearcoat Bernoulli(0.5) with coat roughness
    U[0,0.3], sheen s ~ U[0,1] as sheenColor (s,s,s) with sheenRoughness 0.5,
    transmission Bernoulli(0.2) with IOR 1.5 and Beer-Lambert absorption
    (attenuationColor U[0.5,1]^3 over distance 1)."""
    from .scene import Material
    albedo = tuple(float(x) for x in rng.uniform(0.05, 0.95, 3))
    rough = float(rng.uniform(0.02, 1.0))
    metallic = 1.0 if rng.random() < 0.3 else 0.0
    coat = rng.random() < 0.5
    coat_r = float(rng.uniform(0.0, 0.3))
    sheen = float(rng.uniform(0.0, 1.0))
    trans = rng.random() < 0.2
    att = tuple(float(x) for x in rng.uniform(0.5, 1.0, 3))
    return Material(albedo=albedo, roughness=rough, metallic=metallic, clearcoat=1.0 if coat else 0.0,
                    clearcoat_roughness=coat_r if coat else 0.0, sheen_color=(sheen, sheen, sheen),
                    sheen_roughness=0.5, transmission=1.0 if trans else 0.0, ior=1.5,
                    attenuation_color=att if trans else (1.0, 1.0, 1.0),
                    attenuation_distance=1.0 if trans else float("inf"))


def knot_scene(seed: int = 7, knots: int = 6, tri_range=(30000, 40000), radial: int = 40,
               hdri_size=(1024, 512)):
    """Seeded displaced torus knots (30-40k triangles each) on a 2-triangle
    ground quad, one emissive quad light, and the synthetic HDRI."""
    from . import scene as S
    rng = np.random.Generator(np.random.PCG64(seed))
    meshes = []
    for k in range(knots):
        target = int(rng.integers(tri_range[0], tri_range[1] + 1))
        tubular = max(8, target // (2 * radial))
        p, q = KNOT_PQ[k % len(KNOT_PQ)]
        m = S.torus_knot(p=p, q=q, radius=1.0, tube=0.3, tubular_segments=tubular, radial_segments=radial)
        m = displace_along_normals(m, 0.06, seed * 100 + k, 3.0)
        col, row = k % 3, k // 3
        m = S.transform(m, translate=(-3.6 + 3.6 * col, 1.15, -1.8 + 3.6 * row), scale=1.0,
                        rotate_y=float(rng.uniform(0, 2 * np.pi)))
        meshes.append(S.MeshInstance(m, random_material(rng)))
    meshes.append(S.MeshInstance(S.floor((0.0, 0.0, 0.0), 12.0), S.matte((0.5, 0.5, 0.5), 1.0)))
    # Winding chosen so the lamp's normal points down (-Y).
    lamp = S.quad((-1.0, 5.5, -1.0), (1.0, 5.5, -1.0), (1.0, 5.5, 1.0), (-1.0, 5.5, 1.0))
    meshes.append(S.MeshInstance(lamp, S.light((1.0, 0.9, 0.75), 12.0)))
    w, h = hdri_size
    env = S.Environment(map=(w, h, synthetic_hdri(w, h)), rotation=0.0, intensity=1.0, background=True)
    cam = S.Camera(position=(0.0, 5.0, 11.5), target=(0.0, 0.9, 0.0), fov=40.0, aperture=0.0, focus_distance=11.0)
    return S.Scene(camera=cam, meshes=meshes, env=env, name="knots")


PT_WIDTH, PT_HEIGHT, PT_SPP, PT_DEPTH = 1440, 900, 64, 8


# ---------------------------------------------------------------------------
# Differentiable-rendering workload (512x512, 32 spp, depth 6)
# ---------------------------------------------------------------------------
DIFF_WIDTH, DIFF_HEIGHT, DIFF_SPP, DIFF_DEPTH = 512, 512, 32, 6
DIFF_ENV = (256, 128)


def diff_material(rng: np.random.Generator):
    """Opaque layered BSDF with continuous parameters: albedo U[0.05,0.95]^3,
    roughness U[0.15,0.9], metallic U[0,1], clearcoat Bernoulli(0.5) with coat
    roughness U[0,0.3], sheen s ~ U[0,1] as (s,s,s) with roughness 0.5.  No
    transmission: the reference's opaque/glass choice (trace.frag:186) is a
    stochastic blend, not an importance-sampled lobe, and needs a boundary
    term that this adjoint does not model."""
    from .scene import Material
    coat = rng.random() < 0.5
    sheen = float(rng.uniform(0, 1))
    return Material(albedo=tuple(float(x) for x in rng.uniform(0.05, 0.95, 3)), roughness=float(rng.uniform(0.15, 0.9)),
                    metallic=float(rng.uniform(0, 1)), clearcoat=1.0 if coat else 0.0,
                    clearcoat_roughness=float(rng.uniform(0, 0.3)) if coat else 0.0,
                    sheen_color=(sheen, sheen, sheen), sheen_roughness=0.5)


def spheres_scene(seed: int = 11):
    """16 fBm-displaced icospheres (subdivision 3, 1,280 triangles each: 20,480
    in all) on a 4x4 grid, 16 materials, lit by a 256x128 synthetic HDRI (sky
    gradient plus a sun lobe with sigma 3 deg, peak 500)."""
    from . import scene as S
    rng = np.random.Generator(np.random.PCG64(seed))
    meshes = []
    v, f = icosphere(3)
    for k in range(16):
        r = 0.55 * (1.0 + 0.12 * fbm3(v * 1.5, seed * 31 + k))
        pos = (v * r[:, None])[f]
        nrm = v[f]
        m = S.Mesh(pos.reshape(-1).astype(np.float32), nrm.reshape(-1).astype(np.float32))
        col, row = k % 4, k // 4
        m = S.transform(m, translate=(-1.8 + 1.2 * col, -1.8 + 1.2 * row, 0.0))
        meshes.append(S.MeshInstance(m, diff_material(rng)))
    w, h = DIFF_ENV
    env = S.Environment(map=(w, h, synthetic_hdri(w, h, sun_elevation=40.0, sun_azimuth=-30.0, sigma_deg=3.0,
                                                  peak=500.0)), rotation=0.0, intensity=1.0, background=True)
    cam = S.Camera(position=(0.0, 0.0, 6.5), target=(0.0, 0.0, 0.0), fov=45.0, aperture=0.0, focus_distance=6.5)
    return S.Scene(camera=cam, meshes=meshes, env=env, name="spheres")


def perturb_params(materials, env_rgb, seed: int = 5, amount: float = 0.2):
    """Seeded ±20% perturbation of every material's albedo, roughness and
    metallic (clamped to their valid ranges) and of every env texel channel."""
    from dataclasses import replace
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in materials:
        a = np.clip(np.array(m.albedo) * rng.uniform(1 - amount, 1 + amount, 3), 0.0, 1.0)
        out.append(replace(m, albedo=tuple(float(x) for x in a),
                           roughness=float(np.clip(m.roughness * rng.uniform(1 - amount, 1 + amount), 0.05, 1.0)),
                           metallic=float(np.clip(m.metallic * rng.uniform(1 - amount, 1 + amount), 0.0, 1.0))))
    env = (np.asarray(env_rgb, np.float32) * rng.uniform(1 - amount, 1 + amount, np.asarray(env_rgb).size)
           ).astype(np.float32)
    return out, env


# ---------------------------------------------------------------------------
# Batched multi-scene workload: 64 seeded scenes of 1k..1.5M triangles
# ---------------------------------------------------------------------------
BATCH_SCENES = 64
BATCH_RESOLUTIONS = ((256, 256), (640, 480), (1280, 720), (1920, 1080))
BATCH_TRI_RANGE = (1000, 1500000)
BATCH_DEPTH = 8
WORLEY_SIZE = 256


def worley_alpha(size: int = WORLEY_SIZE, cells: int = 8, seed: int = 3, radius: float = 0.35) -> np.ndarray:
    """Tileable 0/1 Worley mask as an RGBA8 image (size x size x 4): one
    jittered feature point per cell of a cells x cells grid; alpha = 0 where
    the distance to the nearest point (F1, in cell units, with wrap-around) is
    below `radius`, 255 elsewhere; RGB white."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = (np.stack(np.meshgrid(np.arange(cells), np.arange(cells), indexing="ij"), -1)
           + rng.random((cells, cells, 2)))  # [cx, cy] -> point in cell units
    c = (np.arange(size) + 0.5) * cells / size
    px, py = np.meshgrid(c, c, indexing="xy")  # py = row (v), px = column (u)
    f1 = np.full((size, size), np.inf)
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            cx = np.floor(px).astype(np.int64) + ox
            cy = np.floor(py).astype(np.int64) + oy
            p = pts[np.mod(cx, cells), np.mod(cy, cells)]
            dx = px - (cx + (p[..., 0] - np.floor(p[..., 0])))
            dy = py - (cy + (p[..., 1] - np.floor(p[..., 1])))
            f1 = np.minimum(f1, np.sqrt(dx * dx + dy * dy))
    img = np.full((size, size, 4), 255, np.uint8)
    img[..., 3] = np.where(f1 < radius, 0, 255).astype(np.uint8)
    return img


def _grid_displacement(amplitude: float, seed: int, freq: float):
    """fBm offsets (3 octaves) of ring/grid vertices along their normals."""
    def f(points, normals):
        return amplitude * fbm3(points * freq, seed, octaves=3)
    return f


def batch_scene(index: int, seed: int = 2024):
    """Scene `index` of the batched workload: a log-uniform triangle budget in
    [1k, 1.5M] split over 1-4 fBm-displaced torus knots or UV spheres (each
    alpha-tested with probability 0.1: mask mode, cutoff 0.5, the Worley
    texture), a 2-triangle ground, a quad light and the synthetic HDRI at a
    per-scene rotation; materials as random_material.  Returns (scene, width,
    height, spp) with resolution uniform over BATCH_RESOLUTIONS and spp
    uniform in 1..64."""
    from . import scene as S
    rng = np.random.Generator(np.random.PCG64([seed, index]))
    lo, hi = BATCH_TRI_RANGE
    total = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    width, height = BATCH_RESOLUTIONS[int(rng.integers(0, len(BATCH_RESOLUTIONS)))]
    spp = int(rng.integers(1, 65))
    k = int(rng.integers(1, 5))
    meshes, textures = [], [worley_alpha()]
    for j in range(k):
        budget = max(64, total // k)
        disp = _grid_displacement(0.05, seed * 1000 + index * 8 + j, 3.0)
        if rng.random() < 0.5:
            radial = int(min(64, max(8, round(math.sqrt(budget / 8)))))
            tubular = max(8, budget // (2 * radial))
            p, q = KNOT_PQ[int(rng.integers(0, len(KNOT_PQ)))]
            m = S.torus_knot(p, q, 1.0, 0.3, tubular, radial, uvs=True, displace=disp)
        else:
            rings = max(4, int(round(math.sqrt(budget / 4))))
            segments = max(8, budget // (2 * rings))
            m = S.uv_sphere(rings, segments, 1.1, displace=disp)
        ang = 2 * math.pi * j / k
        m = S.transform(m, translate=(2.6 * math.cos(ang) if k > 1 else 0.0, 1.2,
                                      2.6 * math.sin(ang) if k > 1 else 0.0),
                        rotate_y=float(rng.uniform(0, 2 * math.pi)))
        mat = random_material(rng)
        if rng.random() < 0.1:
            mat = replace(mat, alpha_mode="mask", alpha_cutoff=0.5, base_color_map=0, transmission=0.0)
        meshes.append(S.MeshInstance(m, mat))
    meshes.append(S.MeshInstance(S.floor((0.0, 0.0, 0.0), 10.0), S.matte((0.5, 0.5, 0.5), 1.0)))
    lamp = S.quad((-1.0, 5.5, -1.0), (1.0, 5.5, -1.0), (1.0, 5.5, 1.0), (-1.0, 5.5, 1.0))
    meshes.append(S.MeshInstance(lamp, S.light((1.0, 0.9, 0.75), 12.0)))
    env = S.Environment(map=(1024, 512, _batch_hdri()), rotation=float(rng.uniform(0, 360)), intensity=1.0,
                        background=True)
    cam = S.Camera(position=(0.0, 4.5, 9.0), target=(0.0, 1.0, 0.0), fov=45.0, aperture=0.0, focus_distance=9.0)
    return S.Scene(camera=cam, meshes=meshes, env=env, name=f"batch{index}", textures=textures), width, height, spp


_HDRI_CACHE: dict = {}


def _batch_hdri():
    if "map" not in _HDRI_CACHE:
        _HDRI_CACHE["map"] = synthetic_hdri(1024, 512)
    return _HDRI_CACHE["map"]
