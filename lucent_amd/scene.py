"""Scene description and packing — the host mirror of src/scene/*.ts.

* ``Material`` / ``material()`` / ``pack_materials``: material.ts:17-149
  (10 RGBA32F texels per material).
* ``Mesh`` generators ``torus_knot``, ``transform``, ``quad``, ``floor``,
  ``box``, ``from_triangles``, ``merge_meshes``: mesh.ts:9-182.
* ``pack_scene``: packScene (scene.ts:162-305) — triangles in BVH leaf order,
  GEO/ATTR texel layouts, deduplicated materials, emissive 2-triangle quads
  turned into explicit quad lights (quadFromMesh, scene.ts:312-324).
* ``Camera``, ``Sun``, ``Sky``, ``Environment``: scene.ts:10-38,
  environment.ts:12-31.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Optional

import numpy as np

from .bvh import BVH, build_bvh

MAX_SPHERES = 64
MAX_LIGHTS = 8
GEO_TEXELS = 3
ATTR_TEXELS = 4
MATERIAL_TEXELS = 10
ALPHA_CODE = {"opaque": 0, "mask": 1, "blend": 2}


@dataclass(frozen=True)
class Material:
    albedo: tuple = (0.8, 0.8, 0.8)
    metallic: float = 0.0
    roughness: float = 0.5
    transmission: float = 0.0
    ior: float = 1.5
    emission: tuple = (0.0, 0.0, 0.0)
    checker_scale: float = 0.0
    albedo2: tuple = (0.2, 0.2, 0.2)
    alpha: float = 1.0
    alpha_mode: str = "opaque"
    alpha_cutoff: float = 0.5
    clearcoat: float = 0.0
    clearcoat_roughness: float = 0.0
    sheen_color: tuple = (0.0, 0.0, 0.0)
    sheen_roughness: float = 0.0
    thin_walled: bool = False
    attenuation_color: tuple = (1.0, 1.0, 1.0)
    attenuation_distance: float = math.inf
    base_color_map: int = -1
    metal_rough_map: int = -1
    normal_map: int = -1
    emissive_map: int = -1
    normal_scale: float = 1.0
    uv_transform: tuple = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0)


def material(**overrides) -> Material:
    return Material(**overrides)


def matte(albedo, roughness=1.0) -> Material:
    return Material(albedo=tuple(albedo), roughness=roughness)


def metal(albedo, roughness=0.1) -> Material:
    return Material(albedo=tuple(albedo), metallic=1.0, roughness=roughness)


def glass(albedo=(1.0, 1.0, 1.0), ior=1.5, roughness=0.0) -> Material:
    return Material(albedo=tuple(albedo), transmission=1.0, ior=ior, roughness=roughness)


def light(color, strength) -> Material:
    return Material(albedo=(0.0, 0.0, 0.0), emission=tuple(c * strength for c in color))


def is_emissive(m: Material) -> bool:
    return m.emission[0] > 0 or m.emission[1] > 0 or m.emission[2] > 0


def needs_alpha_test(m: Material) -> bool:
    return m.alpha_mode != "opaque" and (m.alpha < 1 or m.base_color_map >= 0)


def pack_materials(materials) -> np.ndarray:
    """material.ts:132-149 — 40 float32 per material."""
    out = np.zeros(len(materials) * MATERIAL_TEXELS * 4, np.float32)
    for i, m in enumerate(materials):
        o = i * 40
        uv = m.uv_transform
        att = m.attenuation_distance if math.isfinite(m.attenuation_distance) else 0.0
        out[o:o + 40] = [*m.albedo, m.metallic,
                         m.roughness, m.ior, m.transmission, m.checker_scale,
                         *m.emission, m.alpha,
                         *m.albedo2, ALPHA_CODE[m.alpha_mode],
                         m.base_color_map, m.metal_rough_map, m.normal_map, m.emissive_map,
                         m.normal_scale, m.clearcoat, m.clearcoat_roughness, m.alpha_cutoff,
                         *m.sheen_color, m.sheen_roughness,
                         *m.attenuation_color, att,
                         uv[0], uv[1], uv[2], 1.0 if m.thin_walled else 0.0,
                         uv[3], uv[4], uv[5], 0.0]
    return out


# ---------------------------------------------------------------------------
# Meshes (mesh.ts)
# ---------------------------------------------------------------------------
@dataclass
class Mesh:
    positions: np.ndarray  # float32, 9 per triangle
    normals: np.ndarray  # float32, 9 per triangle
    uvs: Optional[np.ndarray] = None  # float32, 6 per triangle

    @property
    def triangle_count(self) -> int:
        return self.positions.size // 9


def _normalize_rows(v: np.ndarray) -> np.ndarray:
    n = np.sqrt((v * v).sum(axis=-1, keepdims=True))
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n == 0, 0.0, v * (1.0 / np.where(n == 0, 1.0, n)))


def from_triangles(tris) -> Mesh:
    t = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    n = _normalize_rows(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]))
    nrm = np.repeat(n[:, None, :], 3, axis=1)
    return Mesh(t.reshape(-1).astype(np.float32), nrm.reshape(-1).astype(np.float32))


def quad(a, b, c, d) -> Mesh:
    return from_triangles([[a, b, c], [a, c, d]])


def floor(center, half) -> Mesh:
    x, y, z = center
    return quad([x - half, y, z + half], [x + half, y, z + half], [x + half, y, z - half], [x - half, y, z - half])


def merge_meshes(meshes) -> Mesh:
    pos = np.concatenate([m.positions for m in meshes]) if meshes else np.zeros(0, np.float32)
    nrm = np.concatenate([m.normals for m in meshes]) if meshes else np.zeros(0, np.float32)
    uvs = None
    if any(m.uvs is not None for m in meshes):
        uvs = np.concatenate([m.uvs if m.uvs is not None else np.zeros(m.triangle_count * 6, np.float32)
                              for m in meshes])
    return Mesh(pos.astype(np.float32), nrm.astype(np.float32), uvs)


def box(mn, mx, inward=False) -> Mesh:
    x0, y0, z0 = mn
    x1, y1, z1 = mx
    faces = [[(x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)],
             [(x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0)],
             [(x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1)],
             [(x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)],
             [(x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)],
             [(x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)]]
    return merge_meshes([quad(d, c, b, a) if inward else quad(a, b, c, d) for a, b, c, d in faces])


def torus_knot(p=2, q=3, radius=1.0, tube=0.32, tubular_segments=200, radial_segments=20, uvs=False,
               displace=None) -> Mesh:
    """mesh.ts:95-156: (p, q) knot, ring normals, 2*tubular*radial triangles.
    uvs=True adds (i / tubular, j / radial) texture coordinates per vertex (the
    reference's knot has none).  displace(points, normals) -> offsets, if
    given, moves the ring vertices along their normals before triangulation."""
    def curve(t):
        cu, su = np.cos(t), np.sin(t)
        qp = (q / p) * t
        cs = np.cos(qp)
        return np.stack([radius * (2 + cs) * 0.5 * cu, radius * (2 + cs) * su * 0.5, radius * np.sin(qp) * 0.5], -1)

    i = np.arange(tubular_segments, dtype=np.float64)
    u = (i / tubular_segments) * p * math.pi * 2
    P1, P2 = curve(u), curve(u + 0.01)
    T = P2 - P1
    N0 = P2 + P1
    B = _normalize_rows(np.cross(T, N0))
    N = _normalize_rows(np.cross(B, T))
    j = np.arange(radial_segments, dtype=np.float64)
    v = (j / radial_segments) * math.pi * 2
    cx = -tube * np.cos(v)
    cy = tube * np.sin(v)
    off = N[:, None, :] * cx[None, :, None] + B[:, None, :] * cy[None, :, None]
    ring = P1[:, None, :] + off
    ringn = _normalize_rows(off)
    if displace is not None:
        flat = ring.reshape(-1, 3)
        d = np.asarray(displace(flat, ringn.reshape(-1, 3)), np.float64).reshape(ring.shape[:2])
        ring = ring + ringn * d[..., None]
    ii, jj = np.meshgrid(np.arange(tubular_segments), np.arange(radial_segments), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)
    i1 = (ii + 1) % tubular_segments
    j1 = (jj + 1) % radial_segments
    tri_a = np.stack([np.stack([ii, jj], -1), np.stack([i1, jj], -1), np.stack([i1, j1], -1)], 1)
    tri_b = np.stack([np.stack([ii, jj], -1), np.stack([i1, j1], -1), np.stack([ii, j1], -1)], 1)
    tris = np.stack([tri_a, tri_b], 1).reshape(-1, 3, 2)
    pos = ring[tris[..., 0], tris[..., 1]]
    nrm = ringn[tris[..., 0], tris[..., 1]]
    uv = None
    if uvs:
        # Seam-free: the wrapped index (i + 1) % n keeps its unwrapped coordinate.
        ui = np.stack([np.stack([ii, ii + 1, ii + 1], -1), np.stack([ii, ii + 1, ii], -1)], 1).reshape(-1, 3)
        vi = np.stack([np.stack([jj, jj, jj + 1], -1), np.stack([jj, jj + 1, jj + 1], -1)], 1).reshape(-1, 3)
        uv = np.stack([ui / tubular_segments, vi / radial_segments], -1).reshape(-1).astype(np.float32)
    return Mesh(pos.reshape(-1).astype(np.float32), nrm.reshape(-1).astype(np.float32), uv)


def uv_sphere(rings: int, segments: int, radius: float = 1.0, displace=None) -> Mesh:
    """Latitude-longitude sphere, 2 * rings * segments triangles (the pole rows
    degenerate), with (longitude, latitude) texture coordinates; displace as
    for torus_knot, along the radial normal."""
    th = np.linspace(0.0, math.pi, rings + 1)
    ph = np.linspace(0.0, 2 * math.pi, segments + 1)
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    n = np.stack([st * np.cos(ph)[None, :], np.broadcast_to(ct, (rings + 1, segments + 1)),
                  st * np.sin(ph)[None, :]], -1)
    r = np.full((rings + 1, segments + 1), radius)
    if displace is not None:
        r = r + np.asarray(displace((n * radius).reshape(-1, 3), n.reshape(-1, 3)), np.float64).reshape(r.shape)
        r[:, -1] = r[:, 0]  # the seam column is the same points
    pts = n * r[..., None]
    ii, jj = np.meshgrid(np.arange(rings), np.arange(segments), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)
    a = np.stack([np.stack([ii, jj], -1), np.stack([ii + 1, jj], -1), np.stack([ii + 1, jj + 1], -1)], 1)
    b = np.stack([np.stack([ii, jj], -1), np.stack([ii + 1, jj + 1], -1), np.stack([ii, jj + 1], -1)], 1)
    tris = np.stack([a, b], 1).reshape(-1, 3, 2)
    pos = pts[tris[..., 0], tris[..., 1]]
    nrm = n[tris[..., 0], tris[..., 1]]
    uv = np.stack([tris[..., 1] / segments, tris[..., 0] / rings], -1)
    return Mesh(pos.reshape(-1).astype(np.float32), nrm.reshape(-1).astype(np.float32),
                uv.reshape(-1).astype(np.float32))


def transform(mesh: Mesh, translate=(0.0, 0.0, 0.0), scale=1.0, rotate_y=0.0) -> Mesh:
    """mesh.ts:159-182: scale, rotate about Y, translate (normals only rotate)."""
    c, s = math.cos(rotate_y), math.sin(rotate_y)
    p = mesh.positions.reshape(-1, 3).astype(np.float64) * scale
    n = mesh.normals.reshape(-1, 3).astype(np.float64)
    pos = np.stack([p[:, 0] * c + p[:, 2] * s + translate[0], p[:, 1] + translate[1],
                    -p[:, 0] * s + p[:, 2] * c + translate[2]], -1)
    nrm = np.stack([n[:, 0] * c + n[:, 2] * s, n[:, 1], -n[:, 0] * s + n[:, 2] * c], -1)
    return Mesh(pos.reshape(-1).astype(np.float32), nrm.reshape(-1).astype(np.float32), mesh.uvs)


def quad_from_mesh(mesh: Mesh):
    """scene.ts:312-324: (position, edge1, edge2) of a `quad()` mesh, else None."""
    if mesh.triangle_count != 2:
        return None
    v = mesh.positions.reshape(6, 3).astype(np.float64)
    a, b, c, a2, c2, d = v

    def same(p, q):
        return math.sqrt(float(((p - q) ** 2).sum())) < 1e-6

    if not same(a, a2) or not same(c, c2):
        return None
    e1, e2 = b - a, d - a
    corner = a + e1 + e2
    if not same(corner, c) or np.linalg.norm(np.cross(e1, e2)) < 1e-9:
        return None
    return a, e1, e2


# ---------------------------------------------------------------------------
# Scene (scene.ts)
# ---------------------------------------------------------------------------
@dataclass
class Camera:
    position: tuple
    target: tuple
    fov: float = 45.0
    aperture: float = 0.0
    focus_distance: float = 5.0


@dataclass
class Sun:
    elevation: float = 45.0
    azimuth: float = 0.0
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 0.0
    angular_radius: float = 0.27


@dataclass
class Sky:
    enabled: bool = False
    zenith: tuple = (0.0, 0.0, 0.0)
    horizon: tuple = (0.0, 0.0, 0.0)
    ground: tuple = (0.0, 0.0, 0.0)
    intensity: float = 0.0


@dataclass
class Environment:
    """environment.ts:12-22; ``map`` is (width, height, float32 rgb rows top to bottom) or None."""
    map: Optional[tuple] = None
    rotation: float = 0.0
    intensity: float = 1.0
    background: bool = True
    backdrop: tuple = (0.02, 0.02, 0.025)


@dataclass
class Ground:
    enabled: bool = False
    height: float = 0.0
    material: Material = field(default_factory=lambda: matte((0.5, 0.5, 0.5)))


@dataclass
class Sphere:
    center: tuple
    radius: float
    material: Material


@dataclass
class MeshInstance:
    mesh: Mesh
    material: Material


@dataclass
class Scene:
    camera: Camera
    meshes: list
    spheres: list = field(default_factory=list)
    sun: Sun = field(default_factory=Sun)
    sky: Sky = field(default_factory=Sky)
    env: Environment = field(default_factory=Environment)
    ground: Ground = field(default_factory=Ground)
    name: str = ""
    # SceneTexture images (scene.ts:59-62): uint8 (H, W, 4) RGBA arrays, rows in v order.
    textures: list = field(default_factory=list)


@dataclass
class PackedScene:
    tri_geometry: np.ndarray
    tri_attributes: np.ndarray
    tri_count: int
    bvh_nodes: np.ndarray
    bvh_node_count: int
    tri_indices: np.ndarray
    materials: np.ndarray
    material_count: int
    spheres: np.ndarray
    sphere_materials: np.ndarray
    sphere_lights: np.ndarray
    sphere_count: int
    lights: np.ndarray
    light_count: int
    ground_material: int
    ground_material_active: int  # uGroundMaterial: -1 while the ground is disabled
    ground_height: float
    bvh: BVH
    positions: np.ndarray
    material_list: list
    color_layers: list = field(default_factory=list)  # scene texture index per sRGB layer
    data_layers: list = field(default_factory=list)   # scene texture index per linear layer


def pack_scene(scene: Scene, prebuilt: Optional[BVH] = None) -> PackedScene:
    """packScene (scene.ts:162-305)."""
    if len(scene.spheres) > MAX_SPHERES:
        raise ValueError(f"scene has {len(scene.spheres)} spheres; max is {MAX_SPHERES}")
    materials: list = []
    ids: dict = {}

    def material_index(m: Material) -> int:
        key = id(m)
        if key not in ids:
            ids[key] = len(materials)
            materials.append(m)
        return ids[key]

    lights: list = []

    def add_light(l) -> int:
        if len(lights) >= MAX_LIGHTS:
            return -1
        lights.append(l)
        return len(lights) - 1

    merged = merge_meshes([m.mesh for m in scene.meshes])
    n = merged.triangle_count
    tri_mat = np.zeros(n, np.int32)
    tri_light = np.full(n, -1, np.int32)
    tri_alpha = np.zeros(n, np.uint8)
    cursor = 0
    for inst in scene.meshes:
        k = inst.mesh.triangle_count
        tri_mat[cursor:cursor + k] = material_index(inst.material)
        if needs_alpha_test(inst.material):
            tri_alpha[cursor:cursor + k] = 1
        q = quad_from_mesh(inst.mesh) if is_emissive(inst.material) else None
        if q is not None:
            li = add_light(("quad", q[0], 0.0, q[1], q[2], inst.material.emission))
            tri_light[cursor:cursor + k] = li
        cursor += k
    if prebuilt is not None and prebuilt.tri_indices.size != n:
        raise ValueError(f"prebuilt BVH covers {prebuilt.tri_indices.size} triangles; scene has {n}")
    bvh = prebuilt if prebuilt is not None else build_bvh(merged.positions)
    src = bvh.tri_indices.astype(np.int64)
    pos = merged.positions.reshape(-1, 3, 3)[src]
    nrm = merged.normals.reshape(-1, 3, 3)[src]
    geo = np.zeros((n, 3, 4), np.float32)
    geo[:, :, :3] = pos
    geo[:, 0, 3] = tri_mat[src]
    geo[:, 1, 3] = tri_light[src]
    geo[:, 2, 3] = tri_alpha[src]
    attr = np.zeros((n, 4, 4), np.float32)
    attr[:, :3, :3] = nrm
    if merged.uvs is not None:
        uv = merged.uvs.reshape(-1, 3, 2)[src]
        attr[:, :3, 3] = uv[:, :, 0]
        attr[:, 3, :3] = uv[:, :, 1]
    spheres = np.zeros(MAX_SPHERES * 4, np.float32)
    sphere_mats = np.zeros(MAX_SPHERES, np.int32)
    sphere_lights = np.full(MAX_SPHERES, -1, np.int32)
    for i, s in enumerate(scene.spheres):
        spheres[i * 4:i * 4 + 4] = [*s.center, s.radius]
        sphere_mats[i] = material_index(s.material)
        if is_emissive(s.material):
            sphere_lights[i] = add_light(("sphere", np.array(s.center, np.float64), s.radius, np.zeros(3),
                                          np.zeros(3), s.material.emission))
    ground_mat = material_index(scene.ground.material)
    lt = np.zeros(MAX_LIGHTS * 16, np.float32)
    for i, (kind, p, r, e1, e2, rad) in enumerate(lights):
        o = i * 16
        lt[o:o + 16] = [*p, 0 if kind == "sphere" else 1, *e1, r, *e2, 0, *rad, 0]
    # layerMap (scene.ts:136-153, 272-283): each referenced texture gets one
    # layer of its array, in first-use order; materials store layer indices.
    color_layers: list = []
    data_layers: list = []

    def layer(table: list, tex: int) -> int:
        if tex < 0:
            return -1
        if tex >= len(scene.textures):
            raise ValueError(f"material references texture {tex}, scene has {len(scene.textures)}")
        if tex not in table:
            table.append(tex)
        return table.index(tex)

    materials = [replace(m, base_color_map=layer(color_layers, m.base_color_map),
                         emissive_map=layer(color_layers, m.emissive_map),
                         metal_rough_map=layer(data_layers, m.metal_rough_map),
                         normal_map=layer(data_layers, m.normal_map)) for m in materials]
    return PackedScene(
        tri_geometry=geo.reshape(-1), tri_attributes=attr.reshape(-1), tri_count=n, bvh_nodes=bvh.nodes,
        bvh_node_count=bvh.node_count, tri_indices=bvh.tri_indices, materials=pack_materials(materials),
        material_count=len(materials), spheres=spheres, sphere_materials=sphere_mats, sphere_lights=sphere_lights,
        sphere_count=len(scene.spheres), lights=lt, light_count=len(lights), ground_material=ground_mat,
        ground_material_active=ground_mat if scene.ground.enabled else -1, ground_height=scene.ground.height,
        bvh=bvh, positions=merged.positions, material_list=materials, color_layers=color_layers,
        data_layers=data_layers)


TEXTURE_BUDGET_BYTES = 256 * 1024 * 1024


def _next_pow2(n: int) -> int:
    p = 1
    while p < n:
        p *= 2
    return p


def _lerp_weights(n_src: int, n_dst: int) -> np.ndarray:
    """(n_dst, n_src) LINEAR-filter weights sampling n_src texels at the
    centres of n_dst target texels, clamped to the edge."""
    x = (np.arange(n_dst) + 0.5) * (n_src / n_dst) - 0.5
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    w = np.zeros((n_dst, n_src))
    rows = np.arange(n_dst)
    np.add.at(w, (rows, np.clip(i0, 0, n_src - 1)), 1.0 - f)
    np.add.at(w, (rows, np.clip(i0 + 1, 0, n_src - 1)), f)
    return w


def _halve(a: np.ndarray) -> np.ndarray:
    """Next mip level: 2x2 box average (edge texels repeat on odd sizes)."""
    h, w = a.shape[:2]
    nh, nw = max(1, h // 2), max(1, w // 2)
    y0 = np.minimum(2 * np.arange(nh), h - 1)
    y1 = np.minimum(2 * np.arange(nh) + 1, h - 1)
    x0 = np.minimum(2 * np.arange(nw), w - 1)
    x1 = np.minimum(2 * np.arange(nw) + 1, w - 1)
    return (a[y0][:, x0] + a[y0][:, x1] + a[y1][:, x0] + a[y1][:, x1]) / 4


def resample_rgba8(img: np.ndarray, size: int, srgb: bool = False) -> np.ndarray:
    """One texture-array layer as createTextureArray draws it
    (context.ts:225-273, resample.frag): `texture(src, fragCoord / size)`
    over a mipmapped, edge-clamped source, LINEAR_MIPMAP_LINEAR.  An image of
    the layer size is copied; a smaller one is bilinearly magnified; a
    larger one is sampled from its box-filtered mip chain at
    lod = log2(max(w, h) / size), blending the two nearest levels.  sRGB
    layers filter in linear space, as SRGB8_ALPHA8 textures do."""
    img = np.asarray(img, np.uint8)
    h, w = img.shape[:2]
    if h == size and w == size:
        return img.copy()
    # Unorm channels filter in units of one code (exact for the dyadic
    # weights of 2x scaling, so ties round half to even alike everywhere);
    # sRGB colour channels filter decoded to linear [0, 1].
    v = img.astype(np.float64)
    if srgb:
        c = v[..., :3] / 255.0
        v[..., :3] = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    lod = math.log2(max(w, h) / size)

    def sample(level):
        wy = _lerp_weights(level.shape[0], size)
        wx = _lerp_weights(level.shape[1], size)
        return np.einsum("yi,ijc,xj->yxc", wy, level, wx)

    if lod <= 0:
        out = sample(v)
    else:
        lo = int(math.floor(lod))
        levels = [v]
        while len(levels) <= lo + 1 and max(levels[-1].shape[:2]) > 1:
            levels.append(_halve(levels[-1]))
        a = levels[min(lo, len(levels) - 1)]
        b = levels[min(lo + 1, len(levels) - 1)]
        t = lod - lo
        out = sample(a) * (1 - t) + sample(b) * t
    if srgb:
        c = np.clip(out[..., :3], 0.0, 1.0)
        out[..., :3] = np.where(c < 0.0031308, c * 12.92, 1.055 * np.power(c, 1 / 2.4) - 0.055) * 255.0
    return np.rint(np.clip(out, 0.0, 255.0)).astype(np.uint8)


def texture_arrays(scene: Scene, packed: PackedScene):
    """Renderer.buildTextureArrays (renderer.ts:315-335): the shared layer size
    (next power of two of the largest used image, at most 2048, halved down to
    256 while the arrays exceed the 256 MiB budget) and the two RGBA8 arrays
    as uint32 texels (R in the low byte).  Returns (size, color, data)."""
    used = list(packed.color_layers) + list(packed.data_layers)
    if not used:
        return 0, np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    largest = max(max(np.asarray(scene.textures[i]).shape[:2]) for i in used)
    size = min(2048, _next_pow2(largest))
    while size > 256 and len(used) * size * size * 4 > TEXTURE_BUDGET_BYTES:
        size //= 2

    def array(layers, srgb):
        if not layers:
            return np.zeros(0, np.uint32)
        out = np.stack([resample_rgba8(scene.textures[i], size, srgb) for i in layers])
        return np.ascontiguousarray(out).view(np.uint32).reshape(-1)

    return size, array(packed.color_layers, True), array(packed.data_layers, False)
