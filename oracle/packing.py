"""TEST INFRASTRUCTURE — the oracle's own restatement of the reference's
scene packing, so that a packing mistake in the product cannot hide in a
parity test by being made on both sides.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module; the product never imports it, and it imports nothing from the
product.  It reads a scene *description* by attribute (meshes with
positions/normals/uvs, materials with the reference's Material fields in
snake_case, spheres, ground, textures), the same data a caller hands to
packScene.

Restated here (reference files cited per function):
  * pack_materials / unpack_material   src/scene/material.ts:117-182
  * needs_alpha_test / is_emissive      src/scene/material.ts:106-112
  * merge_meshes                        src/scene/mesh.ts:17-31
  * quad_from_mesh                      src/scene/scene.ts:307-324
  * pack_scene (layerMap included)      src/scene/scene.ts:136-305
  * pick_distance                       src/scene/scene.ts:326-341
  * texture_layers                      src/renderer.ts:314-335 with
                                        src/gl/context.ts:225-273 and
                                        src/shaders/resample.frag
The BVH is the oracle's C restatement of buildBVH (oracle.build_bvh).
"""
from __future__ import annotations

import math
from types import SimpleNamespace

import numpy as np

MAX_SPHERES = 64
MAX_LIGHTS = 8
MATERIAL_FLOATS = 40
ALPHA_MODES = ("opaque", "mask", "blend")
TEXTURE_BUDGET_BYTES = 256 * 1024 * 1024


# ---------------------------------------------------------------------------
# material.ts
# ---------------------------------------------------------------------------
def is_emissive(m) -> bool:
    """material.ts:106-107."""
    e = m.emission
    return e[0] > 0 or e[1] > 0 or e[2] > 0


def needs_alpha_test(m) -> bool:
    """material.ts:110-112."""
    return m.alpha_mode != "opaque" and (m.alpha < 1 or m.base_color_map >= 0)


def _material_row(m, maps) -> list:
    """One material's 10 texels (material.ts:117-131); maps = the four
    texture-layer indices (base colour, metal-rough, normal, emissive)."""
    uv = m.uv_transform
    dist = m.attenuation_distance
    return [
        m.albedo[0], m.albedo[1], m.albedo[2], m.metallic,
        m.roughness, m.ior, m.transmission, m.checker_scale,
        m.emission[0], m.emission[1], m.emission[2], m.alpha,
        m.albedo2[0], m.albedo2[1], m.albedo2[2], float(ALPHA_MODES.index(m.alpha_mode)),
        maps[0], maps[1], maps[2], maps[3],
        m.normal_scale, m.clearcoat, m.clearcoat_roughness, m.alpha_cutoff,
        m.sheen_color[0], m.sheen_color[1], m.sheen_color[2], m.sheen_roughness,
        m.attenuation_color[0], m.attenuation_color[1], m.attenuation_color[2],
        dist if math.isfinite(dist) else 0.0,
        uv[0], uv[1], uv[2], 1.0 if m.thin_walled else 0.0,
        uv[3], uv[4], uv[5], 0.0,
    ]


def pack_materials(materials, layer_maps=None) -> np.ndarray:
    """packMaterials (material.ts:132-149): 40 float32 per material.
    layer_maps[i] overrides material i's four map fields (packScene passes
    texture-array layers there)."""
    rows = []
    for i, m in enumerate(materials):
        maps = layer_maps[i] if layer_maps is not None else (m.base_color_map, m.metal_rough_map, m.normal_map,
                                                              m.emissive_map)
        rows.append(_material_row(m, maps))
    return np.asarray(rows, np.float32).reshape(-1) if rows else np.zeros(0, np.float32)


def unpack_material(data, index: int) -> dict:
    """unpackMaterial (material.ts:152-182), as a dict of the Material fields."""
    f = np.asarray(data, np.float32)[index * MATERIAL_FLOATS:(index + 1) * MATERIAL_FLOATS].astype(np.float64)
    code = int(f[15])
    return {
        "albedo": tuple(f[0:3]), "metallic": f[3], "roughness": f[4], "ior": f[5], "transmission": f[6],
        "checker_scale": f[7], "emission": tuple(f[8:11]), "alpha": f[11], "albedo2": tuple(f[12:15]),
        "alpha_mode": ALPHA_MODES[code] if 0 <= code < 3 else "opaque",
        "base_color_map": f[16], "metal_rough_map": f[17], "normal_map": f[18], "emissive_map": f[19],
        "normal_scale": f[20], "clearcoat": f[21], "clearcoat_roughness": f[22], "alpha_cutoff": f[23],
        "sheen_color": tuple(f[24:27]), "sheen_roughness": f[27], "attenuation_color": tuple(f[28:31]),
        "attenuation_distance": f[31] if f[31] > 0 else math.inf,
        "uv_transform": (f[32], f[33], f[34], f[36], f[37], f[38]), "thin_walled": bool(f[35] > 0.5),
    }


# ---------------------------------------------------------------------------
# mesh.ts / scene.ts
# ---------------------------------------------------------------------------
def merge_meshes(meshes):
    """mergeMeshes (mesh.ts:17-31): concatenation in instance order; uvs kept,
    zero-filled for meshes without them, when any mesh has them."""
    total = sum(np.asarray(m.positions).size for m in meshes)
    positions = np.zeros(total, np.float32)
    normals = np.zeros(total, np.float32)
    has_uv = any(getattr(m, "uvs", None) is not None for m in meshes)
    uvs = np.zeros(total // 9 * 6, np.float32) if has_uv else None
    offset = 0
    for m in meshes:
        p = np.asarray(m.positions, np.float32).reshape(-1)
        positions[offset:offset + p.size] = p
        normals[offset:offset + p.size] = np.asarray(m.normals, np.float32).reshape(-1)
        if has_uv and getattr(m, "uvs", None) is not None:
            uvs[offset // 9 * 6: offset // 9 * 6 + p.size // 9 * 6] = np.asarray(m.uvs, np.float32).reshape(-1)
        offset += p.size
    return positions, normals, uvs


def quad_from_mesh(positions):
    """quadFromMesh (scene.ts:312-324): (position, edge1, edge2) of the
    two-triangle layout (a, b, c), (a, c, d) that quad() produces, else None.
    Float64 arithmetic over the float32 corners, as in JavaScript."""
    p = np.asarray(positions, np.float32).reshape(-1)
    if p.size != 18:
        return None
    v = [tuple(float(x) for x in p[i * 3:i * 3 + 3]) for i in range(6)]
    a, b, c, a2, c2, d = v

    def sub(x, y):
        return (x[0] - y[0], x[1] - y[1], x[2] - y[2])

    def length(x):
        return math.sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2])

    def same(x, y):
        return length(sub(x, y)) < 1e-6

    if not same(a, a2) or not same(c, c2):
        return None
    e1 = sub(b, a)
    e2 = sub(d, a)
    corner = (a[0] + e1[0] + e2[0], a[1] + e1[1] + e2[1], a[2] + e1[2] + e2[2])
    cr = (e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0])
    if not same(corner, c) or length(cr) < 1e-9:
        return None
    return a, e1, e2


def pack_scene(scene, bvh=None):
    """packScene (scene.ts:162-305).  Returns a namespace with the PackedScene
    arrays (tri_geometry, tri_attributes, bvh_nodes, tri_indices, materials,
    spheres, sphere_materials, sphere_lights, lights), their counts,
    light_list, ground_material, color_layers, data_layers and the CPU copies
    (positions, bvh) that pick_distance uses.  bvh: a prebuilt tree over the
    merged meshes (an oracle BVH), else the oracle's buildBVH."""
    import oracle as O
    if len(scene.spheres) > MAX_SPHERES:
        raise ValueError(f"scene has {len(scene.spheres)} spheres; max is {MAX_SPHERES}")
    materials: list = []
    index_of: dict = {}

    def material_index(m) -> int:
        # Deduplicated by object identity (a Map keyed by the Material object).
        k = id(m)
        if k not in index_of:
            index_of[k] = len(materials)
            materials.append(m)
        return index_of[k]

    light_list: list = []

    def add_light(light) -> int:
        if len(light_list) >= MAX_LIGHTS:
            return -1
        light_list.append(light)
        return len(light_list) - 1

    positions, normals, uvs = merge_meshes([inst.mesh for inst in scene.meshes])
    tri_count = positions.size // 9
    tri_material = np.zeros(tri_count, np.int32)
    tri_light = np.full(tri_count, -1, np.int32)
    tri_alpha = np.zeros(tri_count, np.uint8)
    cursor = 0
    for inst in scene.meshes:
        n = np.asarray(inst.mesh.positions).size // 9
        tri_material[cursor:cursor + n] = material_index(inst.material)
        if needs_alpha_test(inst.material):
            tri_alpha[cursor:cursor + n] = 1
        q = quad_from_mesh(inst.mesh.positions) if is_emissive(inst.material) else None
        if q is not None:
            li = add_light({"type": "quad", "position": q[0], "edge1": q[1], "edge2": q[2], "radius": 0.0,
                            "radiance": tuple(inst.material.emission)})
            tri_light[cursor:cursor + n] = li
        cursor += n
    if bvh is not None and np.asarray(bvh.tri_indices).size != tri_count:
        raise ValueError(f"prebuilt BVH covers {np.asarray(bvh.tri_indices).size} triangles; scene has {tri_count}")
    if bvh is None:
        bvh = O.build_bvh(positions)
    src = np.asarray(bvh.tri_indices, np.int64)
    # Leaf order: slot s holds triangle src[s] (scene.ts:215-235).
    geo = np.zeros((tri_count, 12), np.float32)
    attr = np.zeros((tri_count, 16), np.float32)
    P = positions.reshape(-1, 9)
    N = normals.reshape(-1, 9)
    for v in range(3):
        geo[:, v * 4:v * 4 + 3] = P[src, v * 3:v * 3 + 3]
        attr[:, v * 4:v * 4 + 3] = N[src, v * 3:v * 3 + 3]
        if uvs is not None:
            U = uvs.reshape(-1, 6)
            attr[:, v * 4 + 3] = U[src, v * 2]
            attr[:, 12 + v] = U[src, v * 2 + 1]
    geo[:, 3] = tri_material[src]
    geo[:, 7] = tri_light[src]
    geo[:, 11] = tri_alpha[src]

    spheres = np.zeros(MAX_SPHERES * 4, np.float32)
    sphere_materials = np.zeros(MAX_SPHERES, np.int32)
    sphere_lights = np.full(MAX_SPHERES, -1, np.int32)
    for i, s in enumerate(scene.spheres):
        spheres[i * 4:i * 4 + 4] = [s.center[0], s.center[1], s.center[2], s.radius]
        sphere_materials[i] = material_index(s.material)
        if is_emissive(s.material):
            sphere_lights[i] = add_light({"type": "sphere", "position": tuple(s.center), "radius": s.radius,
                                          "edge1": (0.0, 0.0, 0.0), "edge2": (0.0, 0.0, 0.0),
                                          "radiance": tuple(s.material.emission)})
    ground_material = material_index(scene.ground.material)

    lights = np.zeros(MAX_LIGHTS * 16, np.float32)
    for i, l in enumerate(light_list):
        row = [*l["position"], 0.0 if l["type"] == "sphere" else 1.0, *l["edge1"], l["radius"], *l["edge2"], 0.0,
               *l["radiance"], 0.0]
        lights[i * 16:(i + 1) * 16] = row

    # layerMap (scene.ts:136-153): colour (base colour, emissive) and data
    # (metal-rough, normal) arrays, each texture one layer in first-use order.
    color_layers: list = []
    data_layers: list = []

    def layer(table: list, tex) -> float:
        tex = int(tex)
        if tex < 0:
            return -1.0
        if tex not in table:
            table.append(tex)
        return float(table.index(tex))

    maps = []
    for m in materials:
        base = layer(color_layers, m.base_color_map)
        emis = layer(color_layers, m.emissive_map)
        mr = layer(data_layers, m.metal_rough_map)
        nm = layer(data_layers, m.normal_map)
        maps.append((base, mr, nm, emis))
    for t in color_layers + data_layers:
        if t >= len(scene.textures):
            raise ValueError(f"material references texture {t}, scene has {len(scene.textures)}")

    return SimpleNamespace(
        tri_geometry=geo.reshape(-1), tri_attributes=attr.reshape(-1), tri_count=tri_count,
        bvh_nodes=np.asarray(bvh.nodes, np.float32), bvh_node_count=int(bvh.node_count),
        tri_indices=np.asarray(bvh.tri_indices, np.uint32), materials=pack_materials(materials, maps),
        material_count=len(materials), spheres=spheres, sphere_materials=sphere_materials,
        sphere_lights=sphere_lights, sphere_count=len(scene.spheres), lights=lights, light_count=len(light_list),
        light_list=light_list, material_list=materials, ground_material=ground_material,
        ground_material_active=ground_material if scene.ground.enabled else -1,
        ground_height=float(scene.ground.height), color_layers=color_layers, data_layers=data_layers,
        positions=positions, bvh=bvh, sphere_list=list(scene.spheres), ground=scene.ground)


def pick_distance(packed, origin, direction):
    """pickDistance (scene.ts:327-341): triangles (intersectBVH), then
    spheres with tMax = best, then the ground plane; None on a miss."""
    import oracle as O
    o = np.asarray(origin, np.float64).reshape(1, 3)
    d = np.asarray(direction, np.float64).reshape(1, 3)
    best = math.inf
    if packed.tri_count > 0:
        t, tri, _, _ = O.intersect_bvh_batch(packed.bvh, packed.positions, o, d)
        if tri[0] != 0xFFFFFFFF:
            best = float(t[0])
    for s in packed.sphere_list:
        t = O.intersect_sphere(o[0], d[0], s.center, s.radius, 1e-5, best)
        if t is not None:
            best = t
    g = packed.ground
    if g.enabled and abs(d[0, 1]) > 1e-9:
        t = (g.height - o[0, 1]) / d[0, 1]
        if 1e-5 < t < best:
            best = t
    return best if math.isfinite(best) else None


# ---------------------------------------------------------------------------
# Texture arrays (renderer.ts:315-335, context.ts:225-273, resample.frag)
# ---------------------------------------------------------------------------
def _next_pow2(n: int) -> int:
    p = 1
    while p < n:
        p <<= 1
    return p


def _srgb_decode(c8: np.ndarray) -> np.ndarray:
    c = c8 / 255.0
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def _srgb_encode(c: np.ndarray) -> np.ndarray:
    c = np.clip(c, 0.0, 1.0)
    return np.where(c < 0.0031308, c * 12.92, 1.055 * np.power(c, 1 / 2.4) - 0.055)


def _bilinear_clamped(level: np.ndarray, size: int) -> np.ndarray:
    """LINEAR filtering of `level` (h, w, 4) at the centres of a size x size
    target (gl_FragCoord.xy / uSize), CLAMP_TO_EDGE."""
    h, w = level.shape[:2]
    sx = (np.arange(size) + 0.5) / size * w - 0.5
    sy = (np.arange(size) + 0.5) / size * h - 0.5
    x0 = np.floor(sx).astype(np.int64)
    y0 = np.floor(sy).astype(np.int64)
    fx = (sx - x0)[None, :, None]
    fy = (sy - y0)[:, None, None]
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    top = level[ya][:, xa] * (1 - fx) + level[ya][:, xb] * fx
    bot = level[yb][:, xa] * (1 - fx) + level[yb][:, xb] * fx
    return top * (1 - fy) + bot * fy


def resample_layer(img, size: int, srgb: bool) -> np.ndarray:
    """One texture-array layer: the image drawn into a size x size target by
    resample.frag (`texture(uSource, gl_FragCoord.xy / uSize)`) from a
    mipmapped source with LINEAR_MIPMAP_LINEAR / LINEAR filters and
    CLAMP_TO_EDGE (context.ts:256-265).  sRGB layers filter in linear space
    (SRGB8_ALPHA8 source and target).  An image of the layer size maps texel
    centre to texel centre (the identity).  Magnification is bilinear on
    level 0; minification samples the mip chain at lambda =
    log2(max(w, h) / size) (GL 3.0 sec. 3.8.8), the chain built by 2x2 box
    averages (generateMipmap's filter is implementation-defined: that case
    is parity unpinned).  Returns uint8 (size, size, 4)."""
    img = np.asarray(img, np.uint8)
    h, w = img.shape[:2]
    if h == size and w == size:
        return img.copy()
    # Filtering runs on code values (0..255) for unorm channels and on
    # decoded linear values for sRGB colour.
    lin = img.astype(np.float64)
    if srgb:
        lin[..., :3] = _srgb_decode(lin[..., :3])
    lam = math.log2(max(w, h) / size)
    if lam <= 0:
        out = _bilinear_clamped(lin, size)
    else:
        chain = [lin]
        while max(chain[-1].shape[:2]) > 1 and len(chain) <= int(lam) + 1:
            c = chain[-1]
            hh, ww = c.shape[:2]
            nh, nw = max(1, hh // 2), max(1, ww // 2)
            ys = np.minimum(np.arange(nh)[:, None] * 2 + np.array([0, 1])[None, :], hh - 1)
            xs = np.minimum(np.arange(nw)[:, None] * 2 + np.array([0, 1])[None, :], ww - 1)
            acc = np.zeros((nh, nw, 4))
            for dy in range(2):
                for dx in range(2):
                    acc += c[ys[:, dy]][:, xs[:, dx]]
            chain.append(acc / 4)
        lo = min(int(math.floor(lam)), len(chain) - 1)
        hi = min(lo + 1, len(chain) - 1)
        t = lam - math.floor(lam)
        out = _bilinear_clamped(chain[lo], size) * (1 - t) + _bilinear_clamped(chain[hi], size) * t
    if srgb:
        out[..., :3] = _srgb_encode(out[..., :3]) * 255.0
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def texture_layers(scene, packed):
    """buildTextureArrays (renderer.ts:315-335): the shared layer size — the
    next power of two of the largest used image, at most 2048, halved while
    above 256 and the arrays exceed the 256 MiB budget — and the colour and
    data arrays as uint32 texels (R in the low byte).  Returns (size,
    color, data); size 0 when no texture is used."""
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
        stack = np.stack([resample_layer(scene.textures[i], size, srgb) for i in layers])
        return np.ascontiguousarray(stack).view(np.uint32).reshape(-1)

    return size, array(packed.color_layers, True), array(packed.data_layers, False)
