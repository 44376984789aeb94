"""glTF 2.0 / GLB ingestion into this package's meshes and materials — the
host mirror of loadGLTF (reference src/loaders/gltf.ts:1-502).

Meshes come out as triangle soups in world space (node transforms applied,
normals by the inverse-transpose, winding kept under mirroring transforms,
flat normals where NORMAL is missing), one (Mesh, material index) per
primitive; materials follow the metallic-roughness model with the
emissive-strength, transmission, IOR, volume, clearcoat, sheen and unlit
extensions and KHR_texture_transform; images come back as encoded bytes
(the caller decodes them into RGBA8 texture layers, as the reference leaves
decoding to the browser).  Arithmetic follows the reference: accessors are
read into float32, transforms are composed and applied in float64 and the
results stored as float32.
"""
from __future__ import annotations

import base64
import json as _json
import math
import re
import struct
import urllib.parse
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from .scene import Material, Mesh, material

GLB_MAGIC = 0x46546C67
CHUNK_JSON = 0x4E4F534A
CHUNK_BIN = 0x004E4942

# gltf.ts:38-55
UNSUPPORTED_REQUIRED = {"KHR_draco_mesh_compression", "EXT_meshopt_compression", "KHR_texture_basisu",
                        "MSFT_texture_dds", "EXT_mesh_gpu_instancing"}
HANDLED_MATERIAL_EXTENSIONS = {"KHR_materials_emissive_strength", "KHR_materials_transmission", "KHR_materials_ior",
                               "KHR_materials_volume", "KHR_materials_clearcoat", "KHR_materials_sheen",
                               "KHR_materials_unlit"}
COMPONENTS = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
# componentType -> (numpy little-endian dtype, normalisation divisor, signed)
COMPONENT_TYPES = {5120: ("<i1", 127.0, True), 5121: ("<u1", 255.0, False), 5122: ("<i2", 32767.0, True),
                   5123: ("<u2", 65535.0, False), 5125: ("<u4", None, False), 5126: ("<f4", None, False)}


class GLTFError(ValueError):
    pass


@dataclass
class GLTFImage:
    data: bytes
    mime_type: str
    name: str


@dataclass
class GLTFAsset:
    meshes: list            # [(Mesh, material index)], world space
    materials: list         # [Material]; map fields index `images`
    images: list            # [GLTFImage], only those some material samples
    bounds: tuple           # ((min xyz), (max xyz))
    triangle_count: int
    warnings: list = field(default_factory=list)


def is_glb(data: bytes) -> bool:
    return len(data) >= 12 and struct.unpack_from("<I", data, 0)[0] == GLB_MAGIC


def parse_glb(data: bytes):
    """parseGLB (gltf.ts:62-82): (json, bin chunk or None)."""
    if not is_glb(data):
        raise GLTFError("not a GLB file (bad magic)")
    version = struct.unpack_from("<I", data, 4)[0]
    if version != 2:
        raise GLTFError(f"GLB version {version} is not supported; only glTF 2.0")
    length = min(struct.unpack_from("<I", data, 8)[0], len(data))
    offset, js, binary = 12, None, None
    while offset + 8 <= length:
        chunk_length, chunk_type = struct.unpack_from("<II", data, offset)
        start = offset + 8
        if start + chunk_length > length:
            raise GLTFError("GLB chunk runs past the end of the file")
        if chunk_type == CHUNK_JSON:
            js = _json.loads(bytes(data[start:start + chunk_length]).decode("utf-8"))
        elif chunk_type == CHUNK_BIN and binary is None:
            binary = bytes(data[start:start + chunk_length])
        offset = start + chunk_length + ((4 - chunk_length % 4) % 4)
    if not js:
        raise GLTFError("GLB has no JSON chunk")
    return js, binary


def decode_data_uri(uri: str) -> bytes:
    """decodeDataUri (gltf.ts:84-94)."""
    comma = uri.find(",")
    if comma < 0:
        raise GLTFError("malformed data URI")
    meta, payload = uri[5:comma], uri[comma + 1:]
    if not meta.endswith(";base64"):
        return urllib.parse.unquote(payload).encode("utf-8")
    return base64.b64decode(payload)


def _mime_from_data_uri(uri: str) -> str:
    m = re.match(r"^data:([^;,]+)", uri)
    return m.group(1) if m else ""


def _mime_from_name(name: str) -> str:
    ext = re.split(r"[?#]", name)[0].split(".")[-1].lower()
    return {"png": "image/png", "jpg": "image/jpeg", "jpeg": "image/jpeg", "webp": "image/webp",
            "ktx2": "image/ktx2"}.get(ext, "application/octet-stream")


def compose_trs(t, q, s) -> np.ndarray:
    """Column-major T · R · S (gltf.ts:140-158) as 16 float64."""
    x, y, z, w = q
    m = np.zeros(16)
    m[0] = (1 - 2 * (y * y + z * z)) * s[0]
    m[1] = 2 * (x * y + z * w) * s[0]
    m[2] = 2 * (x * z - y * w) * s[0]
    m[4] = 2 * (x * y - z * w) * s[1]
    m[5] = (1 - 2 * (x * x + z * z)) * s[1]
    m[6] = 2 * (y * z + x * w) * s[1]
    m[8] = 2 * (x * z + y * w) * s[2]
    m[9] = 2 * (y * z - x * w) * s[2]
    m[10] = (1 - 2 * (x * x + y * y)) * s[2]
    m[12], m[13], m[14], m[15] = t[0], t[1], t[2], 1.0
    return m


def _multiply(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major a · b, summed in k order as gltf.ts:129-138."""
    out = np.zeros(16)
    for c in range(4):
        for r in range(4):
            s = 0.0
            for k in range(4):
                s += a[k * 4 + r] * b[c * 4 + k]
            out[c * 4 + r] = s
    return out


def _normal_matrix(m: np.ndarray):
    """Inverse-transpose of the upper 3x3 (row-major) and its determinant (gltf.ts:161-175)."""
    a, b, c = m[0], m[4], m[8]
    d, e, f = m[1], m[5], m[9]
    g, h, i = m[2], m[6], m[10]
    A, B, C = e * i - f * h, -(d * i - f * g), d * h - e * g
    det = a * A + b * B + c * C
    inv = 0.0 if det == 0 else 1.0 / det
    n = [A * inv, B * inv, C * inv,
         -(b * i - c * h) * inv, (a * i - c * g) * inv, -(a * h - b * g) * inv,
         (b * f - c * e) * inv, -(a * f - c * d) * inv, (a * e - b * d) * inv]
    return n, det


def texture_transform(ext) -> tuple:
    """KHR_texture_transform as a row-major 2x3 matrix T · R · S (gltf.ts:178-186)."""
    if not ext:
        return (1.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    ox, oy = ext.get("offset", [0, 0])
    sx, sy = ext.get("scale", [1, 1])
    r = ext.get("rotation", 0)
    c, s = math.cos(r), math.sin(r)
    return (c * sx, -s * sy, ox, s * sx, c * sy, oy)


def load_gltf(source: bytes, resolve: Callable[[str], bytes]) -> GLTFAsset:
    """loadGLTF (gltf.ts:188-502): a .glb or .gltf file's bytes; `resolve`
    fetches external resources by relative URI (buffers, images that a
    material uses)."""
    source = bytes(source)
    glb_bin = None
    if is_glb(source):
        js, glb_bin = parse_glb(source)
    else:
        js = _json.loads(source.decode("utf-8"))
    version = str(((js or {}).get("asset") or {}).get("version", "") or "")
    if not version.startswith("2"):
        raise GLTFError(f'glTF version "{version or "unknown"}" is not supported; only 2.0')
    warnings: list = []

    def warn(msg: str) -> None:
        if msg not in warnings:
            warnings.append(msg)

    for ext in js.get("extensionsRequired", []) or []:
        if ext in UNSUPPORTED_REQUIRED:
            raise GLTFError(f"this file requires {ext}, which Lucent does not support")

    buffers = []
    for i, b in enumerate(js.get("buffers", []) or []):
        if "uri" not in b:
            if i == 0 and glb_bin is not None:
                buffers.append(glb_bin)
                continue
            raise GLTFError(f"buffer {i} has no uri and there is no GLB binary chunk")
        uri = str(b["uri"])
        buffers.append(decode_data_uri(uri) if uri.startswith("data:") else bytes(resolve(uri)))

    def buffer_view(index):
        views = js.get("bufferViews") or []
        view = views[index] if isinstance(index, int) and 0 <= index < len(views) else None
        if not view:
            raise GLTFError(f"missing bufferView {index}")
        bi = view.get("buffer")
        buf = buffers[bi] if isinstance(bi, int) and 0 <= bi < len(buffers) else None
        if buf is None:
            raise GLTFError(f"bufferView {index} references missing buffer {bi}")
        start = view.get("byteOffset", 0)
        if start + view["byteLength"] > len(buf):
            raise GLTFError(f"bufferView {index} runs past its buffer")
        return buf[start:start + view["byteLength"]], view.get("byteStride", 0)

    def read_components(data: bytes, comp_type: int, normalized: bool, offsets: np.ndarray) -> np.ndarray:
        dt, div, signed = COMPONENT_TYPES[comp_type]
        raw = np.frombuffer(data, np.uint8)
        size = np.dtype(dt).itemsize
        idx = offsets[..., None] + np.arange(size)
        vals = raw[idx].copy().view(dt)[..., 0] if size > 1 else raw[idx][..., 0].view(dt)
        v = vals.astype(np.float64)
        if normalized and div is not None:
            v = v / div
            if signed:
                v = np.maximum(v, -1.0)
        return v.astype(np.float32)

    accessor_cache: dict = {}

    def read_accessor(index):
        if index in accessor_cache:
            return accessor_cache[index]
        accessors = js.get("accessors") or []
        acc = accessors[index] if isinstance(index, int) and 0 <= index < len(accessors) else None
        if not acc:
            raise GLTFError(f"missing accessor {index}")
        size = COMPONENTS.get(acc.get("type"))
        ct = acc.get("componentType")
        if not size or ct not in COMPONENT_TYPES:
            raise GLTFError(f"accessor {index} has unsupported type {acc.get('type')}/{ct}")
        comp_bytes = np.dtype(COMPONENT_TYPES[ct][0]).itemsize
        count = int(acc["count"])
        out = np.zeros((count, size), np.float32)
        if "bufferView" in acc:
            data, stride = buffer_view(acc["bufferView"])
            element = size * comp_bytes
            step = stride or element
            base = acc.get("byteOffset", 0)
            if count > 0 and base + step * (count - 1) + element > len(data):
                raise GLTFError(f"accessor {index} reads past its bufferView")
            offs = base + np.arange(count)[:, None] * step + np.arange(size)[None, :] * comp_bytes
            if count:
                out[:] = read_components(data, ct, acc.get("normalized") is True, offs)
        sparse = acc.get("sparse")
        if sparse:
            n = sparse["count"]
            ind, val = sparse["indices"], sparse["values"]
            idata, _ = buffer_view(ind["bufferView"])
            vdata, _ = buffer_view(val["bufferView"])
            ib = np.dtype(COMPONENT_TYPES[ind["componentType"]][0]).itemsize
            targets = read_components(idata, ind["componentType"], False,
                                      ind.get("byteOffset", 0) + np.arange(n) * ib)
            voffs = val.get("byteOffset", 0) + (np.arange(n)[:, None] * size + np.arange(size)[None, :]) * comp_bytes
            values = read_components(vdata, ct, acc.get("normalized") is True, voffs)
            for k in range(n):
                t = int(targets[k])
                if t >= count:
                    raise GLTFError(f"sparse index {t} out of range in accessor {index}")
                out[t] = values[k]
        res = (out, size, count)
        accessor_cache[index] = res
        return res

    images_json = js.get("images") or []
    image_remap: dict = {}
    image_list: list = []

    def use_image(image_index) -> int:
        if image_index in image_remap:
            return image_remap[image_index]
        img = images_json[image_index] if isinstance(image_index, int) and 0 <= image_index < len(images_json) else None
        if not img:
            return -1
        slot = len(image_list)
        image_remap[image_index] = slot
        name = str(img.get("name", img.get("uri", f"image {image_index}")))
        if "bufferView" in img:
            data, _ = buffer_view(img["bufferView"])
            image_list.append(GLTFImage(bytes(data), img.get("mimeType", "image/png"), name))
        else:
            uri = str(img.get("uri"))
            if uri.startswith("data:"):
                image_list.append(GLTFImage(decode_data_uri(uri), _mime_from_data_uri(uri), name))
            else:
                image_list.append(GLTFImage(bytes(resolve(uri)), img.get("mimeType", _mime_from_name(uri)), name))
        return slot

    def texture_slot(info) -> int:
        if not info or info.get("index") is None:
            return -1
        textures = js.get("textures") or []
        ti = info["index"]
        tex = textures[ti] if isinstance(ti, int) and 0 <= ti < len(textures) else None
        if not tex:
            return -1
        if info.get("texCoord", 0) != 0:
            warn("textures on TEXCOORD_1 or higher are sampled with TEXCOORD_0")
        ext = tex.get("extensions") or {}
        src = (ext.get("EXT_texture_webp") or {}).get("source", tex.get("source"))
        if src is None:
            if ext.get("KHR_texture_basisu"):
                warn("KTX2/Basis textures are not supported and were skipped")
            return -1
        img = images_json[src] if isinstance(src, int) and 0 <= src < len(images_json) else {}
        mime = img.get("mimeType", _mime_from_name(str(img.get("uri", ""))))
        if mime == "image/ktx2":
            warn("KTX2/Basis textures are not supported and were skipped")
            return -1
        return use_image(src)

    def convert_material(m) -> Material:
        pbr = m.get("pbrMetallicRoughness") or {}
        ext = m.get("extensions") or {}
        for name in ext:
            if name not in HANDLED_MATERIAL_EXTENSIONS:
                warn(f"material extension {name} is ignored")
        bc = pbr.get("baseColorFactor", [1, 1, 1, 1])
        strength = (ext.get("KHR_materials_emissive_strength") or {}).get("emissiveStrength", 1)
        ef = m.get("emissiveFactor", [0, 0, 0])
        transform_src = next((t for t in (pbr.get("baseColorTexture"), pbr.get("metallicRoughnessTexture"),
                                          m.get("normalTexture"), m.get("emissiveTexture"))
                              if t and (t.get("extensions") or {}).get("KHR_texture_transform")), None)
        volume = ext.get("KHR_materials_volume")
        mode = m.get("alphaMode")
        alpha_mode = "mask" if mode == "MASK" else "blend" if mode == "BLEND" else "opaque"
        normal_tex = m.get("normalTexture")
        out = material(
            albedo=(bc[0], bc[1], bc[2]), alpha=bc[3] if len(bc) > 3 else 1, alpha_mode=alpha_mode,
            alpha_cutoff=m.get("alphaCutoff", 0.5), metallic=pbr.get("metallicFactor", 1),
            roughness=pbr.get("roughnessFactor", 1),
            base_color_map=texture_slot(pbr.get("baseColorTexture")),
            metal_rough_map=texture_slot(pbr.get("metallicRoughnessTexture")),
            normal_map=texture_slot(normal_tex),
            normal_scale=(normal_tex or {}).get("scale", 1),
            emissive_map=texture_slot(m.get("emissiveTexture")),
            emission=(ef[0] * strength, ef[1] * strength, ef[2] * strength),
            transmission=(ext.get("KHR_materials_transmission") or {}).get("transmissionFactor", 0),
            ior=(ext.get("KHR_materials_ior") or {}).get("ior", 1.5),
            thin_walled=not volume or volume.get("thicknessFactor", 0) == 0,
            attenuation_color=tuple((volume or {}).get("attenuationColor", [1, 1, 1])),
            attenuation_distance=(volume or {}).get("attenuationDistance", math.inf),
            clearcoat=(ext.get("KHR_materials_clearcoat") or {}).get("clearcoatFactor", 0),
            clearcoat_roughness=(ext.get("KHR_materials_clearcoat") or {}).get("clearcoatRoughnessFactor", 0),
            sheen_color=tuple((ext.get("KHR_materials_sheen") or {}).get("sheenColorFactor", [0, 0, 0])),
            sheen_roughness=(ext.get("KHR_materials_sheen") or {}).get("sheenRoughnessFactor", 0),
            uv_transform=texture_transform(((transform_src or {}).get("extensions") or {}).get(
                "KHR_texture_transform")),
        )
        if "KHR_materials_unlit" in ext:
            from dataclasses import replace
            return replace(out, emission=out.albedo, emissive_map=out.base_color_map, albedo=(0.0, 0.0, 0.0),
                           metallic=0.0)
        return out

    materials = [convert_material(m) for m in js.get("materials", []) or []]
    default_material = [-1]

    def material_for(index) -> int:
        if isinstance(index, int) and 0 <= index < len(materials):
            return index
        if default_material[0] < 0:
            # The glTF default material: white, fully metallic, fully rough.
            default_material[0] = len(materials)
            materials.append(material(albedo=(1.0, 1.0, 1.0), metallic=1.0, roughness=1.0))
        return default_material[0]

    meshes: list = []
    lo = [math.inf] * 3
    hi = [-math.inf] * 3
    tri_total = [0]

    def emit_primitive(prim, world: np.ndarray) -> None:
        mode = prim.get("mode", 4)
        if mode not in (4, 5, 6):
            warn("point and line primitives are skipped")
            return
        if (prim.get("extensions") or {}).get("KHR_draco_mesh_compression"):
            raise GLTFError("Draco-compressed meshes are not supported")
        attrs = prim.get("attributes") or {}
        if attrs.get("POSITION") is None:
            return
        pos, _, vertex_count = read_accessor(attrs["POSITION"])
        nrm = read_accessor(attrs["NORMAL"])[0] if attrs.get("NORMAL") is not None else None
        uv = read_accessor(attrs["TEXCOORD_0"])[0] if attrs.get("TEXCOORD_0") is not None else None
        # Indices pass through float32 as the reference's accessor arrays do.
        idx = read_accessor(prim["indices"])[0].reshape(-1) if prim.get("indices") is not None else None
        n = idx.size if idx is not None else vertex_count
        vertex = (idx.astype(np.int64) if idx is not None else np.arange(n, dtype=np.int64))
        if mode == 4:
            k = np.arange(0, n - 2, 3)
            tris = np.stack([vertex[k], vertex[k + 1], vertex[k + 2]], -1)
        elif mode == 5:
            k = np.arange(0, max(n - 2, 0))
            even = (k % 2) == 0
            tris = np.stack([vertex[k], np.where(even, vertex[k + 1], vertex[k + 2]),
                             np.where(even, vertex[k + 2], vertex[k + 1])], -1)
        else:
            k = np.arange(1, max(n - 1, 1))
            tris = np.stack([vertex[k], vertex[k + 1], np.full_like(k, vertex[0] if n else 0)], -1)
        tris = tris.reshape(-1, 3)
        if tris.shape[0] == 0:
            return
        nm, det = _normal_matrix(world)
        if det < 0:
            # Mirroring transforms reverse winding; swap two corners to keep normals and winding agreeing.
            tris = tris[:, [0, 2, 1]]
        bad = tris >= vertex_count
        if bad.any():
            raise GLTFError(f"index {int(tris[bad][0])} out of range ({vertex_count} vertices)")
        p = pos.astype(np.float64)[tris]  # (T, 3, 3)
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        wx = world[0] * x + world[4] * y + world[8] * z + world[12]
        wy = world[1] * x + world[5] * y + world[9] * z + world[13]
        wz = world[2] * x + world[6] * y + world[10] * z + world[14]
        positions = np.stack([wx, wy, wz], -1).astype(np.float32)
        w3 = np.stack([wx, wy, wz], -1).reshape(-1, 3)
        for a in range(3):
            lo[a] = min(lo[a], float(w3[:, a].min()))
            hi[a] = max(hi[a], float(w3[:, a].max()))
        if nrm is not None:
            q = nrm.astype(np.float64)[tris]
            a_, b_, d_ = q[..., 0], q[..., 1], q[..., 2]
            nx = nm[0] * a_ + nm[1] * b_ + nm[2] * d_
            ny = nm[3] * a_ + nm[4] * b_ + nm[5] * d_
            nz = nm[6] * a_ + nm[7] * b_ + nm[8] * d_
            ln = np.sqrt(nx * nx + ny * ny + nz * nz)
            ln = np.where(ln == 0, 1.0, ln)
            normals = np.stack([nx / ln, ny / ln, nz / ln], -1).astype(np.float32)
        else:
            # glTF: missing normals mean flat shading, from the stored float32 positions.
            pf = positions.astype(np.float64)
            e1, e2 = pf[:, 1] - pf[:, 0], pf[:, 2] - pf[:, 0]
            c = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                          e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], -1)
            ln = np.sqrt((c * c).sum(-1, keepdims=True))
            c = c / np.where(ln == 0, 1.0, ln)
            normals = np.repeat(c[:, None, :], 3, axis=1).astype(np.float32)
        uvs = uv[tris][..., :2].astype(np.float32).reshape(-1) if uv is not None else None
        tri_total[0] += tris.shape[0]
        meshes.append((Mesh(positions.reshape(-1), normals.reshape(-1), uvs), material_for(prim.get("material"))))

    nodes = js.get("nodes") or []

    def visit(node_index, parent: np.ndarray, depth: int) -> None:
        if depth > 64:
            raise GLTFError("node hierarchy is deeper than 64 levels (cycle?)")
        node = nodes[node_index] if isinstance(node_index, int) and 0 <= node_index < len(nodes) else None
        if node is None:
            raise GLTFError(f"missing node {node_index}")
        if node.get("matrix"):
            local = np.asarray(node["matrix"], np.float64)
        else:
            local = compose_trs(node.get("translation", [0, 0, 0]), node.get("rotation", [0, 0, 0, 1]),
                                node.get("scale", [1, 1, 1]))
        world = _multiply(parent, local)
        if node.get("mesh") is not None:
            ms = js.get("meshes") or []
            mi = node["mesh"]
            mesh = ms[mi] if isinstance(mi, int) and 0 <= mi < len(ms) else None
            if not mesh:
                raise GLTFError(f"node {node_index} references missing mesh {mi}")
            for prim in mesh.get("primitives", []) or []:
                emit_primitive(prim, world)
        for child in node.get("children", []) or []:
            visit(child, world, depth + 1)

    scenes = js.get("scenes") or []
    si = js.get("scene", 0)
    if isinstance(si, int) and 0 <= si < len(scenes) and scenes[si]:
        roots = scenes[si].get("nodes", []) or []
    else:
        children = {c for nd in nodes for c in (nd.get("children") or [])}
        roots = [i for i in range(len(nodes)) if i not in children]
    identity = np.zeros(16)
    identity[[0, 5, 10, 15]] = 1.0
    for root in roots:
        visit(root, identity, 0)
    if tri_total[0] == 0:
        raise GLTFError("the file contains no triangle meshes")
    return GLTFAsset(meshes, materials, image_list, (tuple(lo), tuple(hi)), tri_total[0], warnings)
