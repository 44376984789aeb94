"""glTF 2.0 / GLB ingestion through the C ABI (lucent_gltf_*, the native
loader in csrc/gltf.cpp) — the behaviour of loadGLTF (reference
src/loaders/gltf.ts:188-502) for a Python caller.

``load_gltf(source, resolve)`` parses the file natively, asks ``resolve``
for each external file the native parser lists (buffers, then the images
some material samples, in first-use order), hands the bytes back and
collects the result: one (Mesh, material index) per triangle primitive in
world space, Materials, the sampled images still encoded, bounds, the
triangle count and the warnings.  Failures raise GLTFError with the native
loader's message.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import struct
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _native
from .scene import ALPHA_CODE, Material, Mesh

_ALPHA_MODES = {code: mode for mode, code in ALPHA_CODE.items()}


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


_CHUNK_JSON, _CHUNK_BIN = struct.unpack("<II", b"JSONBIN\0")


def parse_glb(data: bytes):
    """The JSON and first binary chunk of a GLB container (the layout
    parseGLB reads, gltf.ts:62-82): (dict, bytes or None)."""
    view = memoryview(bytes(data))
    if len(view) < 12 or view[:4].tobytes() != b"glTF":
        raise GLTFError("not a GLB file (bad magic)")
    _, version, total = struct.unpack_from("<4sII", view)
    if version != 2:
        raise GLTFError(f"GLB version {version} is not supported; only glTF 2.0")
    chunks, at, end = {}, 12, min(total, len(view))
    while at + 8 <= end:
        size, kind = struct.unpack_from("<II", view, at)
        if at + 8 + size > end:
            raise GLTFError("GLB chunk runs past the end of the file")
        chunks.setdefault(kind, view[at + 8:at + 8 + size].tobytes())
        at += 8 + size + (-size) % 4
    if _CHUNK_JSON not in chunks:
        raise GLTFError("GLB has no JSON chunk")
    return json.loads(chunks[_CHUNK_JSON]), chunks.get(_CHUNK_BIN)


def texture_transform(ext: Optional[dict]) -> tuple:
    """KHR_texture_transform as the row-major 2x3 matrix T * R * S that the
    material's uv_transform holds (gltf.ts:178-186)."""
    if not ext:
        return (1.0, 0.0, 0.0, 0.0, 1.0, 0.0)
    (ox, oy), (sx, sy) = ext.get("offset", (0, 0)), ext.get("scale", (1, 1))
    angle = ext.get("rotation", 0)
    c, s = math.cos(angle), math.sin(angle)
    return (c * sx, -s * sy, ox, s * sx, c * sy, oy)


def _material(t) -> Material:
    """A Material from its 40 texel values (pack_materials' order)."""
    return Material(
        albedo=(t[0], t[1], t[2]), metallic=t[3], roughness=t[4], ior=t[5], transmission=t[6], checker_scale=t[7],
        emission=(t[8], t[9], t[10]), alpha=t[11], albedo2=(t[12], t[13], t[14]),
        alpha_mode=_ALPHA_MODES[int(t[15])], base_color_map=int(t[16]), metal_rough_map=int(t[17]),
        normal_map=int(t[18]), emissive_map=int(t[19]), normal_scale=t[20], clearcoat=t[21],
        clearcoat_roughness=t[22], alpha_cutoff=t[23], sheen_color=(t[24], t[25], t[26]), sheen_roughness=t[27],
        attenuation_color=(t[28], t[29], t[30]), attenuation_distance=t[31] if t[31] > 0 else math.inf,
        uv_transform=(t[32], t[33], t[34], t[36], t[37], t[38]), thin_walled=t[35] != 0)


def _check(status: int) -> None:
    try:
        _native.check(status)
    except _native.LucentError as e:
        msg = _native.lib().lucent_last_error(None)
        raise GLTFError(msg.decode() if msg else str(e)) from e


def load_gltf(source: bytes, resolve: Callable[[str], bytes]) -> GLTFAsset:
    """A .glb or .gltf file's bytes -> GLTFAsset; ``resolve`` returns the
    bytes of an external file by its URI as written in the JSON."""
    lib = _native.lib()
    src = bytes(source)
    doc = C.c_void_p()
    _check(lib.lucent_gltf_parse(src, len(src), C.byref(doc)))
    try:
        for i in range(lib.lucent_gltf_resource_count(doc)):
            data = bytes(resolve(lib.lucent_gltf_resource_uri(doc, i).decode()))
            _check(lib.lucent_gltf_set_resource(doc, i, data, len(data)))
        _check(lib.lucent_gltf_build(doc))
        info = _native.GltfSummaryStruct()
        _check(lib.lucent_gltf_info(doc, C.byref(info)))
        meshes = []
        for i in range(info.mesh_count):
            n, mat = C.c_uint32(), C.c_int32()
            ptrs = [C.POINTER(C.c_float)() for _ in range(3)]
            _check(lib.lucent_gltf_mesh(doc, i, C.byref(n), C.byref(mat), *(C.byref(p) for p in ptrs)))
            pos, nrm, uv = (np.ctypeslib.as_array(p, (n.value * k,)).copy() if p else None
                            for p, k in zip(ptrs, (9, 9, 6)))
            meshes.append((Mesh(pos, nrm, uv), mat.value))
        materials = []
        for i in range(info.material_count):
            texels = (C.c_double * 40)()
            _check(lib.lucent_gltf_material(doc, i, texels))
            materials.append(_material(list(texels)))
        images = []
        for i in range(info.image_count):
            data, size = C.POINTER(C.c_uint8)(), C.c_size_t()
            mime, name = C.c_char_p(), C.c_char_p()
            _check(lib.lucent_gltf_image(doc, i, C.byref(data), C.byref(size), C.byref(mime), C.byref(name)))
            images.append(GLTFImage(C.string_at(data, size.value) if size.value else b"", mime.value.decode(),
                                    name.value.decode()))
        warnings = [lib.lucent_gltf_warning(doc, i).decode() for i in range(info.warning_count)]
        b = list(info.bounds)
        return GLTFAsset(meshes, materials, images, (tuple(b[:3]), tuple(b[3:])), int(info.triangle_count), warnings)
    finally:
        lib.lucent_gltf_destroy(doc)
