"""glTF / GLB ingestion (lucent_amd.gltf) against the reference loader's
test cases (test/gltf.test.ts:1-285): GLB chunking, an indexed textured
triangle, TRS and matrix hierarchies, mirroring transforms, flat normals and
the default material, interleaved / normalized / strip data, sparse
accessors, data-URI buffers with lazily fetched images, the PBR extensions,
refusals, and KHR_texture_transform."""
import base64
import json
import math
import struct

import numpy as np
import pytest

from lucent_amd.gltf import GLTFError, load_gltf, parse_glb, texture_transform

FLOAT = 5126


def make_glb(js, binary: bytes | None = None) -> bytes:
    jb = json.dumps(js).encode()
    jb += b" " * ((4 - len(jb) % 4) % 4)
    out = struct.pack("<III", 0x46546C67, 2, 0)
    out += struct.pack("<II", len(jb), 0x4E4F534A) + jb
    if binary is not None:
        pad = (4 - len(binary) % 4) % 4
        out += struct.pack("<II", len(binary) + pad, 0x004E4942) + binary + b"\0" * pad
    return out[:8] + struct.pack("<I", len(out)) + out[12:]


def pack(*parts):
    offsets, blob = [], b""
    for p in parts:
        b = np.ascontiguousarray(p).tobytes()
        offsets.append(len(blob))
        blob += b + b"\0" * ((4 - len(b) % 4) % 4)
    return blob, offsets


def no_fetch(uri):
    raise AssertionError(f"unexpected fetch of {uri}")


def triangle_glb(node=None, extra=None):
    positions = np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0])
    normals = np.float32([0, 0, 1, 0, 0, 1, 0, 0, 1])
    uvs = np.float32([0, 0, 1, 0, 0, 1])
    indices = np.uint16([0, 1, 2])
    png = np.uint8([0x89, 0x50, 0x4E, 0x47])
    blob, off = pack(positions, normals, uvs, indices, png)
    js = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0]}],
        "nodes": [{"mesh": 0, **(node or {})}],
        "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}, "indices": 3,
                                    "material": 0}]}],
        "materials": [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0},
                                                "baseColorFactor": [1, 0.5, 0.25, 1]}}],
        "textures": [{"source": 0}],
        "images": [{"bufferView": 4, "mimeType": "image/png"}],
        "accessors": [
            {"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3"},
            {"bufferView": 1, "componentType": FLOAT, "count": 3, "type": "VEC3"},
            {"bufferView": 2, "componentType": FLOAT, "count": 3, "type": "VEC2"},
            {"bufferView": 3, "componentType": 5123, "count": 3, "type": "SCALAR"},
        ],
        "bufferViews": [
            {"buffer": 0, "byteOffset": off[0], "byteLength": 36},
            {"buffer": 0, "byteOffset": off[1], "byteLength": 36},
            {"buffer": 0, "byteOffset": off[2], "byteLength": 24},
            {"buffer": 0, "byteOffset": off[3], "byteLength": 6},
            {"buffer": 0, "byteOffset": off[4], "byteLength": 4},
        ],
        "buffers": [{"byteLength": len(blob)}],
    }
    js.update(extra or {})
    return make_glb(js, blob)


def vertex(p, i):
    return [round(float(v), 5) + 0.0 for v in p[i * 3:i * 3 + 3]]


def test_parse_glb_splits_chunks_and_rejects_other_files():
    js, binary = parse_glb(make_glb({"asset": {"version": "2.0"}}, bytes([1, 2, 3])))
    assert js["asset"]["version"] == "2.0" and list(binary[:3]) == [1, 2, 3]
    with pytest.raises(GLTFError, match="magic"):
        parse_glb(b'{"asset":{}}')


def test_indexed_textured_triangle():
    a = load_gltf(triangle_glb(), no_fetch)
    assert a.triangle_count == 1
    mesh, mi = a.meshes[0]
    assert vertex(mesh.positions, 1) == [1, 0, 0]
    assert vertex(mesh.normals, 0) == [0, 0, 1]
    assert mesh.uvs.tolist() == [0, 0, 1, 0, 0, 1]
    m = a.materials[mi]
    assert m.albedo == (1, 0.5, 0.25) and m.base_color_map == 0
    assert len(a.images) == 1 and a.images[0].mime_type == "image/png"
    assert list(a.images[0].data) == [0x89, 0x50, 0x4E, 0x47]
    assert a.bounds == ((0, 0, 0), (1, 1, 0))


def test_trs_rotates_normals_with_the_geometry():
    s = math.sqrt(0.5)  # 90 degrees about +Y maps +X to -Z and +Z to +X
    a = load_gltf(triangle_glb({"translation": [1, 2, 3], "rotation": [0, s, 0, s], "scale": [2, 2, 2]}), no_fetch)
    mesh = a.meshes[0][0]
    assert vertex(mesh.positions, 0) == [1, 2, 3]
    assert vertex(mesh.positions, 1) == [1, 2, 1]
    assert vertex(mesh.positions, 2) == [1, 4, 3]
    assert vertex(mesh.normals, 0) == [1, 0, 0]


def test_hierarchy_accumulates():
    glb = triangle_glb({}, {"scenes": [{"nodes": [1]}],
                            "nodes": [{"mesh": 0, "translation": [0, 0, 5]},
                                      {"children": [0], "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 10, 0, 0, 1]}]})
    assert vertex(load_gltf(glb, no_fetch).meshes[0][0].positions, 0) == [10, 0, 5]


def test_mirroring_keeps_winding_and_normals_consistent():
    mesh = load_gltf(triangle_glb({"scale": [-1, 1, 1]}), no_fetch).meshes[0][0]
    p, n = mesh.positions.astype(np.float64), mesh.normals
    e1, e2 = p[3:6] - p[0:3], p[6:9] - p[0:3]
    gz = e1[0] * e2[1] - e1[1] * e2[0]
    assert np.sign(gz) == np.sign(n[2]) and abs(n[2] - 1) < 1e-6


def test_flat_normals_and_default_material():
    blob, _ = pack(np.float32([0, 0, 0, 0, 0, 1, 1, 0, 0]))
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
          "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3"}],
          "bufferViews": [{"buffer": 0, "byteLength": 36}], "buffers": [{"byteLength": len(blob)}]}
    a = load_gltf(make_glb(js, blob), no_fetch)
    assert vertex(a.meshes[0][0].normals, 2) == [0, 1, 0]
    m = a.materials[a.meshes[0][1]]
    assert (m.albedo, m.metallic, m.roughness) == ((1, 1, 1), 1, 1)


def test_interleaved_normalized_strip():
    stride = 16
    buf = bytearray(stride * 4)
    for i, v in enumerate([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]):
        struct.pack_into("<fff", buf, i * stride, *v)
        struct.pack_into("<HH", buf, i * stride + 12, v[0] * 65535, 32768)
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
          "meshes": [{"primitives": [{"mode": 5, "attributes": {"POSITION": 0, "TEXCOORD_0": 1}}]}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 4, "type": "VEC3"},
                        {"bufferView": 0, "byteOffset": 12, "componentType": 5123, "normalized": True, "count": 4,
                         "type": "VEC2"}],
          "bufferViews": [{"buffer": 0, "byteLength": len(buf), "byteStride": stride}],
          "buffers": [{"byteLength": len(buf)}]}
    a = load_gltf(make_glb(js, bytes(buf)), no_fetch)
    mesh = a.meshes[0][0]
    assert a.triangle_count == 2
    assert abs(mesh.normals[2] - 1) < 1e-6 and abs(mesh.normals[11] - 1) < 1e-6  # strip triangles face +Z
    assert vertex(mesh.positions, 4) == [1, 1, 0]
    assert abs(mesh.uvs[2] - 1) < 1e-6 and mesh.uvs[3] == np.float32(32768 / 65535)


def test_sparse_accessor_overrides():
    blob, off = pack(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]), np.uint16([2]), np.float32([0, 7, 0]))
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
          "meshes": [{"primitives": [{"attributes": {"POSITION": 0}}]}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3",
                         "sparse": {"count": 1, "indices": {"bufferView": 1, "componentType": 5123},
                                    "values": {"bufferView": 2}}}],
          "bufferViews": [{"buffer": 0, "byteOffset": off[0], "byteLength": 36},
                          {"buffer": 0, "byteOffset": off[1], "byteLength": 2},
                          {"buffer": 0, "byteOffset": off[2], "byteLength": 12}],
          "buffers": [{"byteLength": len(blob)}]}
    assert vertex(load_gltf(make_glb(js, blob), no_fetch).meshes[0][0].positions, 2) == [0, 7, 0]


def test_gltf_json_with_data_uri_fetches_only_used_images():
    blob, off = pack(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]), np.float32([0, 0, 1, 0, 0, 1]))
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
          "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "material": 0}]}],
          "materials": [{"normalTexture": {"index": 0, "scale": 0.5}}],
          "textures": [{"source": 1}],
          "images": [{"uri": "unused.png"}, {"uri": "tex%20normal.jpg"}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3"},
                        {"bufferView": 1, "componentType": FLOAT, "count": 3, "type": "VEC2"}],
          "bufferViews": [{"buffer": 0, "byteOffset": off[0], "byteLength": 36},
                          {"buffer": 0, "byteOffset": off[1], "byteLength": 24}],
          "buffers": [{"byteLength": len(blob),
                       "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}]}
    fetched = []

    def fetch(uri):
        fetched.append(uri)
        return bytes([9])

    a = load_gltf(json.dumps(js).encode(), fetch)
    assert fetched == ["tex%20normal.jpg"]
    assert a.images[0].mime_type == "image/jpeg"
    m = a.materials[0]
    assert (m.normal_map, m.normal_scale, m.base_color_map) == (0, 0.5, -1)


def test_pbr_extensions_map_onto_the_material_model():
    glb = triangle_glb({}, {"materials": [{
        "alphaMode": "MASK", "alphaCutoff": 0.3, "emissiveFactor": [1, 0.5, 0],
        "pbrMetallicRoughness": {"metallicFactor": 0.25, "roughnessFactor": 0.75},
        "extensions": {
            "KHR_materials_emissive_strength": {"emissiveStrength": 4},
            "KHR_materials_transmission": {"transmissionFactor": 0.9},
            "KHR_materials_ior": {"ior": 1.33},
            "KHR_materials_volume": {"thicknessFactor": 0.1, "attenuationDistance": 2,
                                     "attenuationColor": [1, 0.5, 0.5]},
            "KHR_materials_clearcoat": {"clearcoatFactor": 1, "clearcoatRoughnessFactor": 0.1},
            "KHR_materials_sheen": {"sheenColorFactor": [0.2, 0.3, 0.4], "sheenRoughnessFactor": 0.5},
            "KHR_materials_iridescence": {"iridescenceFactor": 1},
        }}]})
    a = load_gltf(glb, no_fetch)
    m = a.materials[0]
    assert m.alpha_mode == "mask" and m.alpha_cutoff == 0.3
    assert m.emission == (4, 2, 0)
    assert (m.metallic, m.roughness, m.transmission, m.ior) == (0.25, 0.75, 0.9, 1.33)
    assert (m.thin_walled, m.attenuation_distance, m.attenuation_color) == (False, 2, (1, 0.5, 0.5))
    assert (m.clearcoat, m.clearcoat_roughness, m.sheen_roughness) == (1, 0.1, 0.5)
    assert m.sheen_color == (0.2, 0.3, 0.4)
    assert any("KHR_materials_iridescence" in w for w in a.warnings)
    assert len(a.images) == 0  # no texture referenced, so the embedded image is not returned


def test_transmission_without_volume_is_thin_walled():
    a = load_gltf(triangle_glb({}, {"materials": [{"extensions": {
        "KHR_materials_transmission": {"transmissionFactor": 1}}}]}), no_fetch)
    assert a.materials[0].thin_walled is True


def test_refusals():
    with pytest.raises(GLTFError, match="KHR_draco_mesh_compression"):
        load_gltf(triangle_glb({}, {"extensionsRequired": ["KHR_draco_mesh_compression"]}), no_fetch)
    with pytest.raises(GLTFError, match="only 2.0"):
        load_gltf(json.dumps({"asset": {"version": "1.0"}}).encode(), no_fetch)


def test_texture_transform():
    def apply(m, u, v):
        return [round(m[0] * u + m[1] * v + m[2], 6) + 0.0, round(m[3] * u + m[4] * v + m[5], 6) + 0.0]

    assert apply(texture_transform(None), 0.3, 0.7) == [0.3, 0.7]
    m = texture_transform({"offset": [0.5, 0.25], "scale": [2, 3], "rotation": math.pi / 2})
    assert apply(m, 1, 0) == [0.5, 2.25]
    assert apply(m, 0, 1) == [-2.5, 0.25]


def test_loaded_asset_packs_into_a_scene():
    """The loader's meshes and materials feed pack_scene like the generators'
    do; the caller supplies the decoded images (here a 4x4 RGBA8 stand-in for
    the embedded PNG bytes)."""
    from lucent_amd import scene as S
    a = load_gltf(triangle_glb(), no_fetch)
    decoded = [np.full((4, 4, 4), 200, np.uint8) for _ in a.images]
    sc = S.Scene(camera=S.Camera((0.3, 0.3, 2.0), (0.3, 0.3, 0.0), 45.0, 0.0, 2.0),
                 meshes=[S.MeshInstance(m, a.materials[i]) for m, i in a.meshes], textures=decoded)
    p = S.pack_scene(sc)
    assert p.tri_count == 1 and p.material_count >= 1


# ---------------------------------------------------------------------------
# Native loader edge cases beyond gltf.test.ts: fans, skipped primitives,
# warnings, unlit and transformed materials, and the error paths.
# ---------------------------------------------------------------------------
def quad_glb(prims, extra=None, positions=None):
    pos = np.float32(positions if positions is not None else [0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0])
    blob, _ = pack(pos)
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}], "meshes": [{"primitives": prims}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": pos.size // 3, "type": "VEC3"}],
          "bufferViews": [{"buffer": 0, "byteLength": pos.nbytes}], "buffers": [{"byteLength": len(blob)}]}
    js.update(extra or {})
    return make_glb(js, blob)


def test_triangle_fan_and_skipped_points():
    a = load_gltf(quad_glb([{"mode": 6, "attributes": {"POSITION": 0}}, {"mode": 0, "attributes": {"POSITION": 0}}]),
                  no_fetch)
    assert a.triangle_count == 2 and len(a.meshes) == 1
    n = a.meshes[0][0].normals
    assert abs(n[2] - 1) < 1e-6 and abs(n[11] - 1) < 1e-6  # fan triangles keep +Z winding
    assert a.warnings == ["point and line primitives are skipped"]
    assert a.bounds == ((0, 0, 0), (1, 1, 0))


def test_unlit_material_emits_its_base_colour_and_transform_is_taken():
    glb = triangle_glb({}, {"materials": [{
        "pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 1, 1], "baseColorTexture": {
            "index": 0, "texCoord": 1,
            "extensions": {"KHR_texture_transform": {"offset": [0.5, 0.25], "scale": [2, 3],
                                                     "rotation": math.pi / 2}}}},
        "extensions": {"KHR_materials_unlit": {}}}]})
    a = load_gltf(glb, no_fetch)
    m = a.materials[0]
    assert m.emission == (0.5, 0.25, 1) and m.albedo == (0, 0, 0) and m.metallic == 0
    assert m.emissive_map == 0 and m.base_color_map == 0
    assert np.allclose(m.uv_transform, texture_transform({"offset": [0.5, 0.25], "scale": [2, 3],
                                                          "rotation": math.pi / 2}))
    assert a.warnings == ["textures on TEXCOORD_1 or higher are sampled with TEXCOORD_0"]


def test_ktx2_images_are_skipped_with_a_warning():
    glb = triangle_glb({}, {"images": [{"uri": "t.ktx2"}]})
    a = load_gltf(glb, no_fetch)
    assert a.materials[0].base_color_map == -1 and a.images == []
    assert a.warnings == ["KTX2/Basis textures are not supported and were skipped"]


def test_primitives_share_one_default_material():
    a = load_gltf(quad_glb([{"attributes": {"POSITION": 0}}, {"attributes": {"POSITION": 0}, "material": 5}]),
                  no_fetch)
    assert [i for _, i in a.meshes] == [0, 0] and len(a.materials) == 1


def test_external_buffer_is_requested_before_images():
    blob, off = pack(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]))
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
          "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "material": 0}]}],
          "materials": [{"emissiveTexture": {"index": 0}}], "textures": [{"source": 0}],
          "images": [{"uri": "glow.png", "name": "glow"}],
          "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3"}],
          "bufferViews": [{"buffer": 0, "byteLength": 36}], "buffers": [{"byteLength": 36, "uri": "geo.bin"}]}
    fetched = []

    def fetch(uri):
        fetched.append(uri)
        return blob if uri == "geo.bin" else b"\x01\x02"

    a = load_gltf(json.dumps(js).encode(), fetch)
    assert fetched == ["geo.bin", "glow.png"]
    assert (a.images[0].name, a.images[0].mime_type, a.images[0].data) == ("glow", "image/png", b"\x01\x02")


@pytest.mark.parametrize("case,match", [
    ("index", "out of range"),
    ("cycle", "deeper than 64"),
    ("empty", "no triangle meshes"),
    ("view", "runs past its buffer"),
    ("accessor", "missing accessor"),
    ("json", "glTF JSON"),
])
def test_malformed_files_raise(case, match):
    if case == "index":
        blob, off = pack(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]), np.uint16([0, 1, 3]))
        js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0}],
              "meshes": [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}],
              "accessors": [{"bufferView": 0, "componentType": FLOAT, "count": 3, "type": "VEC3"},
                            {"bufferView": 1, "componentType": 5123, "count": 3, "type": "SCALAR"}],
              "bufferViews": [{"buffer": 0, "byteOffset": off[0], "byteLength": 36},
                              {"buffer": 0, "byteOffset": off[1], "byteLength": 6}],
              "buffers": [{"byteLength": len(blob)}]}
        data = make_glb(js, blob)
    elif case == "cycle":
        data = quad_glb([{"attributes": {"POSITION": 0}}],
                        {"nodes": [{"mesh": 0, "children": [1]}, {"children": [0]}], "scenes": [{"nodes": [0]}]})
    elif case == "empty":
        data = quad_glb([{"mode": 1, "attributes": {"POSITION": 0}}])
    elif case == "view":
        data = quad_glb([{"attributes": {"POSITION": 0}}], {"bufferViews": [{"buffer": 0, "byteLength": 4096}]})
    elif case == "accessor":
        data = quad_glb([{"attributes": {"POSITION": 3}}])
    else:
        data = b'{"asset": {"version": "2.0"}, '
    with pytest.raises(GLTFError, match=match):
        load_gltf(data, no_fetch)


def test_native_loader_agrees_with_the_scene_packer_material_layout():
    """The 40 texel values the native loader returns pack back to the same
    texels pack_materials writes for the Material built from them."""
    from lucent_amd import scene as S
    a = load_gltf(triangle_glb({}, {"materials": [{"alphaMode": "BLEND", "emissiveFactor": [0.1, 0.2, 0.3],
                                                   "normalTexture": {"index": 0, "scale": 2}}]}), no_fetch)
    m = a.materials[0]
    assert (m.alpha_mode, m.normal_map, m.normal_scale) == ("blend", 0, 2)
    t = S.pack_materials([m])
    assert t[15] == 2 and t[18] == 0 and t[31] == 0  # blend, normal map 0, no attenuation distance
