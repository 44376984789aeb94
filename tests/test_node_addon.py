"""The Node host path: the N-API addon and node/lucent.js facade over the same
C ABI.  CPU: buildBVH through Node equals the product and oracle builders and
throws the reference's Error on a bad length.  GPU: intersect through Node is
bit-identical to the oracle, and a LucentRenderer render equals the Python
binding's render of the same packed scene."""
import json
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from lucent_amd.bvh import build_bvh

ROOT = Path(__file__).resolve().parents[1]
CHECK = ROOT / "node" / "test" / "check.js"
pytestmark = pytest.mark.skipif(shutil.which("node") is None or not (ROOT / "node/build/lucent_napi.node").exists(),
                                reason="node or the built addon is missing")


def node(mode, d):
    r = subprocess.run(["node", str(CHECK), mode, str(d)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_build_bvh_through_node(tmp_path):
    soup = O.random_soup(3000, O.Lcg(21))
    soup.tofile(tmp_path / "positions.f32")
    meta = node("bvh", tmp_path)
    nodes = np.fromfile(tmp_path / "nodes.f32", np.float32)
    tri = np.fromfile(tmp_path / "tri.u32", np.uint32)
    ref = O.build_bvh(soup)
    assert meta["nodeCount"] == ref.node_count
    assert np.array_equal(nodes.view(np.uint32), ref.nodes.view(np.uint32))
    assert np.array_equal(tri, ref.tri_indices)
    assert np.array_equal(tri, build_bvh(soup).tri_indices)
    assert meta["error"] == "positions length 10 is not a multiple of 9"
    # buildBVHAsync (bvh-async.ts): the same tree from a worker thread, the same error as a rejection.
    assert meta["asyncSame"] is True
    assert meta["asyncErr"] == "positions length 10 is not a multiple of 9"


def test_sync_members_refuse_while_an_async_call_runs(tmp_path):
    """A libuv worker may be running lucent_trace_samples on the renderer's
    context: every synchronous member that touches the context, scene or film
    (uploadLighting frees env buffers the worker reads; reset races the film's
    seed and sample count) throws instead of running."""
    out = node("guard", tmp_path)
    for name, msg in out.items():
        assert "asynchronous call is still running" in msg, f"{name}: {msg}"
    assert set(out) >= {"loadScene", "uploadLighting", "uploadCamera", "reset", "traceSamples", "render",
                        "screenshot", "benchmark", "runDenoiser", "read", "stats", "dispose"}


@pytest.mark.gpu
def test_intersect_through_node_bit_identical(tmp_path):
    rand = O.Lcg(42)
    soup = O.random_soup(2000, rand)
    org = np.zeros((400, 3))
    dirs = np.zeros((400, 3))
    for i in range(400):
        org[i] = [rand() * 16 - 8, rand() * 16 - 8, rand() * 16 - 8]
        d = np.array([rand() - 0.5, rand() - 0.5, rand() - 0.5])
        dirs[i] = d / math.sqrt(float(d @ d))
    b = build_bvh(soup)
    soup.tofile(tmp_path / "positions.f32")
    b.nodes.tofile(tmp_path / "nodes.f32")
    b.tri_indices.tofile(tmp_path / "tri.u32")
    org.tofile(tmp_path / "org.f64")
    dirs.tofile(tmp_path / "dir.f64")
    meta = node("hits", tmp_path)
    t, tri, u, v = O.intersect_bvh_batch(b, soup, org, dirs)
    assert np.array_equal(np.fromfile(tmp_path / "hit.u32", np.uint32), tri)
    for name, ref in (("t", t), ("u", u), ("v", v)):
        got = np.fromfile(tmp_path / f"{name}.f64", np.float64)
        assert np.array_equal(got, ref), name
    if tri[0] != 0xFFFFFFFF:
        assert meta["first"]["triIndex"] == int(tri[0]) and meta["first"]["t"] == t[0]


def test_parse_hdr_through_node(tmp_path):
    """parseHDR through the addon gives the C-ABI decoder's floats and its error messages."""
    from lucent_amd.assets import decode_hdr
    width = 8
    data = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 8\n" + bytes(
        [2, 2, 0, width, 128 + 8, 200, 4, 1, 2, 3, 4, 128 + 4, 9, 128 + 8, 0, 128 + 8, 128] +
        [7, 8, 9, 130] * width)
    (tmp_path / "in.hdr").write_bytes(data)
    meta = node("hdr", tmp_path)
    w, h, rgb = decode_hdr(data)
    assert (meta["width"], meta["height"]) == (w, h) == (8, 2)
    assert np.array_equal(np.fromfile(tmp_path / "rgb.f32", np.float32).reshape(h, w, 3), rgb)
    assert "Radiance" in meta["error"]


def test_load_gltf_through_node(tmp_path):
    """loadGLTF through the addon: the same world-space positions, material
    and images as the Python binding of the native loader, external files
    fetched buffers first, and the reference's version error."""
    import base64
    from lucent_amd.gltf import load_gltf
    pos = np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0])
    (tmp_path / "geo.bin").write_bytes(pos.tobytes())
    (tmp_path / "t.png").write_bytes(b"\x89PNG")
    js = {"asset": {"version": "2.0"}, "nodes": [{"mesh": 0, "translation": [0, 0, 2], "scale": [-1, 1, 1]}],
          "meshes": [{"primitives": [{"mode": 5, "attributes": {"POSITION": 0}, "material": 0}]}],
          "materials": [{"emissiveTexture": {"index": 0}, "emissiveFactor": [1, 1, 1],
                         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 3}}}],
          "textures": [{"source": 0}], "images": [{"uri": "t.png"}],
          "accessors": [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"}],
          "bufferViews": [{"buffer": 0, "byteLength": pos.nbytes}],
          "buffers": [{"byteLength": pos.nbytes, "uri": "geo.bin"}]}
    data = json.dumps(js).encode()
    (tmp_path / "in.gltf").write_bytes(data)
    meta = node("gltf", tmp_path)
    a = load_gltf(data, lambda uri: (tmp_path / uri).read_bytes())
    assert meta["fetched"] == ["geo.bin", "t.png"]
    assert meta["triangleCount"] == a.triangle_count == 2
    assert np.array_equal(np.fromfile(tmp_path / "positions.f32", np.float32), a.meshes[0][0].positions)
    assert meta["bounds"] == {"min": list(a.bounds[0]), "max": list(a.bounds[1])}
    m = a.materials[0]
    assert meta["material"]["emission"] == list(m.emission) == [3, 3, 3]
    assert meta["material"]["emissiveMap"] == m.emissive_map == 0
    assert meta["material"]["attenuationDistance"] is None  # Infinity serialises as null
    assert meta["images"] == [{"mimeType": "image/png", "name": "t.png", "data": list(b"\x89PNG")}]
    assert meta["hasUvs"] is False and meta["warnings"] == []
    assert "only 2.0" in meta["error"]


def write_render_inputs(tmp_path, W=48, H=32, SPP=3, SEED=9):
    """The feature scene's PackedScene arrays, lighting and cameras as files for node/test/check.js."""
    import sys
    sys.path.insert(0, str(Path(__file__).parent))
    from test_pathtrace_gpu import feature_scene
    from lucent_amd.scene import pack_scene
    from lucent_amd import scene as S
    sc = feature_scene()
    p = pack_scene(sc)
    for name, arr in (("geo.f32", p.tri_geometry), ("attr.f32", p.tri_attributes), ("nodes.f32", p.bvh_nodes),
                      ("mat.f32", p.materials), ("sph.f32", p.spheres), ("sphm.i32", p.sphere_materials),
                      ("sphl.i32", p.sphere_lights), ("lights.f32", p.lights), ("tri.u32", p.tri_indices)):
        np.ascontiguousarray(arr).tofile(tmp_path / name)
    w, h, env = sc.env.map
    np.asarray(env, np.float32).tofile(tmp_path / "env.f32")
    scene_json = {
        "camera": {"position": list(sc.camera.position), "target": list(sc.camera.target), "fov": sc.camera.fov,
                   "aperture": sc.camera.aperture, "focusDistance": sc.camera.focus_distance},
        "sun": {"elevation": sc.sun.elevation, "azimuth": sc.sun.azimuth, "color": list(sc.sun.color),
                "intensity": sc.sun.intensity, "angularRadius": sc.sun.angular_radius},
        "sky": {"enabled": sc.sky.enabled, "zenith": list(sc.sky.zenith), "horizon": list(sc.sky.horizon),
                "ground": list(sc.sky.ground), "intensity": sc.sky.intensity},
        "env": {"rotation": sc.env.rotation, "intensity": sc.env.intensity, "background": sc.env.background,
                "backdrop": list(sc.env.backdrop)},
        "ground": {"enabled": sc.ground.enabled, "height": sc.ground.height},
    }
    c2 = S.Camera(position=(0.4, 2.6, 5.9), target=sc.camera.target, fov=sc.camera.fov, aperture=sc.camera.aperture,
                  focus_distance=sc.camera.focus_distance)
    cam2 = {"position": list(c2.position), "target": list(c2.target), "fov": c2.fov, "aperture": c2.aperture,
            "focusDistance": c2.focus_distance}
    meta = {"triCount": p.tri_count, "nodeCount": p.bvh_node_count, "materialCount": p.material_count,
            "sphereCount": p.sphere_count, "lightCount": p.light_count, "groundMaterial": p.ground_material,
            "envW": w, "envH": h, "width": W, "height": H, "spp": SPP, "seed": SEED, "scene": scene_json,
            "camera2": cam2}
    (tmp_path / "meta.json").write_text(json.dumps(meta))
    return sc, c2


@pytest.mark.gpu
def test_renderer_through_node_matches_python_binding(tmp_path, ctx):
    from lucent_amd.renderer import Renderer, RenderSettings
    W, H, SPP, SEED = 48, 32, 3, 9
    sc, c2 = write_render_inputs(tmp_path, W, H, SPP, SEED)
    stats = node("render", tmp_path)
    assert stats["samples"] == SPP and stats["seed"] == SEED + SPP
    r = Renderer(ctx, W, H, RenderSettings(8, 0.0))
    r.load_scene(sc)
    r.film.set_seed(SEED)
    r.reset()
    r.trace_samples(SPP)
    ref = r.film.read()
    r.run_denoiser(4.0, 64.0, 1.5, 3)
    den = r.read_denoised()
    r.upload_camera(c2)
    r.trace_samples(2)
    moved = r.film.read()["color"]
    shot = r.screenshot()
    r.close()
    assert stats["movedSamples"] == SPP + 2
    assert np.array_equal(np.fromfile(tmp_path / "moved.f32", np.float32).reshape(H, W, 4), moved)
    assert np.array_equal(np.fromfile(tmp_path / "shot.u8", np.uint8).reshape(H, W, 4), shot)
    assert np.array_equal(np.fromfile(tmp_path / "color.f32", np.float32).reshape(H, W, 4), ref["color"])
    assert np.array_equal(np.fromfile(tmp_path / "albedo.f32", np.float32).reshape(H, W, 4), ref["albedo"])
    assert np.array_equal(np.fromfile(tmp_path / "den.f32", np.float32).reshape(H, W, 4), den["illum_var"])


@pytest.mark.gpu
def test_async_calls_through_node_equal_sync_calls(tmp_path, ctx):
    """traceSamplesAsync and runDenoiserAsync on worker threads (napi_async_work),
    queued on one renderer, leave the film and the denoised image bit-identical
    to the synchronous calls; a synchronous call while they run is refused."""
    write_render_inputs(tmp_path)
    meta = node("async", tmp_path)
    assert meta["filmSame"] is True and meta["denSame"] is True
    assert "asynchronous call is still running" in meta["guard"]
    assert meta["samples"] == json.loads((tmp_path / "meta.json").read_text())["spp"]
