"""Scene packing: the reference's packScene / material / quadFromMesh tests
(test/scene.test.ts:33-186) restated against both the product's packer
(lucent_amd.scene) and the oracle's independent one (oracle/packing.py),
then the two packers compared array for array on the parity-test scenes, so
that a packing mistake on one side shows instead of being shared."""
import math

import numpy as np
import pytest

import oracle as O
from oracle import packing as OP
from lucent_amd import scene as S
from lucent_amd import workloads

PACKERS = ["product", "oracle"]


def pack(which, scene, bvh=None):
    if which == "product":
        return S.pack_scene(scene, bvh)
    return OP.pack_scene(scene, bvh)


def base_scene():
    """scene.test.ts:8-21."""
    return S.Scene(camera=S.Camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 45.0, 0.0, 5.0), meshes=[],
                   sun=S.Sun(45.0, 0.0, (1.0, 1.0, 1.0), 1.0, 1.0),
                   sky=S.Sky(True, (0.0, 0.0, 1.0), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), 1.0),
                   ground=S.Ground(False, 0.0, S.matte((0.5, 0.5, 0.5))))


def fake_textures(n):
    return [np.full((4, 4, 4), 17 * (i + 1), np.uint8) for i in range(n)]


def r5(x):
    if isinstance(x, (tuple, list)):
        return tuple(r5(v) for v in x)
    if isinstance(x, (bool, str)):
        return x
    x = float(x)
    return round(x, 5) if math.isfinite(x) else x


@pytest.mark.parametrize("which", PACKERS)
def test_materials_round_trip(which):
    """scene.test.ts:33-46."""
    mats = [S.matte((0.1, 0.2, 0.3)), S.metal((1.0, 0.8, 0.5), 0.2), S.glass((0.9, 1.0, 0.9), 1.33, 0.1),
            S.light((1.0, 0.5, 0.25), 10.0)]
    packed = S.pack_materials(mats) if which == "product" else OP.pack_materials(mats)
    assert packed.size == 4 * 40
    for i, m in enumerate(mats):
        u = OP.unpack_material(packed, i)
        assert r5(u["albedo"]) == r5(m.albedo)
        assert u["metallic"] == pytest.approx(m.metallic, abs=5e-3)
        assert u["roughness"] == pytest.approx(m.roughness, abs=5e-3)
        assert u["ior"] == pytest.approx(m.ior, abs=5e-3)
        assert u["transmission"] == pytest.approx(m.transmission, abs=5e-3)
        assert tuple(round(x, 4) for x in u["emission"]) == tuple(round(x, 4) for x in m.emission)


@pytest.mark.parametrize("which", PACKERS)
def test_material_extension_fields_round_trip(which):
    """scene.test.ts:48-71."""
    m = S.material(alpha=0.5, alpha_mode="blend", alpha_cutoff=0.25, clearcoat=1.0, clearcoat_roughness=0.125,
                   sheen_color=(0.5, 0.25, 0.125), sheen_roughness=0.75, thin_walled=True,
                   attenuation_color=(1.0, 0.5, 0.25), attenuation_distance=3.0, base_color_map=2,
                   metal_rough_map=0, normal_map=1, emissive_map=3, normal_scale=0.5,
                   uv_transform=(2.0, 0.0, 0.5, 0.0, 2.0, 0.25))
    pk = S.pack_materials if which == "product" else OP.pack_materials
    u = OP.unpack_material(pk([m]), 0)
    for k, v in u.items():
        assert r5(v) == r5(getattr(m, k)), k
    # Infinite attenuation distance survives as "no absorption".
    assert OP.unpack_material(pk([S.material()]), 0)["attenuation_distance"] == math.inf


@pytest.mark.parametrize("which", PACKERS)
def test_quad_from_mesh(which):
    """scene.test.ts:74-87."""
    qf = S.quad_from_mesh if which == "product" else (lambda mesh: OP.quad_from_mesh(mesh.positions))
    q = qf(S.quad((0, 2, 0), (1, 2, 0), (1, 2, 1), (0, 2, 1)))
    assert q is not None
    assert [float(x) for x in q[0]] == [0, 2, 0]
    assert [float(x) for x in q[1]] == [1, 0, 0]
    assert [float(x) for x in q[2]] == [0, 0, 1]
    assert qf(S.torus_knot(2, 3, 1.0, 0.32, 4, 3)) is None


@pytest.mark.parametrize("which", PACKERS)
def test_shared_materials_deduplicate(which):
    """scene.test.ts:90-104."""
    white = S.matte((1.0, 1.0, 1.0))
    s = base_scene()
    s.meshes = [S.MeshInstance(S.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), white),
                S.MeshInstance(S.quad((2, 0, 0), (3, 0, 0), (3, 1, 0), (2, 1, 0)), white)]
    s.spheres = [S.Sphere((0.0, 0.0, 0.0), 1.0, white)]
    p = pack(which, s)
    assert p.material_count == 2  # white + the (disabled) ground
    assert p.tri_count == 4
    for t in range(p.tri_count):
        assert p.tri_geometry[t * 12 + 3] == 0
    assert p.sphere_materials[0] == 0
    assert p.light_count == 0


@pytest.mark.parametrize("which", PACKERS)
def test_emissive_spheres_and_quads_become_lights(which):
    """scene.test.ts:106-122."""
    s = base_scene()
    lamp = S.light((1.0, 1.0, 1.0), 5.0)
    s.spheres = [S.Sphere((0.0, 3.0, 0.0), 0.5, lamp)]
    s.meshes = [S.MeshInstance(S.quad((-1, 4, -1), (1, 4, -1), (1, 4, 1), (-1, 4, 1)), lamp)]
    p = pack(which, s)
    assert p.light_count == 2
    assert p.lights[3] == 1 and p.lights[16 + 3] == 0  # quad first, then the sphere
    assert p.sphere_lights[0] == 1
    assert p.tri_geometry[0 * 12 + 7] == 0 and p.tri_geometry[1 * 12 + 7] == 0
    assert p.lights[3] == 1
    assert p.lights[12] == 5


@pytest.mark.parametrize("which", PACKERS)
def test_texture_layers_per_usage(which):
    """scene.test.ts:124-140."""
    s = base_scene()
    s.textures = fake_textures(4)
    a = S.material(base_color_map=3, normal_map=1, metal_rough_map=1)
    b = S.material(base_color_map=0, emissive_map=3, metal_rough_map=2)
    s.meshes = [S.MeshInstance(S.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), a),
                S.MeshInstance(S.quad((2, 0, 0), (3, 0, 0), (3, 1, 0), (2, 1, 0)), b)]
    p = pack(which, s)
    assert list(p.color_layers) == [3, 0]
    assert list(p.data_layers) == [1, 2]
    ua = OP.unpack_material(p.materials, 0)
    ub = OP.unpack_material(p.materials, 1)
    assert [ua["base_color_map"], ua["normal_map"], ua["metal_rough_map"], ua["emissive_map"]] == [0, 0, 0, -1]
    assert [ub["base_color_map"], ub["emissive_map"], ub["metal_rough_map"]] == [1, 0, 1]


@pytest.mark.parametrize("which", PACKERS)
def test_maps_past_the_scene_textures_throw(which):
    """scene.test.ts:142-146."""
    s = base_scene()
    s.meshes = [S.MeshInstance(S.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), S.material(normal_map=0))]
    with pytest.raises(ValueError, match="texture 0"):
        pack(which, s)


@pytest.mark.parametrize("which", PACKERS)
def test_uvs_in_leaf_order_and_alpha_flags(which):
    """scene.test.ts:148-167."""
    s = base_scene()
    s.textures = fake_textures(1)
    mesh = S.quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    uvs = np.array([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.1, 0.2, 0.5, 0.6, 0.7, 0.8], np.float32)
    s.meshes = [S.MeshInstance(S.Mesh(mesh.positions, mesh.normals, uvs),
                               S.material(alpha_mode="mask", base_color_map=0)),
                S.MeshInstance(S.quad((2, 0, 0), (3, 0, 0), (3, 1, 0), (2, 1, 0)), S.matte((1.0, 1.0, 1.0)))]
    p = pack(which, s)
    rows = []
    for slot in range(p.tri_count):
        g, a = slot * 12, slot * 16
        us = [p.tri_attributes[a + 3], p.tri_attributes[a + 7], p.tri_attributes[a + 11]]
        vs = [p.tri_attributes[a + 12], p.tri_attributes[a + 13], p.tri_attributes[a + 14]]
        uv = [round(float(x), 4) for pair in zip(us, vs) for x in pair]
        rows.append((p.tri_geometry[g + 3], p.tri_geometry[g + 11], uv))
    textured = [r for r in rows if r[0] == 0]
    assert len(textured) == 2
    assert all(r[1] == 1 for r in textured)
    assert [0.1, 0.2, 0.3, 0.4, 0.5, 0.6] in [r[2] for r in textured]
    assert all(r[1] == 0 and all(x == 0 for x in r[2]) for r in rows if r[0] == 1)


def test_ground_material_and_pick():
    """scene.test.ts:169-180 (the pick side on the oracle's pickDistance; the
    product's pick runs on the GPU, tests/test_pick_gpu.py)."""
    for which in PACKERS:
        s = base_scene()
        s.ground = S.Ground(True, -1.0, S.matte((0.3, 0.3, 0.3)))
        p = pack(which, s)
        assert p.ground_material == 0
    p = OP.pack_scene(s)
    assert OP.pick_distance(p, (0, 1, 0), (0, -1, 0)) == pytest.approx(2)
    assert OP.pick_distance(p, (0, 1, 0), (0, 1, 0)) is None
    s.ground.enabled = False  # the shared ground object is read at pick time
    assert OP.pick_distance(p, (0, 1, 0), (0, -1, 0)) is None


@pytest.mark.parametrize("which", PACKERS)
def test_too_many_spheres_throw(which):
    """scene.test.ts:182-186."""
    s = base_scene()
    s.spheres = [S.Sphere((0.0, 0.0, 0.0), 1.0, S.matte((1.0, 1.0, 1.0))) for _ in range(65)]
    with pytest.raises(ValueError, match="spheres"):
        pack(which, s)


# ---------------------------------------------------------------------------
# The two packers agree on every parity-test scene.
# ---------------------------------------------------------------------------
def textured_scene():
    """Every map slot in use, images of several sizes (identity, magnified
    and minified layers), UVs on some meshes only, an alpha-masked mesh."""
    rng = np.random.default_rng(5)
    tex = [rng.integers(0, 256, (64, 64, 4), dtype=np.uint8), rng.integers(0, 256, (32, 32, 4), dtype=np.uint8),
           rng.integers(0, 256, (128, 64, 4), dtype=np.uint8), rng.integers(0, 256, (64, 64, 4), dtype=np.uint8)]
    knot = S.torus_knot(2, 3, 1.0, 0.3, 60, 8, uvs=True)
    mats = [S.material(base_color_map=0, metal_rough_map=1, normal_map=2, emissive_map=3, emission=(0.5, 0.5, 0.5)),
            S.material(alpha_mode="mask", base_color_map=2, alpha_cutoff=0.4),
            S.matte((0.2, 0.4, 0.6))]
    return S.Scene(camera=S.Camera((0.0, 1.0, 5.0), (0.0, 0.0, 0.0)),
                   meshes=[S.MeshInstance(knot, mats[0]),
                           S.MeshInstance(S.transform(knot, translate=(2.0, 0.0, 0.0)), mats[1]),
                           S.MeshInstance(S.quad((-3, 0, -3), (3, 0, -3), (3, 0, 3), (-3, 0, 3)), mats[2]),
                           S.MeshInstance(S.quad((-1, 3, -1), (1, 3, -1), (1, 3, 1), (-1, 3, 1)),
                                          S.light((1.0, 1.0, 1.0), 4.0))],
                   spheres=[S.Sphere((0.0, 2.0, 0.0), 0.3, S.light((1.0, 0.5, 0.5), 3.0)),
                            S.Sphere((1.0, 0.5, 1.0), 0.5, mats[2])],
                   ground=S.Ground(True, -0.5, S.matte((0.5, 0.5, 0.5))), textures=tex)


SCENES = {
    "textured": textured_scene,
    "knots": lambda: workloads.knot_scene(),
    "spheres": lambda: workloads.spheres_scene(),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_packers_agree_bit_for_bit(name):
    sc = SCENES[name]()
    a = S.pack_scene(sc)
    b = OP.pack_scene(sc)
    for k in ("tri_geometry", "tri_attributes", "bvh_nodes", "tri_indices", "materials", "spheres",
              "sphere_materials", "sphere_lights", "lights"):
        x, y = np.asarray(getattr(a, k)), np.asarray(getattr(b, k))
        assert x.shape == y.shape, k
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{name}: {k} differs"
    for k in ("tri_count", "bvh_node_count", "material_count", "sphere_count", "light_count",
              "ground_material", "ground_material_active"):
        assert getattr(a, k) == getattr(b, k), k
    assert list(a.color_layers) == list(b.color_layers) and list(a.data_layers) == list(b.data_layers)


def test_texture_layers_agree():
    sc = textured_scene()
    a = S.pack_scene(sc)
    b = OP.pack_scene(sc)
    sa, ca, da = S.texture_arrays(sc, a)
    sb, cb, db = OP.texture_layers(sc, b)
    assert sa == sb == 128
    assert np.array_equal(ca, cb) and np.array_equal(da, db)


def test_resample_identity_and_closed_forms():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (16, 16, 4), dtype=np.uint8)
    # Texel centres map onto texel centres at equal size.
    assert np.array_equal(OP.resample_layer(img, 16, True), img)
    # A constant image stays constant under magnification and minification,
    # in both colour spaces.
    const = np.full((8, 8, 4), 200, np.uint8)
    for size in (4, 16, 32):
        for srgb in (False, True):
            assert np.all(OP.resample_layer(const, size, srgb) == 200)
    # Halving a linear checker of 2x2 blocks averages each block exactly.
    chk = np.zeros((8, 8, 4), np.uint8)
    chk[::2, ::2] = 100
    chk[1::2, 1::2] = 200
    half = OP.resample_layer(chk, 4, False)
    assert np.all(half == 75)
    # Magnifying a linear ramp 2x interpolates it (clamped at the edges).
    ramp = np.zeros((1, 4, 4), np.uint8)
    ramp[0, :, 0] = [0, 40, 80, 120]
    up = OP.resample_layer(np.repeat(ramp, 4, axis=0), 8, False)
    assert list(up[0, :, 0]) == [0, 10, 30, 50, 70, 90, 110, 120]
