"""TEST INFRASTRUCTURE — Python access to the FP64 parity oracle.

Loads ``oracle/liboracle.so`` (built by ``make -C oracle``) through ctypes.
Only the test suite, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg may import this package; the product never does.
"""
from __future__ import annotations

import concurrent.futures as cf
import math
import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None

_P = C.c_void_p
_U32 = C.c_uint32
_D = C.c_double


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.or_lcg_next.restype = _D
        l.or_lcg_next.argtypes = [C.POINTER(_U32)]
        l.or_random_soup.argtypes = [_U32, C.POINTER(_U32), _P]
        l.or_random_soup.restype = None
        l.or_intersect_triangle.argtypes = [_P, _P, _P, _P, _P, _D, _D, _P, _P, _P]
        l.or_intersect_triangle.restype = C.c_int
        l.or_intersect_sphere.argtypes = [_P, _P, _P, _D, _D, _D]
        l.or_intersect_sphere.restype = _D
        l.or_aabb_intersect.argtypes = [_P, _P, _P, _P, _D]
        l.or_aabb_intersect.restype = _D
        l.or_build_bvh.argtypes = [_P, _U32, _U32, _U32, _P, C.POINTER(_U32), _P, _P]
        l.or_build_bvh.restype = C.c_int
        l.or_intersect_bvh_batch.argtypes = [_P, _P, _U32, _P, _P, _P, _U32, _D, _D, _P, _P, _P, _P]
        l.or_intersect_bvh_batch.restype = C.c_int
        l.or_intersect_bvh_batch_f64.argtypes = [_P, _P, _U32, _P, _P, _P, _U32, _D, _D, _P, _P, _P, _P]
        l.or_intersect_bvh_batch_f64.restype = C.c_int
        l.or_intersect_all_batch_f64.argtypes = [_P, _U32, _P, _P, _U32, _D, _D, _P, _P, _P, _P]
        l.or_intersect_all_batch_f64.restype = C.c_int
        l.or_shader_order_counts.argtypes = [_P, _P, _U32, _P, _P, _P, _U32, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        l.or_shader_order_counts.restype = None
        l.or_bvh_depth.argtypes = [_P, _U32]
        l.or_bvh_depth.restype = _U32
        l.or_torus_knot.argtypes = [_D, _D, _D, _D, _U32, _U32, _P, _P]
        l.or_torus_knot.restype = None
        _lib = l
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Lcg:
    """test/bvh.test.ts:7-13 — the reference suite's deterministic LCG."""

    def __init__(self, seed: int):
        self.state = C.c_uint32(seed & 0xFFFFFFFF)

    def __call__(self) -> float:
        return lib().or_lcg_next(C.byref(self.state))


def random_soup(count: int, rand: Lcg) -> np.ndarray:
    out = np.zeros(count * 9, np.float32)
    lib().or_random_soup(count, C.byref(rand.state), _ptr(out))
    return out


def torus_knot(p=2.0, q=3.0, radius=1.0, tube=0.32, tubular=200, radial=20):
    n = tubular * radial * 2
    pos = np.zeros(n * 9, np.float32)
    nrm = np.zeros(n * 9, np.float32)
    lib().or_torus_knot(p, q, radius, tube, tubular, radial, _ptr(pos), _ptr(nrm))
    return pos, nrm


@dataclass
class OracleBVH:
    nodes: np.ndarray
    node_count: int
    tri_indices: np.ndarray
    bounds: np.ndarray


def build_bvh(positions: np.ndarray, max_leaf: int = 4, bins: int = 12) -> OracleBVH:
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    n = pos.size // 9
    nodes = np.zeros(max(1, 2 * n - 1) * 8, np.float32)
    tri = np.zeros(n, np.uint32)
    bounds = np.zeros(6, np.float64)
    cnt = C.c_uint32()
    r = lib().or_build_bvh(_ptr(pos), n, max_leaf, bins, _ptr(nodes), C.byref(cnt), _ptr(tri), _ptr(bounds))
    assert r == 0
    return OracleBVH(nodes[: cnt.value * 8].copy(), cnt.value, tri, bounds)


def intersect_triangle(ro, rd, v0, v1, v2, tmin=1e-5, tmax=float("inf")):
    arrs = [np.ascontiguousarray(x, np.float64) for x in (ro, rd, v0, v1, v2)]
    t = np.zeros(1)
    u = np.zeros(1)
    v = np.zeros(1)
    if lib().or_intersect_triangle(*map(_ptr, arrs), tmin, tmax, _ptr(t), _ptr(u), _ptr(v)):
        return float(t[0]), float(u[0]), float(v[0])
    return None


def intersect_sphere(ro, rd, c, r, tmin=1e-5, tmax=float("inf")):
    arrs = [np.ascontiguousarray(x, np.float64) for x in (ro, rd, c)]
    t = lib().or_intersect_sphere(*map(_ptr, arrs), r, tmin, tmax)
    return None if t != t else t


def aabb_intersect(bmin, bmax, origin, inv_dir, tmax=float("inf")) -> float:
    arrs = [np.ascontiguousarray(x, np.float64) for x in (bmin, bmax, origin, inv_dir)]
    return lib().or_aabb_intersect(*map(_ptr, arrs), tmax)


def _split(n: int, parts: int):
    step = (n + parts - 1) // parts
    return [(i, min(n, i + step)) for i in range(0, n, step)] if n else []


def intersect_bvh_batch(bvh, positions, org, dirs, tmin=1e-5, tmax=float("inf"), threads: int | None = None):
    """Hit records of intersectBVH for every ray; (t, tri, u, v) with tri = 0xFFFFFFFF on a miss.

    Rays may be float32 or float64; ctypes releases the GIL so chunks run in parallel.
    """
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    f64 = np.asarray(org).dtype == np.float64
    dt = np.float64 if f64 else np.float32
    o = np.ascontiguousarray(org, dt).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, dt).reshape(-1, 3)
    n = o.shape[0]
    t = np.empty(n)
    u = np.empty(n)
    v = np.empty(n)
    tri = np.empty(n, np.uint32)
    nodes = np.ascontiguousarray(bvh.nodes, np.float32)
    idx = np.ascontiguousarray(bvh.tri_indices, np.uint32)
    fn = lib().or_intersect_bvh_batch_f64 if f64 else lib().or_intersect_bvh_batch
    threads = threads or min(16, os.cpu_count() or 1)
    esz = o.itemsize * 3

    def run(rng):
        a, b = rng
        return fn(_ptr(nodes), _ptr(idx), idx.size, _ptr(pos), _ptr(o) + a * esz, _ptr(d) + a * esz, b - a, tmin,
                  tmax, _ptr(t) + a * 8, _ptr(tri) + a * 4, _ptr(u) + a * 8, _ptr(v) + a * 8)

    with cf.ThreadPoolExecutor(threads) as ex:
        overflow = sum(ex.map(run, _split(n, threads * 4)))
    if overflow:
        raise RuntimeError(f"oracle traversal stack overflowed on {overflow} rays")
    return t, tri, u, v


def intersect_all_batch(positions, org, dirs, tmin=1e-5, tmax=float("inf")):
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    o = np.ascontiguousarray(org, np.float64).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = o.shape[0]
    t = np.empty(n)
    u = np.empty(n)
    v = np.empty(n)
    tri = np.empty(n, np.uint32)
    lib().or_intersect_all_batch_f64(_ptr(pos), pos.size // 9, _ptr(o), _ptr(d), n, tmin, tmax, _ptr(t), _ptr(tri),
                                     _ptr(u), _ptr(v))
    return t, tri, u, v


def shader_order_counts(bvh, positions, org, dirs) -> dict:
    """Mean interior visits P, triangle tests T and leaf visits per ray in the shader's order."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    o = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    P, T, L = C.c_uint64(), C.c_uint64(), C.c_uint64()
    nodes = np.ascontiguousarray(bvh.nodes, np.float32)
    idx = np.ascontiguousarray(bvh.tri_indices, np.uint32)
    lib().or_shader_order_counts(_ptr(nodes), _ptr(idx), idx.size, _ptr(pos), _ptr(o), _ptr(d), o.shape[0],
                                 C.byref(P), C.byref(T), C.byref(L))
    n = max(1, o.shape[0])
    return {"P": P.value / n, "T": T.value / n, "leaf_visits": L.value / n, "rays": o.shape[0]}


def bvh_depth(bvh) -> int:
    nodes = np.ascontiguousarray(bvh.nodes, np.float32)
    return int(lib().or_bvh_depth(_ptr(nodes), bvh.node_count))


# ---------------------------------------------------------------------------
# Integrator oracle (pt_oracle.c)
# ---------------------------------------------------------------------------
_F = C.c_float
_I = C.c_int32


class OrScene(C.Structure):
    _fields_ = [
        ("geo", _P), ("attr", _P), ("tri_count", _U32), ("nodes", _P), ("node_count", _U32),
        ("materials", _P), ("material_count", _U32), ("spheres", _P), ("sphere_mat", _P), ("sphere_light", _P),
        ("sphere_count", _U32), ("lights", _P), ("light_count", _U32), ("ground_material", _I),
        ("ground_height", _F), ("env_rgb", _P), ("env_w", _I), ("env_h", _I), ("env_table", _P), ("grid_w", _I),
        ("grid_h", _I), ("env_cos", _F), ("env_sin", _F), ("env_intensity", _F), ("env_sampling", _I),
        ("env_background", _I), ("backdrop", _F * 3), ("sun_enabled", _I), ("sun_dir", _F * 3),
        ("sun_radiance", _F * 3), ("sun_cos", _F), ("sun_pdf", _F), ("sheen_lut", _P), ("color_maps", _P),
        ("data_maps", _P), ("tex_size", _I), ("color_layers", _I), ("data_layers", _I),
    ]


class OrFrame(C.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("max_bounces", _U32), ("clamp", _F), ("cam_pos", _F * 3),
                ("cam_right", _F * 3), ("cam_up", _F * 3), ("cam_fwd", _F * 3), ("tan_half_fov", _F),
                ("aperture", _F), ("focus_dist", _F)]


class OrPrev(C.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("cam_pos", _F * 3), ("cam_right", _F * 3), ("cam_up", _F * 3),
                ("cam_fwd", _F * 3), ("tan_half_fov", _F), ("history_cap", _F), ("color4", _P), ("albedo4", _P),
                ("normal4", _P), ("history", _P)]


def _pt_lib():
    l = lib()
    if not getattr(l, "_pt_ready", False):
        l.or_trace_pixels.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), _U32, _U32, _P, _U32, _P, _P, _P]
        l.or_trace_pixels.restype = C.c_int
        l.or_trace_reprojected.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), C.POINTER(OrPrev), _U32, _P, _U32,
                                           _P, _P]
        l.or_trace_reprojected.restype = C.c_int
        l.or_alias_table.argtypes = [_P, _U32, _P, _P, _P, C.POINTER(_D)]
        l.or_alias_table.restype = C.c_int
        l.or_env_sampler.argtypes = [_P, _U32, _U32, _U32, _P, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_D)]
        l.or_env_sampler.restype = C.c_int
        l.or_sheen_albedo.argtypes = [_D, _D, _U32]
        l.or_sheen_albedo.restype = _D
        l.or_round_half.argtypes = [_F]
        l.or_round_half.restype = _F
        l.or_sheen_lut_half.argtypes = [_P]
        l.or_sheen_lut_half.restype = None
        l.or_eval_opaque.argtypes = [_P, _P, _P, _P, _P, C.POINTER(_D)]
        l.or_eval_opaque.restype = None
        l.or_sample_opaque.argtypes = [_P, _P, _P, _U32, _P, _P, C.POINTER(_D)]
        l.or_sample_opaque.restype = None
        l.or_sample_dielectric.argtypes = [_P, C.c_int, _P, _U32, _P, C.POINTER(C.c_int)]
        l.or_sample_dielectric.restype = C.c_int
        l.or_pcg_stream.argtypes = [_U32, _U32, _U32, _P]
        l.or_pcg_stream.restype = _U32
        l.or_d_ggx.argtypes = [_D, _D]
        l.or_d_ggx.restype = _D
        l.or_g1_ggx.argtypes = [_D, _D]
        l.or_g1_ggx.restype = _D
        l.or_sample_vndf.argtypes = [_P, _D, _D, _D, _P]
        l.or_sample_vndf.restype = None
        l._pt_ready = True
    return l


def alias_table(weights):
    w = np.ascontiguousarray(weights, np.float64)
    n = w.size
    prob = np.zeros(n, np.float32)
    alias = np.zeros(n, np.uint32)
    pmf = np.zeros(n, np.float64)
    total = C.c_double()
    if _pt_lib().or_alias_table(_ptr(w), n, _ptr(prob), _ptr(alias), _ptr(pmf), C.byref(total)) != 0:
        raise ValueError("invalid alias weights")
    return prob, alias, pmf, total.value


def sample_alias(prob, alias, u0, u1):
    """alias.ts:68-72."""
    n = prob.size
    i = min(n - 1, int(math.floor(u0 * n)))
    return i if u1 < prob[i] else int(alias[i])


def env_sampler(rgb, width, height, max_width=1024):
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1)
    gw0 = max(1, min(width, max_width))
    table = np.zeros(gw0 * height * 4 + 4, np.float32)
    gw, gh, total = C.c_uint32(), C.c_uint32(), C.c_double()
    rc = _pt_lib().or_env_sampler(_ptr(rgb), width, height, max_width, _ptr(table), C.byref(gw), C.byref(gh),
                                  C.byref(total))
    if rc != 0:
        raise ValueError("invalid environment map")
    return table[: gw.value * gh.value * 4].copy(), gw.value, gh.value, total.value


def sheen_albedo(cos_v, roughness, steps=128):
    return _pt_lib().or_sheen_albedo(cos_v, roughness, steps)


def sheen_lut_half():
    out = np.zeros(256, np.float32)
    _pt_lib().or_sheen_lut_half(_ptr(out))
    return out


def round_half(x: float) -> float:
    return _pt_lib().or_round_half(x)


def pcg_stream(pixel_index: int, seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    _pt_lib().or_pcg_stream(pixel_index, seed, n, _ptr(out))
    return out


def eval_opaque(mat40, wo, wi, lut=None):
    m = np.ascontiguousarray(mat40, np.float32)
    lut = sheen_lut_half() if lut is None else np.ascontiguousarray(lut, np.float32)
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.ascontiguousarray(wi, np.float64)
    f = np.zeros(3)
    pdf = C.c_double()
    _pt_lib().or_eval_opaque(_ptr(m), _ptr(lut), _ptr(wo), _ptr(wi), _ptr(f), C.byref(pdf))
    return f, pdf.value


def eval_opaque_batch(mat40, wo, wi, lut=None):
    """eval_opaque for arrays of local-frame directions (n, 3); returns (f (n, 3), pdf (n,))."""
    l = _pt_lib()
    l.or_eval_opaque_batch.argtypes = [_P, _P, _U32, _P, _P, _P, _P]
    l.or_eval_opaque_batch.restype = None
    m = np.ascontiguousarray(mat40, np.float32)
    lut = sheen_lut_half() if lut is None else np.ascontiguousarray(lut, np.float32)
    wo = np.ascontiguousarray(np.broadcast_to(np.asarray(wo, np.float64), np.shape(wi)), np.float64)
    wi = np.ascontiguousarray(wi, np.float64)
    n = wi.shape[0]
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    l.or_eval_opaque_batch(_ptr(m), _ptr(lut), n, _ptr(wo), _ptr(wi), _ptr(f), _ptr(pdf))
    return f, pdf


def sample_opaque(mat40, wo, state, lut=None):
    m = np.ascontiguousarray(mat40, np.float32)
    lut = sheen_lut_half() if lut is None else np.ascontiguousarray(lut, np.float32)
    wo = np.ascontiguousarray(wo, np.float64)
    d = np.zeros(3)
    w = np.zeros(3)
    pdf = C.c_double()
    _pt_lib().or_sample_opaque(_ptr(m), _ptr(lut), _ptr(wo), state & 0xFFFFFFFF, _ptr(d), _ptr(w), C.byref(pdf))
    return d, w, pdf.value


def sample_dielectric(mat40, front, wo, state):
    m = np.ascontiguousarray(mat40, np.float32)
    wo = np.ascontiguousarray(wo, np.float64)
    wi = np.zeros(3)
    refl = C.c_int()
    ok = _pt_lib().or_sample_dielectric(_ptr(m), 1 if front else 0, _ptr(wo), state & 0xFFFFFFFF, _ptr(wi),
                                        C.byref(refl))
    return bool(ok), wi, bool(refl.value)


def camera_basis(position, target, fov):
    """renderer.ts:156-161 in float64; returns (forward, right, up, tanHalfFov)."""
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f = f / math.sqrt(float(f @ f))
    up0 = np.array([0.0, 0.0, 1.0]) if abs(f[1]) > 0.999 else np.array([0.0, 1.0, 0.0])
    r = np.cross(f, up0)
    r = r / math.sqrt(float(r @ r))
    u = np.cross(r, f)
    return f, r, u, math.tan(fov * math.pi / 360)


def bake_sky(sky, sun, width: int = 1024) -> np.ndarray:
    """bakeSky (environment.ts:106-132), float32 (width/2, width, 3) rows top to bottom."""
    l = _pt_lib()
    if not getattr(l, "_sky_ready", False):
        l.or_bake_sky.argtypes = [C.c_int, _P, _P, _P, _D, _D, _D, _P, _D, _D, C.c_int, _P]
        l.or_bake_sky.restype = None
        l.or_env_pdf_dir.argtypes = [C.POINTER(OrScene), _P]
        l.or_env_pdf_dir.restype = _D
        l._sky_ready = True
    out = np.zeros((width // 2, width, 3), np.float32)
    z, h, g, c = (np.ascontiguousarray(x, np.float64) for x in (sky.zenith, sky.horizon, sky.ground, sun.color))
    l.or_bake_sky(1 if sky.enabled else 0, _ptr(z), _ptr(h), _ptr(g), sky.intensity, sun.elevation, sun.azimuth,
                  _ptr(c), sun.intensity, sun.angular_radius, width, _ptr(out))
    return out


class PtOracle:
    """Holds the arrays an OrScene points to, all built by the oracle itself
    from a scene description: packScene (oracle/packing.py), the oracle's
    buildBVH, its texture layers, env sampler, sky bake and sheen LUT, and
    the lighting uniforms as Renderer.uploadLighting sets them
    (renderer.ts:355-413)."""

    def __init__(self, scene, bvh=None):
        from . import packing
        keep = []

        def arr(a, dt):
            x = np.ascontiguousarray(np.asarray(a, dt).reshape(-1))
            keep.append(x)
            return x.ctypes.data

        packed = packing.pack_scene(scene, bvh)
        self.packed = packed
        s = OrScene()
        s.geo = arr(packed.tri_geometry, np.float32)
        s.attr = arr(packed.tri_attributes, np.float32)
        s.tri_count = packed.tri_count
        s.nodes = arr(packed.bvh_nodes, np.float32)
        s.node_count = packed.bvh_node_count
        s.materials = arr(packed.materials, np.float32)
        s.material_count = packed.material_count
        s.spheres = arr(packed.spheres, np.float32)
        s.sphere_mat = arr(packed.sphere_materials, np.int32)
        s.sphere_light = arr(packed.sphere_lights, np.int32)
        s.sphere_count = packed.sphere_count
        s.lights = arr(packed.lights, np.float32)
        s.light_count = packed.light_count
        s.ground_material = packed.ground_material_active
        s.ground_height = packed.ground_height
        sun, env = scene.sun, scene.env
        el, az = math.radians(sun.elevation), math.radians(sun.azimuth)
        sd = [math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)]
        ca = math.cos(math.radians(sun.angular_radius))
        solid = 2 * math.pi * (1 - ca)
        s.sun_enabled = 1 if (env.map is None and sun.intensity > 0 and sun.elevation > -sun.angular_radius) else 0
        s.sun_dir[:] = sd
        s.sun_radiance[:] = [c * sun.intensity / solid for c in sun.color]
        s.sun_cos = ca
        s.sun_pdf = 1 / solid
        if env.map is not None:
            w, h, rgb = env.map
            rgb = np.ascontiguousarray(rgb, np.float32)
            rot, intensity = math.radians(env.rotation), env.intensity
        else:
            # The procedural sky, baked in world orientation at unit intensity.
            rgb = bake_sky(scene.sky, sun)
            h, w = rgb.shape[:2]
            rot, intensity = 0.0, 1.0
        self.env_rgb = rgb
        table, gw, gh, total = env_sampler(rgb, w, h, 1024)
        s.env_rgb = arr(rgb, np.float32)
        s.env_w, s.env_h = w, h
        s.env_table = arr(table, np.float32)
        s.grid_w, s.grid_h = gw, gh
        s.env_cos, s.env_sin = math.cos(rot), math.sin(rot)
        s.env_intensity = intensity
        s.env_sampling = 1 if total > 0 else 0
        s.env_background = 1 if env.background else 0
        s.backdrop[:] = list(env.backdrop)
        s.sheen_lut = arr(sheen_lut_half(), np.float32)
        size, color, data = packing.texture_layers(scene, packed)
        if size:
            s.color_maps = arr(color, np.uint32)
            s.data_maps = arr(data, np.uint32)
            s.tex_size = size
            s.color_layers = color.size // (size * size)
            s.data_layers = data.size // (size * size)
        self.textures = (size, color, data)
        self.scene = s
        self.keep = keep
        self.env_table = table

    def env_pdf_batch(self, dirs) -> np.ndarray:
        """envPdf of many world directions (n, 3)."""
        l = _pt_lib()
        l.or_env_pdf_batch.argtypes = [C.POINTER(OrScene), _U32, _P, _P]
        l.or_env_pdf_batch.restype = None
        d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
        out = np.zeros(d.shape[0])
        l.or_env_pdf_batch(C.byref(self.scene), d.shape[0], _ptr(d), _ptr(out))
        return out

    def env_pdf(self, d) -> float:
        """envPdf of world direction d (lights.glsl:49-56)."""
        l = _pt_lib()
        l.or_env_pdf_dir.argtypes = [C.POINTER(OrScene), _P]
        l.or_env_pdf_dir.restype = _D
        dd = np.ascontiguousarray(d, np.float64)
        return l.or_env_pdf_dir(C.byref(self.scene), _ptr(dd))

    def frame(self, camera, width, height, max_bounces=8, clamp=0.0) -> OrFrame:
        f = OrFrame()
        f.width, f.height, f.max_bounces, f.clamp = width, height, max_bounces, clamp
        fw, r, u, t = camera_basis(camera.position, camera.target, camera.fov)
        # The reference uploads these as float32 uniforms.
        f.cam_pos[:] = list(np.float32(camera.position))
        f.cam_fwd[:] = list(np.float32(fw))
        f.cam_right[:] = list(np.float32(r))
        f.cam_up[:] = list(np.float32(u))
        f.tan_half_fov = t
        f.aperture = camera.aperture
        f.focus_dist = camera.focus_distance
        return f

    COUNT_NAMES = ["paths", "closest_rays", "closest_interior_visits", "closest_tri_tests", "shadow_rays",
                   "shadow_node_visits", "shadow_tri_tests", "vertices", "env_nee", "escaped", "light_nee"]

    def trace_pixels(self, frame: OrFrame, seed0: int, samples: int, xy: np.ndarray, per_sample: bool = False,
                     threads: int | None = None, counts: bool = False):
        xy = np.ascontiguousarray(xy, np.uint32).reshape(-1, 2)
        n = xy.shape[0]
        out = np.zeros((n, 12))
        ps = np.zeros((n, samples, 3)) if per_sample else None
        threads = threads or min(16, os.cpu_count() or 1)
        chunks = _split(n, threads * 4)
        cnt = np.zeros((max(1, len(chunks)), 16))

        def run(arg):
            i, (a, b) = arg
            return _pt_lib().or_trace_pixels(C.byref(self.scene), C.byref(frame), seed0, samples,
                                             _ptr(xy) + a * 8, b - a, _ptr(out) + a * 96,
                                             (_ptr(ps) + a * samples * 24) if per_sample else None,
                                             (_ptr(cnt) + i * 128) if counts else None)

        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(run, enumerate(chunks)))
        if counts:
            tot = cnt.sum(0)
            return out, ps, dict(zip(self.COUNT_NAMES, tot[: len(self.COUNT_NAMES)].tolist()))
        return out, ps

    def trace_reprojected(self, frame: OrFrame, prev_film: dict, prev_frame: OrFrame, history_cap: float, seed: int,
                          xy: np.ndarray, threads: int | None = None):
        """trace.frag main() with uReproject = 1 (trace.frag:274-355) at the listed
        pixels: one sample with seed `seed`, history reprojected from prev_film
        (float arrays color/albedo/normal (H, W, 4) and history (H, W), rows
        bottom-up) seen through prev_frame's camera.  Returns (out (n, 13): colour
        rgb + lum^2, albedo + coverage, normal + depth, history; reprojected (n,)
        bool)."""
        xy = np.ascontiguousarray(xy, np.uint32).reshape(-1, 2)
        n = xy.shape[0]
        arrs = [np.ascontiguousarray(prev_film[k], np.float32) for k in ("color", "albedo", "normal", "history")]
        pv = OrPrev()
        pv.width, pv.height = prev_frame.width, prev_frame.height
        for k in ("cam_pos", "cam_right", "cam_up", "cam_fwd"):
            getattr(pv, k)[:] = list(getattr(prev_frame, k))
        pv.tan_half_fov = prev_frame.tan_half_fov
        pv.history_cap = history_cap
        pv.color4, pv.albedo4, pv.normal4, pv.history = [a.ctypes.data for a in arrs]
        out = np.zeros((n, 13))
        flags = np.zeros(n, np.uint8)
        threads = threads or min(16, os.cpu_count() or 1)
        chunks = _split(n, threads * 4)

        def run(ab):
            a, b = ab
            return _pt_lib().or_trace_reprojected(C.byref(self.scene), C.byref(frame), C.byref(pv), seed,
                                                  _ptr(xy) + a * 8, b - a, _ptr(out) + a * 104, _ptr(flags) + a)

        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(run, chunks))
        return out, flags.astype(bool)

    def set_materials(self, packed_materials) -> None:
        """Replaces the material texels the oracle scene points to (same count)."""
        m = np.ascontiguousarray(packed_materials, np.float32).reshape(-1)
        self.keep.append(m)
        self.scene.materials = m.ctypes.data

    def set_env_texels(self, rgb) -> None:
        e = np.ascontiguousarray(rgb, np.float32).reshape(-1)
        self.keep.append(e)
        self.scene.env_rgb = e.ctypes.data

    def adjoint_pixels(self, frame: OrFrame, seed0: int, samples: int, xy, dl3, material_count: int,
                       threads: int | None = None):
        """FP64 path-replay gradient of sum_p dl3[p] · mean radiance[p] over the listed pixels:
        (material_count x 5 array, env (env_h, env_w, 3) array)."""
        l = _pt_lib()
        l.or_adjoint_pixels.argtypes = [C.POINTER(OrScene), C.POINTER(OrFrame), _U32, _U32, _P, _P, _U32, _P, _P]
        l.or_adjoint_pixels.restype = C.c_int
        xy = np.ascontiguousarray(xy, np.uint32).reshape(-1, 2)
        dl = np.ascontiguousarray(dl3, np.float64).reshape(-1, 3)
        n = xy.shape[0]
        threads = threads or min(16, os.cpu_count() or 1)
        chunks = _split(n, threads)
        envn = self.scene.env_w * self.scene.env_h * 3
        gm = np.zeros((len(chunks), material_count * 5))
        ge = np.zeros((len(chunks), envn))

        def run(arg):
            i, (a, b) = arg
            return l.or_adjoint_pixels(C.byref(self.scene), C.byref(frame), seed0, samples, _ptr(xy) + a * 8,
                                       _ptr(dl) + a * 24, b - a, _ptr(gm) + i * gm.shape[1] * 8,
                                       _ptr(ge) + i * envn * 8)

        with cf.ThreadPoolExecutor(threads) as ex:
            list(ex.map(run, enumerate(chunks)))
        return gm.sum(0).reshape(material_count, 5), ge.sum(0).reshape(self.scene.env_h, self.scene.env_w, 3)


# ---------------------------------------------------------------------------
# Denoiser oracle (denoise_oracle.c)
# ---------------------------------------------------------------------------
def denoise(color, albedo, normal, history, cam13, sigma_l=4.0, sigma_n=64.0, sigma_p=1.5, iterations=5):
    """FP64 variance + à-trous over a film read back as float arrays (H, W, 4)."""
    l = lib()
    if not getattr(l, "_dn_ready", False):
        l.or_denoise.argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, _P, _D, _D, _D, C.c_int, _P, _P]
        l.or_denoise.restype = C.c_int
        l._dn_ready = True
    h, w = color.shape[:2]
    arrs = [np.ascontiguousarray(x, np.float32) for x in (color, albedo, normal, history)]
    cam = np.ascontiguousarray(cam13, np.float32)
    out = np.zeros((h, w, 4))
    remod = np.zeros((h, w, 3))
    l.or_denoise(*map(_ptr, arrs), w, h, _ptr(cam), sigma_l, sigma_n, sigma_p, iterations, _ptr(out), _ptr(remod))
    return out, remod


def denoise_window(color, albedo, normal, history, cam13, sigma_l, sigma_n, sigma_p, iterations, x0, y0, w, h):
    """The oracle filter over the window [x0, x0+w) x [y0, y0+h) of a full film
    (H, W, 4 arrays, rows bottom-up); pixels outside the window are treated as
    outside the image, camera rays use full-frame coordinates."""
    l = lib()
    l.or_denoise_window.argtypes = [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _D, _D,
                                    _D, C.c_int, _P, _P]
    l.or_denoise_window.restype = C.c_int
    H, W = color.shape[:2]
    arrs = [np.ascontiguousarray(x[y0:y0 + h, x0:x0 + w], np.float32) for x in (color, albedo, normal, history)]
    cam = np.ascontiguousarray(cam13, np.float32)
    out = np.zeros((h, w, 4))
    remod = np.zeros((h, w, 3))
    l.or_denoise_window(*map(_ptr, arrs), w, h, x0, y0, W, H, _ptr(cam), sigma_l, sigma_n, sigma_p, iterations,
                        _ptr(out), _ptr(remod))
    return out, remod


def camera13(camera) -> np.ndarray:
    fw, r, u, t = camera_basis(camera.position, camera.target, camera.fov)
    return np.concatenate([np.float32(camera.position), np.float32(r), np.float32(u), np.float32(fw),
                           [np.float32(t)]]).astype(np.float32)


def display(color, albedo, denoised, view_w, view_h, denoise_weight=0.0, exposure=1.0, tonemap=0):
    """FP64 restatement of display.frag (display.frag:16-67) over a film read
    back as float arrays (H, W, 4), rows bottom-up; denoised is the denoised
    illumination (H, W, >=3) or None.  Returns (rgba8 (view_h, view_w, 4)
    uint8 as a UNORM8 target stores the colour: clamp, x255, round half to
    even; the float colour before that conversion (view_h, view_w, 3))."""
    color = np.asarray(color, np.float64)
    albedo = np.asarray(albedo, np.float64)
    rh, rw = color.shape[:2]

    def radiance(px, py):
        px = np.clip(px, 0, rw - 1)
        py = np.clip(py, 0, rh - 1)
        raw = color[py, px, :3]
        if denoise_weight > 0:
            a = albedo[py, px]
            d = np.asarray(denoised, np.float64)[py, px, :3]
            mixed = raw * (1 - denoise_weight) + d * np.maximum(a[..., :3], 1e-3) * denoise_weight
            return np.where((a[..., 3] >= 0.5)[..., None], mixed, raw)
        return raw

    yy, xx = np.mgrid[0:view_h, 0:view_w]
    if rw == view_w and rh == view_h:
        c = radiance(xx, yy)
    else:
        sx = (xx + 0.5) * (rw / view_w) - 0.5
        sy = (yy + 0.5) * (rh / view_h) - 0.5
        ix, iy = np.floor(sx).astype(np.int64), np.floor(sy).astype(np.int64)
        fx, fy = sx - np.floor(sx), sy - np.floor(sy)

        def cr(f):
            return [f * (-0.5 + f * (1 - 0.5 * f)), 1 + f * f * (-2.5 + 1.5 * f), f * (0.5 + f * (2 - 1.5 * f)),
                    f * f * (-0.5 + 0.5 * f)]

        wx, wy = cr(fx), cr(fy)
        c = np.zeros((view_h, view_w, 3))
        for j in range(4):
            for i in range(4):
                c += radiance(ix + i - 1, iy + j - 1) * (wx[i] * wy[j])[..., None]
        c = np.maximum(c, 0)
    c = c * exposure
    if tonemap == 0:
        c = np.clip((c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14), 0, 1)
    elif tonemap == 1:
        c = c / (1 + c)
    c = np.clip(c, 0, 1)
    c = np.where(c >= 0.0031308, 1.055 * np.power(c, 1 / 2.4) - 0.055, c * 12.92)
    px, py = xx + 0.5, yy + 0.5
    a = 0.06711056 * px + 0.00583715 * py
    b = 52.9829189 * (a - np.floor(a))
    c = c + ((b - np.floor(b)) - 0.5)[..., None] / 255
    rgba = np.empty((view_h, view_w, 4), np.uint8)
    rgba[..., :3] = np.rint(np.clip(c, 0, 1) * 255).astype(np.uint8)
    rgba[..., 3] = 255
    return rgba, c
