"""TEST INFRASTRUCTURE — Python access to the FP64 parity oracle.

Loads ``oracle/liboracle.so`` (built by ``make -C oracle``) through ctypes.
Only the test suite, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg may import this package; the product never does.
"""
from __future__ import annotations

import concurrent.futures as cf
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
