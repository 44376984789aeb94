"""Device context, uploaded scenes and batched ray queries.

``Context`` replaces the WebGL2 context behind Renderer (src/renderer.ts:226);
``Scene`` is an uploaded PackedScene (Renderer.loadScene, renderer.ts:279);
``Scene.intersect`` is intersectBVH (src/math/bvh.ts:265) over a batch of rays.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native
from ._native import (LUCENT_MISS, LUCENT_PREC_EXACT_FP64, LUCENT_PREC_GLSL_FP32, LUCENT_QUERY_ANY,
                      LUCENT_QUERY_CLOSEST, LucentError)
from .bvh import BVH

PRECISIONS = {"exact": LUCENT_PREC_EXACT_FP64, "glsl": LUCENT_PREC_GLSL_FP32}


class Context:
    """One device and one HIP stream (not re-entrant)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _native.check(_native.lib().lucent_create(device, C.byref(h)))
        self._h = h
        # Scenes, films and frame graphs made on this context: the C ABI
        # wants them destroyed before the context, so close() does that for
        # any still alive.
        self._children = weakref.WeakSet()

    def adopt(self, child) -> None:
        """Registers an object with a destroy()/close() method that close() must run first."""
        self._children.add(child)

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise LucentError(-1, "context is closed")
        return self._h

    def device_info(self) -> dict:
        cus = C.c_int32()
        mem = C.c_int64()
        _native.check(_native.lib().lucent_device_info(self.handle, C.byref(cus), C.byref(mem)))
        return {"compute_units": cus.value, "global_mem_bytes": mem.value}

    def synchronize(self) -> None:
        _native.check(_native.lib().lucent_synchronize(self.handle))

    def set_tree_build(self, max_leaf: int = 2, bins: int = 32) -> None:
        """lucent_set_tree_build: the device tree of scenes uploaded from now
        on (max_leaf 0 keeps the caller's BVH as the device tree)."""
        _native.check(_native.lib().lucent_set_tree_build(self.handle, max_leaf, bins))

    def set_record_budget(self, nbytes: int = 0) -> None:
        """lucent_set_record_budget: path-record bytes one single-film tile
        launch may use (0: the default, 4.5 GiB)."""
        _native.check(_native.lib().lucent_set_record_budget(self.handle, nbytes))

    def env_sampler_gpu(self, rgba, max_width: int = 1024, parallel: bool = False):
        """lucent_env_sampler_gpu: buildEnvSampler in float64 on the device
        from `rgba`, a (height, width, 4) float32 torch tensor on this
        context's device (with `parallel`, lucent_env_sampler_gpu_parallel:
        the alias pairs from FP64 prefix sums).  Returns (table (cells, 4)
        float32 tensor, grid_w, grid_h, total, device milliseconds)."""
        import torch
        if rgba.dtype != torch.float32 or rgba.dim() != 3 or rgba.shape[2] != 4 or not rgba.is_contiguous():
            raise ValueError("rgba must be a contiguous (height, width, 4) float32 tensor")
        h, w = int(rgba.shape[0]), int(rgba.shape[1])
        gw, gh, total, ms = C.c_uint32(), C.c_uint32(), C.c_double(), C.c_float()
        cells = min(w, max_width) * h  # upper bound on the grid
        table = torch.empty((max(cells, 1), 4), dtype=torch.float32, device=rgba.device)
        fn = _native.lib().lucent_env_sampler_gpu_parallel if parallel else _native.lib().lucent_env_sampler_gpu
        _native.check(fn(self.handle, C.c_void_p(rgba.data_ptr()), w, h, max_width, C.c_void_p(table.data_ptr()),
                         C.byref(gw), C.byref(gh), C.byref(total), C.byref(ms), None))
        return table[:gw.value * gh.value], gw.value, gh.value, total.value, ms.value

    def close(self) -> None:
        if self._h is not None:
            for child in list(self._children):
                (getattr(child, "close", None) or child.destroy)()
            _native.lib().lucent_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def scene_from_bvh(self, positions, bvh: BVH) -> "Scene":
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float32).reshape(-1))
        if pos.size % 9:
            raise ValueError("positions length is not a multiple of 9")
        tri_count = pos.size // 9
        if bvh.tri_indices.size != tri_count:
            raise ValueError(f"BVH covers {bvh.tri_indices.size} triangles; positions hold {tri_count}")
        nodes = np.ascontiguousarray(bvh.nodes, dtype=np.float32)
        tri = np.ascontiguousarray(bvh.tri_indices, dtype=np.uint32)
        h = C.c_void_p()
        _native.check(_native.lib().lucent_scene_from_bvh(
            self.handle, pos.ctypes.data, tri_count, nodes.ctypes.data, bvh.node_count, tri.ctypes.data,
            C.byref(h)))
        return Scene(self, h)

    def upload_scene(self, packed) -> "Scene":
        """Uploads a ``lucent_amd.scene.PackedScene`` (scene.ts:106-134)."""
        keep = []

        def ptr(a, dtype):
            if a is None:
                return None
            arr = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))
            keep.append(arr)
            return arr.ctypes.data

        s = _native.PackedSceneStruct()
        s.tri_geometry = ptr(packed.tri_geometry, np.float32)
        s.tri_attributes = ptr(packed.tri_attributes, np.float32)
        s.tri_count = packed.tri_count
        s.bvh_nodes = ptr(packed.bvh_nodes, np.float32)
        s.bvh_node_count = packed.bvh_node_count
        s.tri_indices = ptr(packed.tri_indices, np.uint32)
        s.materials = ptr(packed.materials, np.float32)
        s.material_count = packed.material_count
        s.spheres = ptr(packed.spheres, np.float32)
        s.sphere_materials = ptr(packed.sphere_materials, np.int32)
        s.sphere_lights = ptr(packed.sphere_lights, np.int32)
        s.sphere_count = packed.sphere_count
        s.lights = ptr(packed.lights, np.float32)
        s.light_count = packed.light_count
        s.ground_material = packed.ground_material_active
        s.ground_height = packed.ground_height
        h = C.c_void_p()
        _native.check(_native.lib().lucent_scene_upload(self.handle, C.byref(s), C.byref(h)))
        return Scene(self, h)


@dataclass
class Hits:
    """Batched BVHHit (bvh.ts:254-259); tri == LUCENT_MISS marks the reference's null."""

    t: np.ndarray
    tri: np.ndarray
    u: np.ndarray
    v: np.ndarray

    @property
    def hit(self) -> np.ndarray:
        return self.tri != LUCENT_MISS


class Scene:
    def __init__(self, ctx: Context, handle: C.c_void_p):
        self.ctx = ctx
        self._h = handle
        ctx.adopt(self)

    @property
    def handle(self):
        if self._h is None:
            raise LucentError(-1, "scene is destroyed")
        return self._h

    def stats(self) -> dict:
        st = _native.SceneStatsStruct()
        _native.check(_native.lib().lucent_scene_stats_get(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def rebuild_env_sampler(self, on_device: int = 0) -> float:
        """lucent_scene_rebuild_env_sampler: the importance sampler rebuilt
        from the scene's current environment texels, by the host build (the
        map copied back; on_device 0), the FP64 device build (same table
        bits; 1) or the parallel FP64 device build (same distribution, other
        alias pairs; 2); returns the luminance-weighted total (0 switches
        light sampling off)."""
        total = C.c_double()
        _native.check(_native.lib().lucent_scene_rebuild_env_sampler(self.ctx.handle, self.handle, int(on_device),
                                                                     C.byref(total)))
        return total.value

    def destroy(self) -> None:
        if self._h is not None:
            _native.lib().lucent_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    @staticmethod
    def _flags(precision: str, any_hit: bool) -> int:
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
        return PRECISIONS[precision] | (LUCENT_QUERY_ANY if any_hit else LUCENT_QUERY_CLOSEST)

    def intersect(self, origins, dirs, tmin: float = 1e-5, tmax: float = float("inf"),
                  precision: str = "exact", any_hit: bool = False) -> Hits:
        """intersectBVH for a batch of rays held in host arrays (float32 or float64 xyz)."""
        o = np.asarray(origins)
        d = np.asarray(dirs)
        f64 = o.dtype == np.float64 or d.dtype == np.float64
        dt = np.float64 if f64 else np.float32
        o = np.ascontiguousarray(o, dtype=dt).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=dt).reshape(-1, 3)
        if o.shape != d.shape:
            raise ValueError("origins and dirs differ in shape")
        n = o.shape[0]
        t = np.empty(n, np.float64)
        u = np.empty(n, np.float64)
        v = np.empty(n, np.float64)
        tri = np.empty(n, np.uint32)
        fn = _native.lib().lucent_intersect_f64 if f64 else _native.lib().lucent_intersect
        _native.check(fn(self.ctx.handle, self.handle, o.ctypes.data, d.ctypes.data, n, tmin, tmax,
                         self._flags(precision, any_hit), t.ctypes.data, tri.ctypes.data, u.ctypes.data,
                         v.ctypes.data))
        return Hits(t, tri, u, v)

    def intersect_device(self, org, dirs, t, tri, u, v, tmin: float = 1e-5, tmax: float = float("inf"),
                         precision: str = "exact", any_hit: bool = False, stream: Optional[int] = None,
                         noalloc: bool = False, flags: int = 0) -> None:
        """Device-pointer form: arguments are torch CUDA tensors (or raw int addresses).
        noalloc=True calls the variant that allocates no device memory."""
        def addr(x):
            return x if isinstance(x, int) else x.data_ptr()
        n = (org.numel() // 3) if hasattr(org, "numel") else None
        if n is None:
            raise ValueError("pass torch tensors for the device form")
        fn = _native.lib().lucent_intersect_device_noalloc if noalloc else _native.lib().lucent_intersect_device
        _native.check(fn(self.ctx.handle, self.handle, addr(org), addr(dirs), n, tmin, tmax,
                         self._flags(precision, any_hit) | flags, addr(t), addr(tri), addr(u), addr(v), stream))
