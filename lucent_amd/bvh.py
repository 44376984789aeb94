"""BVH construction — the host mirror of src/math/bvh.ts.

``build_bvh`` is buildBVH (bvh.ts:52): the same binned SAH, the same node
layout ``[min.xyz, leftFirst, max.xyz, count]`` (bvh.ts:6-16) and the same
leaf order, computed by the C++ builder inside liblucent_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native

NODE_STRIDE = 8


@dataclass
class BVH:
    """Mirror of the reference's ``BVH`` interface (bvh.ts:18-24)."""

    nodes: np.ndarray  # float32, node_count * 8
    node_count: int
    tri_indices: np.ndarray  # uint32, slot -> original triangle
    bounds: tuple  # ((minx, miny, minz), (maxx, maxy, maxz))

    def depth(self) -> int:
        """Depth of the deepest leaf (root = 0)."""
        depth = np.zeros(self.node_count, dtype=np.int64)
        nodes = self.nodes.reshape(-1, NODE_STRIDE)
        for n in range(self.node_count):
            if nodes[n, 7] == 0 and self.node_count > 1:
                l = int(nodes[n, 3])
                depth[l] = depth[l + 1] = depth[n] + 1
        return int(depth.max()) if self.node_count else 0


def _as_f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def build_bvh(positions, max_leaf_size: int = 4, bins: int = 12) -> BVH:
    """buildBVH(positions, {maxLeafSize, bins}) — bvh.ts:52-252.

    Raises ValueError when the length is not a multiple of 9, as the
    reference throws (bvh.ts:55-57).
    """
    pos = _as_f32(positions).reshape(-1)
    if pos.size % 9 != 0:
        raise ValueError(f"positions length {pos.size} is not a multiple of 9")
    tri_count = pos.size // 9
    nodes = np.zeros(max(1, 2 * tri_count - 1) * NODE_STRIDE, dtype=np.float32)
    tri = np.zeros(tri_count, dtype=np.uint32)
    bounds = np.zeros(6, dtype=np.float64)
    count = C.c_uint32(0)
    _native.check(
        _native.lib().lucent_build_bvh(
            pos.ctypes.data, tri_count, max_leaf_size, bins, nodes.ctypes.data, C.byref(count),
            tri.ctypes.data, bounds.ctypes.data,
        )
    )
    n = count.value
    return BVH(nodes[: n * NODE_STRIDE].copy(), n, tri, (tuple(bounds[:3]), tuple(bounds[3:])))


def build_bvh_gpu(ctx, positions, max_leaf_size: int = 4, radius: int = 0) -> BVH:
    """lucent_build_bvh_gpu: a linear BVH built on the context's GPU (Morton
    order, radix tree), in buildBVH's layout (bvh.ts:6-16).  Same consumers,
    higher traversal cost, far faster build for large meshes.  radius > 0:
    lucent_build_bvh_gpu_clustered, clusters in Morton order merged by
    surface area within that many neighbours (about the SAH tree's cost)."""
    pos = _as_f32(positions).reshape(-1)
    if pos.size % 9 != 0:
        raise ValueError(f"positions length {pos.size} is not a multiple of 9")
    tri_count = pos.size // 9
    nodes = np.zeros(max(1, 2 * tri_count - 1) * NODE_STRIDE, dtype=np.float32)
    tri = np.zeros(tri_count, dtype=np.uint32)
    bounds = np.zeros(6, dtype=np.float64)
    count = C.c_uint32(0)
    if radius > 0:
        _native.check(_native.lib().lucent_build_bvh_gpu_clustered(
            ctx.handle, pos.ctypes.data, tri_count, max_leaf_size, radius, nodes.ctypes.data, C.byref(count),
            tri.ctypes.data, bounds.ctypes.data))
    else:
        _native.check(
            _native.lib().lucent_build_bvh_gpu(
                ctx.handle, pos.ctypes.data, tri_count, max_leaf_size, nodes.ctypes.data, C.byref(count),
                tri.ctypes.data, bounds.ctypes.data,
            )
        )
    n = count.value
    return BVH(nodes[: n * NODE_STRIDE].copy(), n, tri, (tuple(bounds[:3]), tuple(bounds[3:])))


def build_bvh_parallel(positions, max_leaf_size: int = 4, bins: int = 12, threads: int = 0) -> BVH:
    """lucent_build_bvh_parallel: buildBVH's splits with subtrees built on
    `threads` host threads (0 = all); the same nodes, numbered subtree by
    subtree."""
    pos = _as_f32(positions).reshape(-1)
    if pos.size % 9 != 0:
        raise ValueError(f"positions length {pos.size} is not a multiple of 9")
    tri_count = pos.size // 9
    nodes = np.zeros(max(1, 2 * tri_count - 1) * NODE_STRIDE, dtype=np.float32)
    tri = np.zeros(tri_count, dtype=np.uint32)
    bounds = np.zeros(6, dtype=np.float64)
    count = C.c_uint32(0)
    _native.check(
        _native.lib().lucent_build_bvh_parallel(
            pos.ctypes.data, tri_count, max_leaf_size, bins, threads, nodes.ctypes.data, C.byref(count),
            tri.ctypes.data, bounds.ctypes.data,
        )
    )
    n = count.value
    return BVH(nodes[: n * NODE_STRIDE].copy(), n, tri, (tuple(bounds[:3]), tuple(bounds[3:])))
