"""The no-allocation closest-hit query (lucent_intersect_device_noalloc)
against lucent_intersect_device, through the C ABI.

Two calls in a row on the same device (the second with fewer rays, so it
runs over what the first left in the shared scratch array), each checked
bit for bit against the allocating entry point and against the FP64 oracle;
then the forced-spill variant (a 2-entry LDS ring), so the traversal stacks
really live in the scratch array.
"""
import numpy as np
import pytest

import oracle as O
from lucent_amd import workloads
from lucent_amd.bvh import build_bvh

pytestmark = pytest.mark.gpu
TINY_LDS_STACK = 0x100


def run(sc, o, d, noalloc, flags=0):
    import torch
    dev = torch.device("cuda", 0)
    og = torch.from_numpy(np.ascontiguousarray(o, np.float32)).to(dev)
    dg = torch.from_numpy(np.ascontiguousarray(d, np.float32)).to(dev)
    n = o.shape[0]
    # Outputs start as a non-miss pattern so that an unwritten record shows.
    t = torch.full((n,), -7.0, dtype=torch.float64, device=dev)
    tri = torch.full((n,), 12345, dtype=torch.int32, device=dev)
    u = torch.full((n,), -7.0, dtype=torch.float64, device=dev)
    v = torch.full((n,), -7.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    sc.intersect_device(og, dg, t, tri, u, v, noalloc=noalloc, flags=flags)
    sc.ctx.synchronize()
    return (t.cpu().numpy(), tri.cpu().numpy().view(np.uint32), u.cpu().numpy(), v.cpu().numpy())


def same(a, b):
    return all(np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)) for x, y in zip(a, b))


def test_noalloc_matches_allocating_entry_point(ctx):
    w = workloads.config1(1 << 18, seed=9)
    b = build_bvh(w["positions"])
    sc = ctx.scene_from_bvh(w["positions"], b)
    for n in (1 << 18, 3001):
        o, d = w["origins"][:n], w["dirs"][:n]
        print(f"noalloc closest hit, icosphere {w['positions'].size // 9} tris, {n} rays", flush=True)
        ref = run(sc, o, d, noalloc=False)
        got = run(sc, o, d, noalloc=True)
        assert same(got, ref), f"{n} rays: no-allocation records differ from lucent_intersect_device"
        t, tri, u, v = O.intersect_bvh_batch(b, w["positions"], o, d)
        assert np.array_equal(got[1], tri) and np.array_equal(got[0], t)
        assert np.array_equal(got[2], u) and np.array_equal(got[3], v)
        assert (got[1] != 0xFFFFFFFF).mean() > 0.99


def test_noalloc_spill_path(ctx):
    pos, _ = O.torus_knot()
    b = build_bvh(pos)
    sc = ctx.scene_from_bvh(pos, b)
    o, d = workloads.sphere_rays(20000, 7)
    o = o * np.float32(1.5)
    for n in (20000, 777):
        print(f"noalloc forced spill, torus knot, {n} rays", flush=True)
        ref = run(sc, o[:n], d[:n], noalloc=False, flags=TINY_LDS_STACK)
        got = run(sc, o[:n], d[:n], noalloc=True, flags=TINY_LDS_STACK)
        assert same(got, ref), f"{n} rays: spilled no-allocation records differ"
        t, tri, u, v = O.intersect_bvh_batch(b, pos, o[:n], d[:n])
        assert np.array_equal(got[1], tri) and np.array_equal(got[0], t)
