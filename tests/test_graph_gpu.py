"""Frame graphs (lucent_frame_graph_*): a captured frame replayed n times
leaves the film, the denoised image and the display buffer bit-identical to
n eager frames.  Every replay is checked against the eager result, with the
program's own eager work and synchronisations between replays."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

from lucent_amd import _native
from lucent_amd.renderer import FrameGraph, Renderer, RenderSettings, camera_struct

pytestmark = pytest.mark.gpu

sys.path.insert(0, str(Path(__file__).parent))
from test_pathtrace_gpu import feature_scene  # noqa: E402

W, H = 80, 48


def pair(ctx, seed):
    out = []
    for _ in range(2):
        r = Renderer(ctx, W, H, RenderSettings(max_bounces=8, firefly_clamp=0.0))
        r.load_scene(feature_scene())
        r.film.set_seed(seed)
        r.reset()
        out.append(r)
    return out


@pytest.mark.parametrize("with_filters", [False, True])
def test_graph_replays_equal_eager_frames(ctx, with_filters):
    import torch

    g_r, e_r = pair(ctx, 21)
    dn = (4.0, 64.0, 1.5, 3) if with_filters else None
    dp = (1.0, 1.0, 0) if with_filters else None
    buf_g = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda")
    buf_e = torch.zeros((H, W, 4), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    graph = FrameGraph(g_r, 2, dn, dp, buf_g.data_ptr() if with_filters else None)
    lib = _native.lib()
    for rep in range(4):
        graph.launch()
        # Eager frame on the other renderer, with host work in between.
        e_r.trace_samples(2)
        if with_filters:
            e_r.run_denoiser(*dn)
            p = _native.DisplayParamsStruct(*dp)
            _native.check(lib.lucent_display(ctx.handle, e_r.film.handle, C.byref(p), W, H, buf_e.data_ptr(), None))
        ctx.synchronize()
        torch.cuda.synchronize()
        a, b = g_r.film.read(), e_r.film.read()
        for k in a:
            assert np.array_equal(a[k], b[k]), (rep, k)
        assert g_r.film.info() == e_r.film.info()
        if with_filters:
            da, db = g_r.read_denoised(), e_r.read_denoised()
            for k in da:
                assert np.array_equal(da[k], db[k]), (rep, k)
            assert torch.equal(buf_g, buf_e), rep
    graph.close()
    g_r.close()
    e_r.close()


def test_graph_refuses_a_film_advanced_elsewhere(ctx):
    r, _ = pair(ctx, 3)
    graph = FrameGraph(r, 1)
    graph.launch()
    r.trace_samples(1)
    st = _native.lib().lucent_frame_graph_launch(graph._h, None)
    assert st != 0 and b"outside this graph" in _native.lib().lucent_last_error(ctx.handle)
    graph.close()
    r.close()


def batch_set(ctx, seed):
    """Three films of two scenes at different sizes (one with an alpha-tested
    surface), seeded alike, for a batch graph and its eager twin."""
    from lucent_amd import workloads
    from test_pathtrace_gpu import alpha_scene
    scenes = [feature_scene(), workloads.spheres_scene(), alpha_scene()]
    sizes = [(64, 40), (33, 17), (40, 24)]
    rs = []
    for sc, (w, h) in zip(scenes, sizes):
        r = Renderer(ctx, w, h, RenderSettings(max_bounces=8, firefly_clamp=0.0))
        r.load_scene(sc)
        r.film.set_seed(seed)
        r.reset()
        rs.append(r)
    return rs


@pytest.mark.parametrize("budget", [0, 1])
def test_batch_graph_replays_equal_eager_batches(ctx, budget):
    """lucent_batch_graph_*: a batch captured once and replayed n times
    leaves every film bit-identical to n eager lucent_trace_batch calls
    (default record budget: the per-sample hand-out with record rings; a
    budget of 1 byte: whole pixels), with eager work and synchronisations
    between the replays; the films' sample counts and seeds advance."""
    import torch
    from lucent_amd.renderer import BatchGraph, trace_batch

    ctx.set_record_budget(budget)
    try:
        g_rs, e_rs = batch_set(ctx, 9), batch_set(ctx, 9)
        spps = [3, 1, 2]
        settings = RenderSettings(max_bounces=8, firefly_clamp=0.0)
        graph = BatchGraph(ctx, list(zip(g_rs, spps)), settings)
        for rep in range(3):
            print(f"batch graph replay {rep}, record budget {budget}", flush=True)
            graph.launch()
            trace_batch(ctx, list(zip(e_rs, spps)), settings)
            ctx.synchronize()
            torch.cuda.synchronize()
            for g, e, n in zip(g_rs, e_rs, spps):
                a, b = g.film.read(), e.film.read()
                for k in a:
                    assert np.array_equal(a[k], b[k]), (rep, k)
                assert g.film.info() == e.film.info()
                assert g.film.info()["samples"] == n * (rep + 1)
        graph.close()
        for r in g_rs + e_rs:
            r.close()
    finally:
        ctx.set_record_budget(0)


def test_batch_graph_refuses_a_film_advanced_elsewhere(ctx):
    from lucent_amd.renderer import BatchGraph
    rs = batch_set(ctx, 4)
    graph = BatchGraph(ctx, [(r, 1) for r in rs], RenderSettings(max_bounces=8, firefly_clamp=0.0))
    graph.launch()
    rs[1].trace_samples(1)
    st = _native.lib().lucent_batch_graph_launch(graph._h, None)
    assert st != 0 and b"outside this graph" in _native.lib().lucent_last_error(ctx.handle)
    with pytest.raises(_native.LucentError, match="persistent batch launch"):
        BatchGraph(ctx, [(r, 1) for r in rs], RenderSettings(max_bounces=8, firefly_clamp=0.0, per_sample_launch=True))
    graph.close()
    for r in rs:
        r.close()
