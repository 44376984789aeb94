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
