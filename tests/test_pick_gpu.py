"""pickDistance (scene.ts:327-341) on the product side: the reference's
ground-plane case (test/scene.test.ts:169-180) and the oracle's pickDistance
on rays that hit triangles, spheres and the ground in every order."""
import numpy as np
import pytest

from oracle import packing as OP
from lucent_amd import scene as S
from lucent_amd.renderer import Renderer, RenderSettings

from test_packing import base_scene, textured_scene

pytestmark = pytest.mark.gpu


def test_ground_pick(ctx):
    s = base_scene()
    s.ground = S.Ground(True, -1.0, S.matte((0.3, 0.3, 0.3)))
    s.env = S.Environment(map=(4, 2, np.ones(24, np.float32)))
    r = Renderer(ctx, 8, 8, RenderSettings(8, 0.0))
    r.load_scene(s)
    assert r.pick_distance((0, 1, 0), (0, -1, 0)) == pytest.approx(2)
    assert r.pick_distance((0, 1, 0), (0, 1, 0)) is None
    s.ground.enabled = False  # the shared ground object is read at pick time
    assert r.pick_distance((0, 1, 0), (0, -1, 0)) is None
    r.close()


def test_pick_matches_oracle(ctx):
    sc = textured_scene()
    sc.env = S.Environment(map=(4, 2, np.ones(24, np.float32)))
    r = Renderer(ctx, 8, 8, RenderSettings(8, 0.0))
    r.load_scene(sc)
    p = OP.pack_scene(sc)
    rng = np.random.default_rng(9)
    n = 0
    for _ in range(300):
        o = rng.uniform(-4, 4, 3) + np.array([0, 2.0, 0])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a = r.pick_distance(o, d)
        b = OP.pick_distance(p, o, d)
        assert (a is None) == (b is None)
        if a is not None:
            assert a == b
            n += 1
    assert n > 100
    r.close()
