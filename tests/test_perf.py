"""Performance controls (lucent_amd.perf) against the reference's scheduler
tests (test/perf.test.ts): the token bucket holds the average GPU load near
the budget for cheap and expensive samples, never banks a long stall,
affords whole samples only; presets round-trip and get cheaper from max to
eco; recommend picks the tier from the 1080p cost and lowers resolution in
quarter steps."""
from dataclasses import replace

import pytest

from lucent_amd.perf import (PERFORMANCE_PRESETS, GpuBudget, affordable_samples, match_preset, pick_performance,
                             recommend)


def simulate(budget, ms_per_sample, seconds=10):
    b = GpuBudget(budget)
    frame = 1000 / 60
    busy, now = 0.0, 0.0
    while now < seconds * 1000:
        allowance = b.allowance(now, max(frame * budget * 2, ms_per_sample * 1.05))
        n = affordable_samples(allowance, ms_per_sample)
        b.spend(n * ms_per_sample)
        busy += n * ms_per_sample
        now += frame
    return busy / (seconds * 1000)


@pytest.mark.parametrize("budget", [0.1, 0.25, 0.5, 0.8])
@pytest.mark.parametrize("cost", [0.5, 4, 16, 40, 120])
def test_average_load_follows_the_budget(budget, cost):
    assert abs(simulate(budget, cost) - budget) < 0.05  # toBeCloseTo(budget, 1)


def test_whole_gpu_at_full_budget():
    assert simulate(1, 3) > 0.95


def test_no_credit_across_a_long_stall():
    b = GpuBudget(0.5)
    b.allowance(0, 20)
    assert b.allowance(10_000, 20) <= 20


def test_whole_samples_only():
    assert affordable_samples(9.9, 5) == 1
    assert affordable_samples(4.9, 5) == 0
    assert affordable_samples(1000, 1) == 32
    assert affordable_samples(10, 0) == 0


def test_presets_round_trip_and_order():
    for name, p in PERFORMANCE_PRESETS.items():
        assert match_preset(p) == name
    assert match_preset(replace(PERFORMANCE_PRESETS["balanced"], gpu_budget=0.55)) == "custom"
    assert pick_performance(PERFORMANCE_PRESETS["eco"]) == PERFORMANCE_PRESETS["eco"]
    order = ["max", "high", "balanced", "eco"]
    for a, b in zip(order, order[1:]):
        pa, pb = PERFORMANCE_PRESETS[a], PERFORMANCE_PRESETS[b]
        assert pb.gpu_budget < pa.gpu_budget and pb.resolution_scale <= pa.resolution_scale
        assert pb.max_bounces < pa.max_bounces


def test_recommend_tiers_and_resolution():
    assert recommend(10, 1920, 1080).preset == "high"
    assert recommend(60, 1920, 1080).preset == "balanced"
    assert recommend(200, 1920, 1080).preset == "eco"
    assert abs(recommend(40, 3840, 2160).ms_at_1080p - 10) < 1e-9
    r = recommend(80, 2560, 1440)
    assert r.preset == "balanced" and r.settings.resolution_scale == 0.5  # sqrt(25/80) = 0.56 -> 0.5
    assert r.settings.gpu_budget == PERFORMANCE_PRESETS["balanced"].gpu_budget
    assert recommend(5000, 1920, 1080).settings.resolution_scale == 0.25
    assert recommend(1, 1920, 1080).settings.resolution_scale == 1
    assert recommend(100, 1280, 720).settings.resolution_scale <= 0.5


@pytest.mark.gpu
def test_budgeted_render_measures_then_spends_the_allowance(ctx):
    """render() with no sample count: the first call measures one sample, later
    calls trace what the allowance pays for, camera motion forces one sample,
    and max_samples stops accumulation."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent))
    from test_pathtrace_gpu import feature_scene
    from lucent_amd.renderer import Renderer, RenderSettings

    r = Renderer(ctx, 64, 48, RenderSettings(max_bounces=4, firefly_clamp=0.0, gpu_budget=1.0, max_samples=40))
    r.load_scene(feature_scene())
    assert r.render(now_ms=0.0) == 1 and r.frame == 1 and r.ms_per_sample > 0
    r.ms_per_sample = 2.0  # a fixed cost and a clean bucket make the allowance arithmetic exact
    r.budget.credit = 0.0
    got = r.render(now_ms=100.0)  # 100 ms of credit at budget 1, capped by the burst max(2 * 16.7, 2.1)
    assert got == 16 and r.frame == 17
    r.upload_camera(r.scene.camera, moving=True)
    base = r.last_motion
    r.last_sync_measure = base
    # Moving: one sample whatever the credit; the reprojected history keeps min(17, 8) samples.
    assert r.render(now_ms=base + 10) == 1 and r.frame == 9
    t = base + 1000
    while r.frame < 40:
        t += 100
        r.render(now_ms=t)
    assert r.frame == 40 and r.render(now_ms=t + 1e6) == 0
    r.close()
