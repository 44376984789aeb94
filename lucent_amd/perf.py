"""Performance controls — the host mirror of src/perf.ts:1-122: presets that
bundle the GPU-time budget with the other cost knobs, a recommendation from
a measured scan, and the token bucket that spends wall-clock time's share of
GPU work on samples (Renderer.traceBudgeted, renderer.ts:618-649)."""
from __future__ import annotations

import math
from dataclasses import asdict, dataclass, replace
from typing import Optional


@dataclass(frozen=True)
class PerformanceSettings:
    gpu_budget: float          # share of wall-clock time the GPU may spend rendering, 0.05-1
    fps_cap: float             # redraws per second at most; 0 follows the display (60)
    resolution_scale: float    # internal resolution as a fraction of the view
    max_bounces: int
    max_samples: int           # stop accumulating at this many samples per pixel; 0 never stops
    pause_when_unfocused: bool


# perf.ts:24-29
PERFORMANCE_PRESETS = {
    "eco": PerformanceSettings(0.25, 30, 0.5, 4, 1024, True),
    "balanced": PerformanceSettings(0.5, 60, 0.75, 6, 2048, True),
    "high": PerformanceSettings(0.8, 60, 1.0, 8, 4096, True),
    "max": PerformanceSettings(1.0, 0, 1.0, 12, 0, False),
}
PRESET_ORDER = ("eco", "balanced", "high", "max")


def pick_performance(s) -> PerformanceSettings:
    """pickPerformance (perf.ts:39-43): just the performance fields of a larger settings object."""
    return PerformanceSettings(**{k: getattr(s, k) for k in PerformanceSettings.__dataclass_fields__})


def match_preset(s: PerformanceSettings) -> str:
    """matchPreset (perf.ts:45-49): the preset these settings equal, or 'custom'."""
    for name in PRESET_ORDER:
        if asdict(PERFORMANCE_PRESETS[name]) == asdict(pick_performance(s)):
            return name
    return "custom"


@dataclass(frozen=True)
class Recommendation:
    preset: str
    settings: PerformanceSettings
    ms_at_1080p: float


def recommend(ms_per_sample: float, width: int, height: int) -> Recommendation:
    """recommend (perf.ts:71-81): the tier from the cost of a 1080p sample,
    then the resolution lowered in quarter steps (never above the tier's own,
    never below 0.25) until one sample fits the tier's frame-time target."""
    pixels = max(1, width * height)
    ms1080 = ms_per_sample * (1920 * 1080) / pixels
    preset = "high" if ms1080 <= 30 else "balanced" if ms1080 <= 90 else "eco"
    target = 20 if preset == "high" else 25 if preset == "balanced" else 33
    base = PERFORMANCE_PRESETS[preset]
    fit = math.sqrt(target / max(ms_per_sample, 1e-3))
    scale = min(base.resolution_scale, max(0.25, math.floor(fit * 4) / 4))
    return Recommendation(preset, replace(base, resolution_scale=scale), ms1080)


class GpuBudget:
    """Token bucket for GPU time (perf.ts:89-113): wall-clock time earns
    `budget` ms of GPU work per ms; work spends it; a gap longer than 250 ms
    is not banked and savings are capped at burst_ms."""

    def __init__(self, budget: float):
        self.budget = budget
        self.credit = 0.0
        self.last = -1.0

    def allowance(self, now_ms: float, burst_ms: float) -> float:
        dt = 0.0 if self.last < 0 else min(now_ms - self.last, 250.0)
        self.last = now_ms
        self.credit = min(self.credit + dt * self.budget, max(burst_ms, 1.0))
        return self.credit

    def spend(self, ms: float) -> None:
        self.credit -= ms

    def reset(self) -> None:
        self.credit = 0.0
        self.last = -1.0


def affordable_samples(allowance_ms: float, ms_per_sample: float, cap: int = 32) -> int:
    """affordableSamples (perf.ts:116-119): whole samples the allowance pays for."""
    if not (ms_per_sample > 0) or allowance_ms < ms_per_sample:
        return 0
    return min(cap, int(math.floor(allowance_ms / ms_per_sample)))
