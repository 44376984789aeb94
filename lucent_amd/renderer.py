"""Progressive renderer — the host mirror of Renderer (src/renderer.ts:171-871):
load_scene / upload_lighting / upload_camera (with temporal reprojection) /
reset / trace_samples / run_denoiser / render / screenshot / benchmark /
read-back, with the GPU-time budget scheduler (perf.ts) behind render().
The moving preview's automatic resolution drop is not part of it; a
caller that renders a preview film reprojects from it with
lucent_trace_reprojected.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native
from .perf import GpuBudget, affordable_samples
from .runtime import Context, Scene as DeviceScene
from .scene import PackedScene, Scene, pack_scene, texture_arrays


@dataclass
class RenderSettings:
    """Integrator part of RenderSettings (renderer.ts:42-84); the reference's
    default firefly clamp is 8, the benchmark configurations use 0 (unbiased).
    per_sample_launch selects the reference's launch structure (one pass over
    the film per sample) instead of the persistent tile kernel;
    inline_shadows makes the tile kernel trace each shadow ray where the
    integrator asks for it instead of in the lane's common traversal loop.
    All structures give bit-identical films."""
    max_bounces: int = 8
    firefly_clamp: float = 8.0
    per_sample_launch: bool = False
    waves_per_cu: int = 0
    occupancy: int = 0
    wavefront: bool = False
    inline_shadows: bool = False
    # Shape of the persistent per-sample hand-out (0 = each film's tile sized
    # for job_paths paths per job (0 = 256), shade batch 32, 8 traversal
    # turns per check); autotune() picks them.
    tile_w: int = 0
    tile_h: int = 0
    shade_batch: int = 0
    trace_turns: int = 0
    job_paths: int = 0
    # Presentation part (renderer.ts:42-84 defaults).
    temporal: bool = True
    denoise: bool = True
    exposure: float = 1.0
    tonemap: str = "aces"
    # Scheduling part (renderer.ts:42-84, perf.ts): samples per render();
    # 0 lets the GPU-time budget decide.
    samples_per_frame: int = 0
    gpu_budget: float = 0.5
    fps_cap: float = 60.0
    max_samples: int = 0

    def struct(self) -> "_native.RenderSettingsStruct":
        flags = (_native.TRACE_PER_SAMPLE if self.per_sample_launch else 0) | \
                (_native.TRACE_WAVEFRONT if self.wavefront else 0) | \
                (_native.TRACE_INLINE_SHADOWS if self.inline_shadows else 0)
        return _native.RenderSettingsStruct(self.max_bounces, self.firefly_clamp, flags, self.waves_per_cu,
                                            self.occupancy, self.tile_w, self.tile_h, self.shade_batch,
                                            self.trace_turns, self.job_paths)


def lighting_struct(scene: Scene, keep: list) -> _native.LightingStruct:
    L = _native.LightingStruct()
    s = scene.sun
    L.sun.elevation, L.sun.azimuth = s.elevation, s.azimuth
    L.sun.color[:] = list(s.color)
    L.sun.intensity, L.sun.angular_radius = s.intensity, s.angular_radius
    k = scene.sky
    L.sky.enabled = 1 if k.enabled else 0
    L.sky.zenith[:], L.sky.horizon[:], L.sky.ground[:] = list(k.zenith), list(k.horizon), list(k.ground)
    L.sky.intensity = k.intensity
    e = scene.env
    if e.map is not None:
        w, h, data = e.map
        arr = np.ascontiguousarray(np.asarray(data, np.float32).reshape(-1))
        if arr.size != w * h * 3:
            raise ValueError("environment map size mismatch")
        keep.append(arr)
        L.env.map_rgb, L.env.width, L.env.height = arr.ctypes.data, w, h
    else:
        L.env.map_rgb, L.env.width, L.env.height = None, 0, 0
    L.env.rotation, L.env.intensity = e.rotation, e.intensity
    L.env.background = 1 if e.background else 0
    L.env.backdrop[:] = list(e.backdrop)
    return L


HISTORY_CAP = 8  # renderer.ts:139
MOTION_HOLD_MS = 160  # renderer.ts:133
TONEMAPS = {"aces": 0, "reinhard": 1, "none": 2}


def atrous_levels(spp: int) -> int:
    """atrousLevels (renderer.ts:142)."""
    return 5 if spp <= 4 else 4 if spp <= 16 else 3 if spp <= 64 else 2 if spp <= 256 else 1


def denoise_weight(spp: int) -> float:
    """denoiseWeight (renderer.ts:144): full until 128 spp, fading to the raw image by 1024."""
    return min(1.0, max(0.0, (1024 - spp) / (1024 - 128)))


def camera_struct(cam) -> _native.CameraStruct:
    c = _native.CameraStruct()
    c.position[:] = list(cam.position)
    c.target[:] = list(cam.target)
    c.fov, c.aperture, c.focus_distance = cam.fov, cam.aperture, cam.focus_distance
    return c


class Film:
    """The accumulation targets (renderer.ts:476-496); rows bottom to top."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        h = C.c_void_p()
        _native.check(_native.lib().lucent_film_create(ctx.handle, width, height, C.byref(h)))
        self._h = h
        self.width, self.height = width, height
        ctx.adopt(self)

    @property
    def handle(self):
        return self._h

    def reset(self) -> None:
        _native.check(_native.lib().lucent_film_reset(self.ctx.handle, self._h))

    def set_seed(self, seed: int) -> None:
        _native.check(_native.lib().lucent_film_set_seed(self._h, seed))

    def info(self) -> dict:
        w, h, n, s = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        _native.check(_native.lib().lucent_film_info(self._h, C.byref(w), C.byref(h), C.byref(n), C.byref(s)))
        return {"width": w.value, "height": h.value, "samples": n.value, "seed": s.value}

    def read(self) -> dict:
        n = self.width * self.height
        color = np.empty((self.height, self.width, 4), np.float32)
        albedo = np.empty((self.height, self.width, 4), np.float32)
        normal = np.empty((self.height, self.width, 4), np.float32)
        hist = np.empty((self.height, self.width), np.float32)
        _native.check(_native.lib().lucent_film_read(self.ctx.handle, self._h, color.ctypes.data, albedo.ctypes.data,
                                                     normal.ctypes.data, hist.ctypes.data))
        assert color.size == n * 4
        return {"color": color, "albedo": albedo, "normal": normal, "history": hist}

    def device_ptrs(self) -> dict:
        p = [C.c_void_p() for _ in range(4)]
        _native.check(_native.lib().lucent_film_device_ptrs(self._h, *[C.byref(x) for x in p]))
        return dict(zip(["color", "albedo", "normal", "history"], [x.value for x in p]))

    def destroy(self) -> None:
        if self._h is not None:
            _native.lib().lucent_film_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Renderer:
    def __init__(self, ctx: Context, width: int, height: int, settings: Optional[RenderSettings] = None):
        self.ctx = ctx
        self.settings = settings or RenderSettings()
        self.film = Film(ctx, width, height)
        self.scene: Optional[Scene] = None
        self.packed: Optional[PackedScene] = None
        self.device_scene: Optional[DeviceScene] = None
        self.history_camera = None     # the camera the film's samples were traced with
        self.reproject_from = None     # pending reprojection: the previous view's camera
        self.denoised_frame = -1
        self.budget = GpuBudget(self.settings.gpu_budget)
        self.ms_per_sample = 0.0
        self.last_sync_measure = -math.inf
        self.last_motion = -math.inf

    @property
    def frame(self) -> int:
        return self.film.info()["samples"]

    def load_scene(self, scene: Scene, bvh=None) -> None:
        """Renderer.loadScene (renderer.ts:279-312)."""
        self.packed = pack_scene(scene, bvh)
        if self.device_scene is not None:
            self.device_scene.destroy()
        self.device_scene = self.ctx.upload_scene(self.packed)
        self.scene = scene
        self.upload_textures(scene)
        self.upload_lighting(scene)

    def pick_distance(self, origin, direction) -> Optional[float]:
        """pickDistance (scene.ts:327-341): the nearest hit along the ray —
        triangles through the exact GPU query (intersectBVH's FP64 hit
        record), then the spheres with tMax = that distance, then the ground
        plane if it is enabled — or None."""
        o = np.asarray(origin, np.float64).reshape(1, 3)
        d = np.asarray(direction, np.float64).reshape(1, 3)
        best = math.inf
        if self.packed.tri_count > 0:
            hit = self.device_scene.intersect(o, d)
            if hit.tri[0] != 0xFFFFFFFF:
                best = float(hit.t[0])
        for s in self.scene.spheres:
            # intersectSphere (ray.ts:48-67), tMin 1e-5.
            oc = o[0] - np.asarray(s.center, np.float64)
            b = float(oc @ d[0])
            c = float(oc @ oc) - s.radius * s.radius
            disc = b * b - c
            if disc < 0:
                continue
            sq = math.sqrt(disc)
            t = -b - sq
            if t <= 1e-5 or t >= best:
                t = -b + sq
                if t <= 1e-5 or t >= best:
                    continue
            best = t
        g = self.scene.ground
        if g.enabled and abs(d[0, 1]) > 1e-9:
            t = (g.height - o[0, 1]) / d[0, 1]
            if 1e-5 < t < best:
                best = t
        return best if math.isfinite(best) else None

    def upload_textures(self, scene: Scene) -> None:
        """The texture arrays of loadScene (renderer.ts:295, 315-335)."""
        size, color, data = texture_arrays(scene, self.packed)
        self.texture_size = size
        _native.check(_native.lib().lucent_scene_set_textures(
            self.ctx.handle, self.device_scene.handle, size, color.ctypes.data if color.size else None,
            color.size // (size * size) if size else 0, data.ctypes.data if data.size else None,
            data.size // (size * size) if size else 0))

    def upload_lighting(self, scene: Scene) -> None:
        """Renderer.uploadLighting (renderer.ts:355-413)."""
        keep: list = []
        L = lighting_struct(scene, keep)
        _native.check(_native.lib().lucent_scene_set_lighting(self.ctx.handle, self.device_scene.handle, C.byref(L)))
        self.reset()

    def reset(self) -> None:
        """Renderer.reset (renderer.ts:463-474)."""
        self.film.reset()
        self.history_camera = None
        self.reproject_from = None
        self.denoised_frame = -1

    def upload_camera(self, camera, moving: bool = True) -> None:
        """Renderer.uploadCamera + invalidateView (renderer.ts:424-451): with
        temporal reuse and samples in the film, the next sample reprojects the
        history instead of starting over.  `moving` only matters to the
        reference's preview scaling."""
        self.scene.camera = camera
        if moving:
            self.last_motion = time.perf_counter() * 1e3
        if self.settings.temporal and self.history_camera is not None and self.frame > 0:
            self.reproject_from = self.history_camera
            self.denoised_frame = -1
        else:
            self.reset()

    def trace_samples(self, count: int, stream: Optional[int] = None) -> None:
        """Renderer.traceSamples (renderer.ts:702-737); the first sample after a
        camera move reprojects the history (uReproject = 1)."""
        if count <= 0:
            return
        cam = camera_struct(self.scene.camera)
        st = self.settings.struct()
        lib = _native.lib()
        if self.reproject_from is not None:
            prev = camera_struct(self.reproject_from)
            _native.check(lib.lucent_trace_reprojected(self.ctx.handle, self.device_scene.handle, self.film.handle,
                                                       C.byref(cam), self.film.handle, C.byref(prev),
                                                       float(HISTORY_CAP), C.byref(st), stream))
            self.reproject_from = None
            count -= 1
        if count:
            _native.check(lib.lucent_trace_samples(self.ctx.handle, self.device_scene.handle, self.film.handle,
                                                   C.byref(cam), C.byref(st), count, stream))
        self.history_camera = self.scene.camera

    def display_params(self) -> "_native.DisplayParamsStruct":
        spp = self.frame
        w = denoise_weight(spp) if (self.settings.denoise and spp > 0 and self.denoised_frame == spp) else 0.0
        return _native.DisplayParamsStruct(w, self.settings.exposure, TONEMAPS[self.settings.tonemap])

    def _measure_sample(self) -> float:
        """measureSample (renderer.ts:664-670): one sample timed between two waits."""
        self.ctx.synchronize()
        t0 = time.perf_counter()
        self.trace_samples(1)
        self.ctx.synchronize()
        return (time.perf_counter() - t0) * 1e3

    def trace_budgeted(self, now_ms: float) -> int:
        """traceBudgeted (renderer.ts:618-649): spends this frame's share of the
        GPU-time budget on samples; the sample cost is re-measured every 2 s."""
        st = self.settings
        remaining = st.max_samples - self.frame if st.max_samples > 0 else math.inf
        if st.samples_per_frame > 0:
            n = int(min(st.samples_per_frame, remaining))
            self.trace_samples(n)
            return n
        self.budget.budget = st.gpu_budget
        frame_ms = 1000.0 / (st.fps_cap if st.fps_cap > 0 else 60.0)
        if self.ms_per_sample <= 0 or now_ms - self.last_sync_measure > 2000:
            ms = self._measure_sample()
            self.ms_per_sample = self.ms_per_sample * 0.6 + ms * 0.4 if self.ms_per_sample > 0 else ms
            self.last_sync_measure = now_ms
            self.budget.allowance(now_ms, 0)
            self.budget.spend(ms)
            return 1
        allowance = self.budget.allowance(now_ms, max(frame_ms * st.gpu_budget * 2, self.ms_per_sample * 1.05))
        # Moving the camera always gets a sample; the budget is repaid once it stops.
        interactive = now_ms - self.last_motion < MOTION_HOLD_MS
        n = int(min(1 if interactive else affordable_samples(allowance, self.ms_per_sample), remaining))
        if n <= 0:
            return 0
        self.trace_samples(n)
        self.budget.spend(n * self.ms_per_sample)
        return n

    def render(self, samples: Optional[int] = None, stream: Optional[int] = None,
               now_ms: Optional[float] = None) -> int:
        """Renderer.render (renderer.ts:552-600): traces `samples` samples, or
        with samples None what the settings' samples_per_frame or GPU-time
        budget pay for (nothing once max_samples is reached), runs the
        denoiser when its weight is above 0 and the film changed, and leaves
        the image to display()/screenshot().  Returns the samples traced."""
        if samples is None:
            done = self.settings.max_samples > 0 and self.frame >= self.settings.max_samples
            samples = 0 if done else self.trace_budgeted(time.perf_counter() * 1e3 if now_ms is None else now_ms)
        else:
            self.trace_samples(samples, stream)
        spp = self.frame
        if self.settings.denoise and spp > 0 and denoise_weight(spp) > 0 and self.denoised_frame != spp:
            self.run_denoiser(stream=stream)
            self.denoised_frame = spp
        return samples

    def display(self, d_rgba8: int, view_w: Optional[int] = None, view_h: Optional[int] = None,
                stream: Optional[int] = None) -> None:
        """The display draw (display.frag) into device memory d_rgba8."""
        p = self.display_params()
        _native.check(_native.lib().lucent_display(self.ctx.handle, self.film.handle, C.byref(p),
                                                   view_w or self.film.width, view_h or self.film.height, d_rgba8,
                                                   stream))

    def screenshot(self, view_w: Optional[int] = None, view_h: Optional[int] = None) -> np.ndarray:
        """Renderer.screenshot (renderer.ts:856-862): the displayed image as
        RGBA8 (view_h, view_w, 4), rows bottom-up like the film."""
        vw, vh = view_w or self.film.width, view_h or self.film.height
        out = np.empty((vh, vw, 4), np.uint8)
        p = self.display_params()
        _native.check(_native.lib().lucent_screenshot(self.ctx.handle, self.film.handle, C.byref(p), vw, vh,
                                                      out.ctypes.data))
        return out

    def benchmark(self, samples: int) -> dict:
        """Renderer.benchmark (renderer.ts:745-777): 1 warm-up sample, then
        `samples` timed samples; Mpaths/s = samples * W * H / time."""
        self.reset()
        self.trace_samples(1)
        self.ctx.synchronize()
        t0 = time.perf_counter()
        self.trace_samples(samples)
        self.ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        w, h = self.film.width, self.film.height
        return {"traceMs": ms, "mpathsPerSec": samples * w * h / (ms / 1e3) / 1e6, "width": w, "height": h}

    def run_denoiser(self, sigma_l: float = 4.0, sigma_n: float = 64.0, sigma_p: float = 1.5, iterations: int = 0,
                     stream: Optional[int] = None) -> None:
        """Renderer.runDenoiser (renderer.ts:788-830); iterations 0 = atrousLevels(frame)."""
        cam = camera_struct(self.scene.camera)
        p = _native.DenoiseParamsStruct(sigma_l, sigma_n, sigma_p, iterations)
        _native.check(_native.lib().lucent_denoise(self.ctx.handle, self.film.handle, C.byref(cam), C.byref(p), stream))

    def read_denoised(self) -> dict:
        w, h = self.film.width, self.film.height
        iv = np.empty((h, w, 4), np.float32)
        rad = np.empty((h, w, 4), np.float32)
        _native.check(_native.lib().lucent_film_read_denoised(self.ctx.handle, self.film.handle, iv.ctypes.data,
                                                              rad.ctypes.data))
        return {"illum_var": iv, "radiance": rad}

    def close(self) -> None:
        self.film.destroy()
        if self.device_scene is not None:
            self.device_scene.destroy()


class FrameGraph:
    """One progressive frame captured into a HIP graph (lucent_frame_graph_*):
    `samples` samples of the renderer's film, then optionally the denoiser and
    the display pass into device memory d_rgba8.  launch() replays it; n
    replays equal n eager frames bit for bit.  Camera, settings and filter
    parameters are fixed at capture."""

    def __init__(self, renderer: "Renderer", samples: int, denoise: Optional[tuple] = None,
                 display: Optional[tuple] = None, d_rgba8: Optional[int] = None, view: Optional[tuple] = None):
        r = renderer
        self.renderer = r
        cam = camera_struct(r.scene.camera)
        st = r.settings.struct()
        dn = _native.DenoiseParamsStruct(*denoise) if denoise is not None else None
        dp = _native.DisplayParamsStruct(*display) if display is not None else None
        vw, vh = view or (r.film.width, r.film.height)
        h = C.c_void_p()
        _native.check(_native.lib().lucent_frame_graph_create(
            r.ctx.handle, r.device_scene.handle, r.film.handle, C.byref(cam), C.byref(st), samples,
            C.byref(dn) if dn is not None else None, C.byref(dp) if dp is not None else None, vw, vh, d_rgba8,
            C.byref(h)))
        self._h = h
        self.samples = samples
        r.ctx.adopt(self)

    def launch(self, stream: Optional[int] = None) -> None:
        _native.check(_native.lib().lucent_frame_graph_launch(self._h, stream))
        self.renderer.history_camera = self.renderer.scene.camera

    def close(self) -> None:
        if self._h is not None:
            _native.lib().lucent_frame_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_items(ctx: Context, jobs):
    arr = (_native.BatchItemStruct * max(1, len(jobs)))()
    for i, (r, n) in enumerate(jobs):
        if r.ctx is not ctx:
            raise ValueError("every renderer of a batch must use the batch's context")
        arr[i].scene = r.device_scene.handle
        arr[i].film = r.film.handle
        arr[i].camera = camera_struct(r.scene.camera)
        arr[i].samples = int(n)
    return arr


def autotune(ctx: Context, jobs, settings: RenderSettings, trial_samples: int = 2) -> dict:
    """lucent_autotune_launch: times the persistent launch's occupancy
    variants, then its job tile, shade batch and traversal turns, on `jobs`
    (a list of (Renderer, samples)), and sets them in `settings`.  Returns
    {"occupancy", "tile", "job_paths", "shade_batch", "trace_turns",
    "trials": [...]}; a tile of [0, 0] is sized per film for job_paths paths
    per job."""
    st = settings.struct()
    trials = (_native.AutotuneTrialStruct * 64)()
    n = C.c_uint32()
    _native.check(_native.lib().lucent_autotune_launch(ctx.handle, _batch_items(ctx, jobs), len(jobs), C.byref(st),
                                                       trial_samples, trials, 64, C.byref(n)))
    settings.occupancy, settings.tile_w, settings.tile_h = st.occupancy, st.tile_w, st.tile_h
    settings.shade_batch, settings.trace_turns = st.shade_batch, st.trace_turns
    settings.job_paths = st.job_paths
    return {"occupancy": st.occupancy, "tile": [st.tile_w, st.tile_h], "job_paths": st.job_paths,
            "shade_batch": st.shade_batch, "trace_turns": st.trace_turns,
            "trials": [{"occupancy": t.occupancy, "tile": [t.tile_w, t.tile_h], "job_paths": t.job_paths,
                        "shade_batch": t.shade_batch, "trace_turns": t.trace_turns, "ms": round(t.ms, 3)}
                       for t in trials[:n.value]]}


def autotune_occupancy(ctx: Context, jobs, settings: RenderSettings, trial_samples: int = 2) -> dict:
    """lucent_autotune_tiles: the occupancy variant alone (sets
    settings.occupancy).  Returns {"occupancy": n, "ms": {variant: ms}}."""
    best = C.c_uint32()
    ms = (C.c_float * 5)()
    st = settings.struct()
    _native.check(_native.lib().lucent_autotune_tiles(ctx.handle, _batch_items(ctx, jobs), len(jobs), C.byref(st),
                                                      trial_samples, C.byref(best), ms))
    settings.occupancy = best.value
    return {"occupancy": best.value, "ms": {k + 2: float(ms[k]) for k in range(5)}}


def trace_batch(ctx: Context, jobs, settings: RenderSettings, stream: Optional[int] = None) -> None:
    """lucent_trace_batch: `jobs` is a list of (Renderer, samples); every
    renderer's film gets `samples` more samples, all in one persistent launch.
    The renderers must share `ctx`; their own settings are ignored in favour
    of `settings`."""
    st = settings.struct()
    _native.check(_native.lib().lucent_trace_batch(ctx.handle, _batch_items(ctx, jobs), len(jobs), C.byref(st), stream))


class BatchGraph:
    """lucent_batch_graph_*: one trace_batch over `jobs` (a list of
    (Renderer, samples)) captured into a HIP graph with its own device
    buffers; launch() replays it, each replay adding every job's samples to
    its film, bit-identical to the same number of trace_batch calls.
    Cameras and settings are fixed at capture."""

    def __init__(self, ctx: Context, jobs, settings: RenderSettings):
        st = settings.struct()
        h = C.c_void_p()
        _native.check(_native.lib().lucent_batch_graph_create(ctx.handle, _batch_items(ctx, jobs), len(jobs),
                                                              C.byref(st), C.byref(h)))
        self._h = h
        self.jobs = list(jobs)
        ctx.adopt(self)

    def launch(self, stream: Optional[int] = None) -> None:
        _native.check(_native.lib().lucent_batch_graph_launch(self._h, stream))

    def close(self) -> None:
        if self._h is not None:
            _native.lib().lucent_batch_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
