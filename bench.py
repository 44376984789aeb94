"""Benchmark of the MI355X path-tracing hot path (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload closest-hit]

Headline (default): the plain progressive path trace at 1440x900 — one step
is a full 64-spp frame of the 222k-triangle knot scene (Renderer.traceSamples(64)
after a reset: one persistent k_trace_units launch; the reference's launch
structure, one launch per sample, is timed beside it on the same film);
value = camera paths (pixel samples) per second.  The closest-hit benchmark is reported beside it: one step = one
batched exact closest-hit launch of 2^20 rays against the 81,920-triangle
fBm icosphere (inputs resident in HBM).

Timing: W untimed steps, then K steps bracketed by torch.cuda.synchronize();
each launch is also bracketed by HIP events on the stream it runs on (torch's
current stream, handed to the library), which gives the average launch
duration used for the roofline.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s closest-hit; samples/s @1440x900 batched vs plain; adjoint grad/s"
HBM_PEAK_GBS = 8000.0


def profiled_traffic(key: str):
    """HBM bytes per launch for `key` from the committed rocprofv3 --pmc summary, or None."""
    path = ROOT / "profiles/hbm_traffic.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + " "):
            for tok in line.split():
                if tok.startswith("hbm_bytes_per_launch="):
                    return float(tok.split("=", 1)[1])
    return None


def profiled_counters(profile: str, kernel: str):
    """VALU issue busy %, VALU lane utilisation % and L2 hit % of `kernel`'s row in a committed
    rocprofv3 counter summary (profiles/<profile>), or None."""
    path = ROOT / "profiles" / profile
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 15 and kernel in cells[0]:
            try:
                return {"valu_issue_busy_pct": float(cells[7]), "valu_lanes_active_pct": float(cells[8]),
                        "l2_hit_pct": float(cells[12])}
            except ValueError:
                return None
    return None


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_closest_hit(bvh, positions, origins, dirs, sample: int, workers: int = 0) -> dict:
    """Times the JavaScript port of intersectBVH (oracle/js) with Node worker_threads on a
    contiguous sample of the benchmark rays; checks its hit ids against the GPU's."""
    workers = workers or int(os.environ.get("LUCENT_CPU_WORKERS", "0")) or min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as d:
        bvh.nodes.tofile(f"{d}/nodes.f32")
        bvh.tri_indices.tofile(f"{d}/tri_indices.u32")
        positions.tofile(f"{d}/positions.f32")
        origins[:sample].tofile(f"{d}/origins.f32")
        dirs[:sample].tofile(f"{d}/dirs.f32")
        r = subprocess.run(["node", str(ROOT / "oracle/js/cpu_baseline.js"), d, str(workers), "5"],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
        res = json.loads(r.stdout.strip().splitlines()[-1])
        ids = np.fromfile(f"{d}/ids.u32", np.uint32)
    res["ids"] = ids
    return res


def bench_closest_hit(args) -> dict:
    import torch

    from lucent_amd import workloads
    from lucent_amd.bvh import build_bvh
    from lucent_amd.runtime import Context

    log("generating the closest-hit workload (81,920 tris, 2^20 rays)")
    w = workloads.config1()
    pos, org, dirs = w["positions"], w["origins"], w["dirs"]
    n = org.shape[0]
    t0 = time.time()
    bvh = build_bvh(pos)
    build_s = time.time() - t0
    ctx = Context(0)
    scene = ctx.scene_from_bvh(pos, bvh)
    stats = scene.stats()
    dev = torch.device("cuda:0")
    d_org = torch.from_numpy(org).to(dev)
    d_dir = torch.from_numpy(dirs).to(dev)
    d_t = torch.empty(n, dtype=torch.float64, device=dev)
    d_u = torch.empty(n, dtype=torch.float64, device=dev)
    d_v = torch.empty(n, dtype=torch.float64, device=dev)
    d_id = torch.empty(n, dtype=torch.int32, device=dev)
    # A dedicated (non-null) torch stream: the library launches on it and the
    # HIP events below are recorded on it.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    torch.cuda.synchronize()

    def step():
        scene.intersect_device(d_org, d_dir, d_t, d_id, d_u, d_v, stream=stream.cuda_stream)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    avg_launch_s = float(np.mean(launch_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3
    rays_per_s = n * args.steps / wall

    ids = d_id.cpu().numpy().view(np.uint32)
    counts = json.loads((ROOT / "tests/golden/config1_counts.json").read_text())
    P, T = counts["P_interior_visits_per_ray"], counts["T_triangle_tests_per_ray"]
    hit_bytes = 8 + 4 + 8 + 8
    b_ray = 24 + hit_bytes + 32 + 64 * P + 48 * T
    achieved = b_ray * n / avg_launch_s / 1e9
    traffic = profiled_traffic("closest_hit_icosphere81920_2e20rays")
    # Compulsory bytes: each ray in and its hit out (40 B) plus the scene's
    # nodes and triangles read once (the reference layout: 32 B per node,
    # 48 B per triangle).
    compulsory = 40 * n + 32 * stats["node_count"] + 48 * (pos.size // 9)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_ray": round(b_ray, 1), "avg_launch_ms": round(avg_launch_s * 1e3, 4),
                "compulsory_bytes_per_launch": int(compulsory),
                "compulsory_gbs": round(compulsory / avg_launch_s / 1e9, 1)}

    cpu = None
    if not args.no_cpu_baseline:
        sample = min(n, args.cpu_sample)
        log(f"cpu baseline: JS intersectBVH port on {sample} rays")
        res = cpu_baseline_closest_hit(bvh, pos, org, dirs, sample)
        agree = float((res.pop("ids") == ids[:sample]).mean())
        cpu = {"value": round(res["mrays_per_s"], 4), "unit": "Mrays/s", "cores": res["workers"], "kind": "port",
               "sample": (f"first {sample} of the 2^20 benchmark rays, JS port of intersectBVH "
                          f"(Node {res['node']}, worker_threads={res['workers']}, host cpus={res['cpus']}), "
                          f"median of 5 after 1 warm-up; hit ids agree with GPU on {agree:.6f} of rays"),
               "median_s": round(res["median_s"], 4)}
        # The same port on one worker, over a smaller contiguous sample.
        one_n = min(sample, 1 << 16)
        one = cpu_baseline_closest_hit(bvh, pos, org, dirs, one_n, workers=1)
        cpu["one_worker"] = {"value": round(one["mrays_per_s"], 4), "unit": "Mrays/s", "cores": 1,
                             "sample": f"first {one_n} rays, median of 5 after 1 warm-up",
                             "ids_agree_with_gpu": float((one.pop("ids") == ids[:one_n]).mean())}
        # Secondary figure: the FP64 C oracle (intersectBVH restated in C) on host threads.
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        O.intersect_bvh_batch(bvh, pos, org[:4096], dirs[:4096], threads=threads)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            _, otri, _, _ = O.intersect_bvh_batch(bvh, pos, org[:sample], dirs[:sample], threads=threads)
            times.append(time.perf_counter() - t0)
        cpu["c_oracle"] = {"value": round(sample / float(np.median(times)) / 1e6, 4), "unit": "Mrays/s",
                           "cores": threads, "kind": "port",
                           "sample": f"first {sample} rays, FP64 C oracle on {threads} threads, median of 3",
                           "ids_agree_with_gpu": float((otri == ids[:sample]).mean())}
    ctx_info = ctx.device_info()
    scene.destroy()
    ctx.close()
    return {
        "value": round(rays_per_s / 1e6, 2), "unit": "Mrays/s", "ms_per_step": round(ms_per_step, 4),
        "dtype": "f64", "config": {"workload": "closest-hit exact (FP64 hit test): 81,920-tri fBm icosphere, 2^20 "
                                               "rays from radius-3 sphere into the unit ball",
                                   "rays_per_step": n, "bvh_nodes": stats["node_count"],
                                   "bvh_depth": stats["depth"], "host_bvh_build_s": round(build_s, 3),
                                   "compute_units": ctx_info["compute_units"]},
        "roofline": roofline, "cpu_baseline": cpu,
        "hit_fraction": float((ids != 0xFFFFFFFF).mean()),
    }


def bench_path_trace(args) -> dict:
    """Progressive path trace: 1440x900, 64 spp per step, depth 8, RR from
    bounce 2, clamp 0.  The step is Renderer.traceSamples(64) after a reset,
    run by the persistent per-sample tile kernel (one k_trace_units launch
    per step); the reference's launch structure (one launch per sample) is
    timed beside it on the same film and seeds (the two films are
    bit-identical)."""
    import torch

    from lucent_amd import workloads
    from lucent_amd.renderer import Renderer, RenderSettings
    from lucent_amd.runtime import Context

    W, H, SPP = workloads.PT_WIDTH, workloads.PT_HEIGHT, workloads.PT_SPP
    log(f"building the path-trace scene ({W}x{H}, {SPP} spp)")
    scene = workloads.knot_scene()
    ctx = Context(0)
    r = Renderer(ctx, W, H, RenderSettings(max_bounces=workloads.PT_DEPTH, firefly_clamp=0.0))
    t0 = time.time()
    r.load_scene(scene)
    load_s = time.time() - t0
    stats = r.device_scene.stats()
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)

    def timed_steps(n, per_sample):
        r.settings.per_sample_launch = per_sample
        starts = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        ends = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        torch.cuda.synchronize()
        wall0 = time.perf_counter()
        for i in range(n):
            # The reset is ordered on the library's own stream: drain this one first.
            stream.synchronize()
            r.film.set_seed(0)
            r.reset()
            starts[i].record(stream)
            r.trace_samples(SPP, stream=stream.cuda_stream)
            ends[i].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - wall0
        return wall, [s_.elapsed_time(e) for s_, e in zip(starts, ends)]

    from lucent_amd.renderer import autotune
    tune = autotune(ctx, [(r, SPP)], r.settings, trial_samples=SPP)
    log(f"autotuned occupancy {tune['occupancy']}, tile {tune['tile']}, shade batch {tune['shade_batch']}, "
        f"trace turns {tune['trace_turns']}")
    log(f"warmup {args.warmup}")
    if args.warmup:
        timed_steps(args.warmup, False)
    log(f"timing {args.steps} steps")
    wall, trace_ms = timed_steps(args.steps, False)
    paths = W * H * SPP
    value = paths * args.steps / wall
    launch_s = float(np.mean(trace_ms)) / 1e3
    img = r.film.read()
    mean_rgb = [float(x) for x in img["color"][..., :3].reshape(-1, 3).mean(0)]
    log("timing the per-sample launch structure (1 step)")
    wall_ps, trace_ms_ps = timed_steps(1, True)
    img_ps = r.film.read()
    identical = all(np.array_equal(img[k], img_ps[k]) for k in img)
    r.settings.per_sample_launch = False
    counts = json.loads((ROOT / "tests/golden/pathtrace_counts.json").read_text())["per_path"]
    b_path = path_bytes(counts)
    # One launch accumulates SPP samples per pixel: the film (72 B per pixel)
    # is read and written once, everything else is per path.
    b_launch = (b_path - 72) * paths + 72 * W * H
    achieved = b_launch / launch_s / 1e9
    traffic = profiled_traffic("path_trace_1440x900")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "bytes_per_path": round(b_path - 72, 1),
                "film_bytes_per_pixel": 72, "kernel": "k_trace_units", "avg_launch_ms": round(launch_s * 1e3, 3),
                "occupancy": tune["occupancy"], "tile": tune["tile"], "shade_batch": tune["shade_batch"],
                "trace_turns": tune["trace_turns"],
                "note": "achieved counts every node and triangle fetch of the reference's binary-tree traversal "
                        "(per-path counts from the FP64 oracle); the scene's nodes and triangles stay resident in "
                        "L2 and the MALL, so this figure can exceed the HBM peak. Measured HBM traffic is "
                        "`traffic`; the kernel is bound by VALU issue (`counters`)"}
    if traffic:
        roofline["measured_hbm_gbs"] = round(traffic / launch_s / 1e9, 1)
        roofline["measured_hbm_frac"] = round(traffic / launch_s / 1e9 / HBM_PEAK_GBS, 4)
    roofline["counters"] = profiled_counters("path_trace_1440x900_64spp_counters.md", "k_trace_units")
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_path_trace(r.packed, scene, W, H)
    rays_per_path = counts["closest_rays"] + counts["shadow_rays"]
    r.close()
    ctx.close()
    per_sample_value = paths / (float(np.mean(trace_ms_ps)) / 1e3)
    return {"value": round(value, 1), "unit": "samples/s", "ms_per_step": round(wall / args.steps * 1e3, 3),
            "dtype": "f32",
            "config": {"workload": f"progressive path trace {W}x{H}, {SPP} spp, depth {workloads.PT_DEPTH}, "
                                   f"RR from bounce 2, clamp 0; 6 fBm torus knots + ground quad "
                                   f"({stats['tri_count']} tris), quad light, 1024x512 HDRI",
                       "width": W, "height": H, "spp": SPP, "tris": stats["tri_count"],
                       "scene_load_s": round(load_s, 3)},
            "roofline": roofline, "cpu_baseline": cpu, "image_mean_rgb": mean_rgb,
            "mpaths_per_s": round(value / 1e6, 2), "mrays_per_s_all_rays": round(value * rays_per_path / 1e6, 1),
            "autotune": tune,
            "per_sample_launches": {"samples_per_s": round(per_sample_value, 1),
                                    "trace_ms": round(float(np.mean(trace_ms_ps)), 3),
                                    "avg_sample_launch_ms": round(float(np.mean(trace_ms_ps)) / SPP, 4),
                                    "film_bit_identical": identical,
                                    "speedup_of_persistent_tiles": round(float(np.mean(trace_ms_ps)) / (launch_s * 1e3), 3)}}


def path_bytes(c: dict) -> float:
    """Algorithmic bytes of one camera path: accumulate read+write (72), per
    closest-hit ray 24 in + 16 out + 32 root + 64 per child pair + 48 per
    triangle, per shadow ray 24 in + 32 per node + 48 per triangle, per surface
    vertex 48 GEO + 64 ATTR + 160 material, per environment NEE 16 alias + 48
    bilinear, per escaped ray 48 bilinear."""
    return (72 + c["closest_rays"] * 72 + 64 * c["closest_interior_visits"] + 48 * c["closest_tri_tests"]
            + c["shadow_rays"] * 24 + 32 * c["shadow_node_visits"] + 48 * c["shadow_tri_tests"]
            + c["vertices"] * 272 + c["env_nee"] * 64 + c["escaped"] * 48)


def cpu_baseline_path_trace(packed, scene, W, H) -> dict:
    """The FP64 oracle's path tracer (a C restatement of trace.frag) on the
    host cores: a 256x256 crop at the frame centre, 16 spp, one thread per core."""
    import oracle as O

    o = O.PtOracle(scene)
    f = o.frame(scene.camera, W, H, 8, 0.0)
    yy, xx = np.mgrid[H // 2 - 128:H // 2 + 128, W // 2 - 128:W // 2 + 128]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1)
    threads = min(16, os.cpu_count() or 1)
    log(f"cpu baseline: FP64 oracle path tracer, 256x256 crop x 16 spp on {threads} threads")
    t0 = time.perf_counter()
    o.trace_pixels(f, 0, 16, xy, threads=threads)
    dt = time.perf_counter() - t0
    n = xy.shape[0] * 16
    return {"value": round(n / dt, 1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"256x256 centre crop of the 1440x900 frame, 16 spp ({n} paths), FP64 C restatement of "
                      f"trace.frag on {threads} threads, {dt:.2f} s wall"}


def bench_denoise(args) -> dict:
    """Mixed-precision denoise: a 1920x1080 film at 8 spp of the knot scene
    (FP32 radiance, FP16 guides), variance pass + 5 à-trous iterations with
    sigma_n 128, sigma_p 1 (for sigma_z), sigma_l 4; FP32 taps."""
    import torch

    from lucent_amd import workloads
    from lucent_amd.renderer import Renderer, RenderSettings
    from lucent_amd.runtime import Context

    W, H = 1920, 1080
    ctx = Context(0)
    r = Renderer(ctx, W, H, RenderSettings(8, 0.0))
    r.load_scene(workloads.knot_scene())
    r.trace_samples(8)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    torch.cuda.synchronize()
    # Steady state: the GPU's clocks settle over the first tens of
    # milliseconds of back-to-back denoises.
    for _ in range(40):
        r.run_denoiser(4.0, 128.0, 1.0, 5, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    reps = 100
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        r.run_denoiser(4.0, 128.0, 1.0, 5, stream=stream.cuda_stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    px = W * H
    b_px = 44 + 5 * 48  # variance pass + 5 iterations, compulsory bytes per pixel
    achieved = b_px * px / (ms / 1e3) / 1e9
    cpu = None
    error = None
    if not args.no_cpu_baseline:
        import oracle as O
        film = r.film.read()
        cam = O.camera13(r.scene.camera)
        t0 = time.perf_counter()
        ref, _ = O.denoise_window(film["color"], film["albedo"], film["normal"], film["history"], cam, 4.0, 128.0,
                                  1.0, 5, 720, 405, 480, 270)
        dt = time.perf_counter() - t0
        cpu = {"value": round(dt * 1e3 * 16, 1), "unit": "ms per 1920x1080 denoise (extrapolated)", "cores": 1,
               "kind": "port", "sample": f"FP64 oracle of the same filter on a 480x270 window (1/16 of the frame): "
                                         f"{dt * 1e3:.1f} ms, x16"}
        # Error of the GPU result (FP16 guides, FP32 taps) against the FP64
        # filter on the window's pixels beyond the filter's reach (62 px)
        # from its edges: relative to |ref| + 1e-3 of the mean.
        m = 62
        g = r.read_denoised()["illum_var"][405 + m:405 + 270 - m, 720 + m:720 + 480 - m, :3].astype(np.float64)
        rr = ref[m:270 - m, m:480 - m, :3]
        e = np.abs(g - rr) / (np.abs(rr) + 1e-3 * np.abs(rr).mean() + 1e-9)
        error = {"p99_9_rel": float(np.quantile(e, 0.999)), "rms_rel": float(np.sqrt((e ** 2).mean())),
                 "max_rel": float(e.max()), "pixels": int(e.size // 3),
                 "against": "FP64 oracle of the same filter on the same stored film (480x270 window interior)"}
    r.close()
    ctx.close()
    return {"metric": "denoise ms @1920x1080, 8 spp, 5 a-trous iterations", "value": round(ms, 4), "unit": "ms",
            "error_vs_fp64_oracle": error,
            "higher_is_better": False,
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_pixel": b_px,
                         "traffic": profiled_traffic("denoise_1920x1080")},
            "cpu_baseline": cpu}


def bench_adjoint(args) -> dict:
    """Differentiable rendering: 512x512, 32 spp, depth 6, 16 displaced spheres
    (20,480 tris), 16 materials, 256x128 env map.  One gradient evaluation =
    forward render (one persistent launch, per-sample radiance recorded) +
    dL/dI of the L2 loss against a target rendered with ±20% perturbed
    parameters + path-replay backward (one persistent launch + FP64
    reductions), all in device memory on one stream; grad/s = 1 / (forward +
    backward time).  HIP events on that stream time each pass; the
    roofline is the backward's (k_adjoint and its two reductions)."""
    import torch

    from lucent_amd import workloads
    from lucent_amd.diff import DiffRenderer
    from lucent_amd.runtime import Context

    W, H, SPP = workloads.DIFF_WIDTH, workloads.DIFF_HEIGHT, workloads.DIFF_SPP
    ctx = Context(0)
    sc = workloads.spheres_scene()
    d = DiffRenderer(ctx, sc, W, H, workloads.DIFF_DEPTH)
    from lucent_amd.renderer import autotune
    tune = autotune(ctx, [(d.renderer, SPP)], d.renderer.settings, trial_samples=SPP)
    log(f"forward occupancy {tune['occupancy']}")
    base_mats, base_env = list(d.materials), d.env.copy()
    pm, penv = workloads.perturb_params(base_mats, base_env, seed=5)
    d.set_materials(pm)
    d.set_env_texels(penv)
    target = d.forward(128, 1000, record=False, device=True).clone()
    d.set_materials(base_mats)
    d.set_env_texels(base_env)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream
    img = d.forward(SPP, 0, device=True, stream=sp)
    d.backward(2 * (img - target), SPP, 0, device=True, stream=sp)  # warm-up
    torch.cuda.synchronize()
    # The replay's job tile: the fastest of a few shapes (identical gradients
    # up to the order of the FP64 sums).
    tile_ms = {}
    for tile in ((8, 8), (4, 4), (4, 2), (2, 2)):
        d.adjoint_tile = tile
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        best = float("inf")
        for _ in range(2):
            e0.record(stream)
            d.backward(2 * (img - target), SPP, 0, device=True, stream=sp)
            e1.record(stream)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        tile_ms[f"{tile[0]}x{tile[1]}"] = round(best, 3)
    d.adjoint_tile = tuple(int(v) for v in min(tile_ms, key=tile_ms.get).split("x"))
    log(f"replay tile {d.adjoint_tile}: {tile_ms}")
    reps = max(3, min(args.steps, 10))
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(reps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for e in ev:
        e[0].record(stream)
        img = d.forward(SPP, 0, device=True, stream=sp)
        e[1].record(stream)
        gm, ge = d.backward(2 * (img - target), SPP, 0, device=True, stream=sp)
        e[2].record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - wall0) / reps
    t_fwd = float(np.mean([a.elapsed_time(b) for a, b, _ in ev])) / 1e3
    t_bwd = float(np.mean([b.elapsed_time(c) for _, b, c in ev])) / 1e3
    loss = float(((img[..., :3] - target[..., :3]).double() ** 2).sum())
    params = {"material_scalars": int(gm[:16].numel()), "env_channels": int(ge.numel())}
    counts = json.loads((ROOT / "tests/golden/adjoint_counts.json").read_text())["per_path"]
    b_path = path_bytes(counts) - 72
    paths, px = W * H * SPP, W * H
    env_lookups = counts["env_nee"] + counts["escaped"]
    # Replay: the path's fetches again, its recorded radiance (16 B), FP64
    # atomics of 4 texels x 3 channels per environment lookup; dL/dI (16 B
    # per pixel); the reduction of 8 env-gradient replicas into the output.
    b_bwd = paths * (b_path + 16 + 96 * env_lookups) + 16 * px + 9 * 8 * params["env_channels"] + 8 * 80
    achieved = b_bwd / t_bwd / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic("adjoint_512x512"),
                "kernel": "k_adjoint (+ k_reduce_materials, k_reduce_env)", "avg_launch_ms": round(t_bwd * 1e3, 3),
                "bytes_per_path": round(b_path + 16 + 96 * env_lookups, 1)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_adjoint(sc, W, H, SPP, workloads.DIFF_DEPTH)
    d.close()
    ctx.close()
    return {"metric": "adjoint grad/s (forward + backward, 512x512, 32 spp, depth 6)",
            "value": round(1.0 / (t_fwd + t_bwd), 3), "unit": "grad/s", "forward_ms": round(t_fwd * 1e3, 2),
            "backward_ms": round(t_bwd * 1e3, 2), "wall_ms_per_gradient": round(wall * 1e3, 2), "params": params,
            "loss": loss, "forward_launch": {k: v for k, v in tune.items() if k != "trials"},
            "replay_tile": list(d.adjoint_tile), "replay_tile_ms": tile_ms, "roofline": roofline, "cpu_baseline": cpu,
            "backward": "replays the forward pass's recorded per-sample radiance; image, loss gradient and "
                        "parameter gradients stay in device memory"}


def cpu_baseline_adjoint(sc, W, H, SPP, depth) -> dict:
    """The FP64 oracle's path-replay adjoint on 64 sampled pixels (all their
    samples), extrapolated to whole-image gradient evaluations per second."""
    import oracle as O

    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, depth, 0.0)
    n = 16384
    rng = np.random.default_rng(3)
    xy = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], -1).astype(np.uint32)
    dl = rng.standard_normal((n, 3))
    threads = min(16, os.cpu_count() or 1)
    log(f"cpu baseline: FP64 oracle adjoint, {n} pixels x {SPP} spp on {threads} threads")
    t0 = time.perf_counter()
    o.adjoint_pixels(f, 0, SPP, xy, dl, o.packed.material_count, threads=threads)
    dt = time.perf_counter() - t0
    per_image = dt * (W * H) / n
    return {"value": round(1.0 / per_image, 6), "unit": "grad/s", "cores": threads, "kind": "port",
            "sample": f"{n} random pixels x {SPP} spp of the 512x512 gradient (the 64-pixel check, scaled up "
                      f"to a measurable run), FP64 C path-replay adjoint on {threads} threads: {dt:.3f} s, "
                      f"extrapolated x{W * H / n:g} to one whole-image gradient"}


def bench_batch(args) -> dict:
    """Batched multi-scene tracing: 64 seeded scenes (1k..1.5M triangles,
    11.8M in all, 6% of meshes Worley alpha-masked), resolution uniform over
    256^2 / 640x480 / 1280x720 / 1920x1080, spp uniform in 1..64, depth 8.
    Batched: ONE lucent_trace_batch over all 64 films.  Two plain paths on
    the same films and seeds: one lucent_trace_samples(spp) launch per scene
    (the library's one-film persistent launch), and one launch per sample
    per scene (the reference's structure, renderer.ts:702-737).  Same seeds,
    same estimator: the films are compared bit for bit, so the per-pixel
    error is equal by construction."""
    import torch

    from lucent_amd import scene as S
    from lucent_amd import workloads
    from lucent_amd.bvh import build_bvh_gpu
    from lucent_amd.renderer import Renderer, RenderSettings, trace_batch
    from lucent_amd.runtime import Context

    log(f"building {workloads.BATCH_SCENES} batch scenes")
    t0 = time.time()
    ctx = Context(0)
    jobs, tris, gen_s, load_s = [], 0, 0.0, 0.0
    for i in range(workloads.BATCH_SCENES):
        g0 = time.time()
        sc, w, h, spp = workloads.batch_scene(i)
        g1 = time.time()
        r = Renderer(ctx, w, h, RenderSettings(max_bounces=workloads.BATCH_DEPTH, firefly_clamp=0.0))
        # The caller's tree comes from the GPU builder (the role of
        # buildBVHAsync); the upload rebuilds the device tree with the binned
        # SAH on the host's threads.
        r.load_scene(sc, build_bvh_gpu(ctx, S.merge_meshes([m.mesh for m in sc.meshes]).positions))
        load_s += time.time() - g1
        gen_s += g1 - g0
        tris += r.packed.tri_count
        r.packed = None  # host copies are not needed any more
        jobs.append((r, spp))
    setup_s = time.time() - t0
    paths = sum(r.film.width * r.film.height * n for r, n in jobs)
    pixels = sum(r.film.width * r.film.height for r, _ in jobs)
    log(f"{tris} triangles, {paths} paths; set-up {setup_s:.1f} s (scene generation {gen_s:.1f} s, "
        f"pack + BVH + upload {load_s:.1f} s)")
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream

    def reset_all():
        torch.cuda.synchronize()
        for r, _ in jobs:
            r.film.set_seed(0)
            r.reset()
        ctx.synchronize()

    settings = RenderSettings(max_bounces=workloads.BATCH_DEPTH, firefly_clamp=0.0)
    from lucent_amd.renderer import autotune
    tune = autotune(ctx, jobs, settings, trial_samples=8)
    log(f"autotuned {tune} (up to 8 spp of every scene, tiles sized for the real spp)")

    def run(mode, plain_settings=None):
        reset_all()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        e0.record(stream)
        if mode == "batched":
            trace_batch(ctx, jobs, settings, stream=sp)
        else:
            for r, n in jobs:
                r.settings = plain_settings
                r.trace_samples(n, stream=sp)
        e1.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t, e0.elapsed_time(e1) / 1e3

    def films_digest():
        out = []
        for r, _ in jobs:
            f = r.film.read()
            out.append(hash(f["color"].tobytes()) ^ hash(f["albedo"].tobytes()) ^ hash(f["normal"].tobytes()))
        return out

    if args.no_plain:
        run("batched")
        bt = [run("batched") for _ in range(max(1, min(args.steps, 3)))]
        for r, _ in jobs:
            r.close()
        ctx.close()
        tb_wall, tb_gpu = min(bt)
        return {"metric": "samples/s, 64-scene batch, one persistent launch (plain paths not run)",
                "value": round(paths / tb_wall, 1), "unit": "samples/s", "batched_s": round(tb_wall, 3),
                "batched_gpu_s": round(tb_gpu, 3), "autotune": tune, "paths": paths, "setup_s": round(setup_s, 1)}
    def plain(**kw):
        return RenderSettings(max_bounces=workloads.BATCH_DEPTH, firefly_clamp=0.0, **kw)

    log("plain: one launch per sample per scene")
    ps_wall, ps_gpu = run("per_sample", plain(per_sample_launch=True, occupancy=settings.occupancy))
    ref_digest = films_digest()
    log(f"per-sample launches {ps_wall:.2f} s; plain: one lucent_trace_samples launch per scene")
    # One launch per scene under several launch shapes: the batch's autotuned
    # shape, the library default (tiles sized for 256 paths per job) and
    # uniform tiles; the plain path is credited with its fastest.
    plain_shapes = {
        "batch_autotuned": plain(occupancy=settings.occupancy, tile_w=settings.tile_w, tile_h=settings.tile_h,
                                 job_paths=settings.job_paths, shade_batch=settings.shade_batch,
                                 trace_turns=settings.trace_turns),
        "default": plain(),
        "tiles_2x2": plain(tile_w=2, tile_h=2),
        "tiles_8x8": plain(tile_w=8, tile_h=8),
    }
    plain_s, same_plain = {}, True
    for name, ps in plain_shapes.items():
        run("per_scene", ps)
        pl = [run("per_scene", ps) for _ in range(2)]
        same_plain = same_plain and films_digest() == ref_digest
        plain_s[name] = min(pl)[0]
        log(f"one launch per scene, {name}: {plain_s[name]:.3f} s")
    best_plain = min(plain_s, key=plain_s.get)
    pl_wall = plain_s[best_plain]
    log("batched warm-up")
    run("batched")
    log("batched: timing")
    bt = [run("batched") for _ in range(max(1, min(args.steps, 3)))]
    same = films_digest() == ref_digest
    tb_wall, tb_gpu = min(bt)
    for r, _ in jobs:
        r.close()
    ctx.close()
    roofline = None
    cpath = ROOT / "tests/golden/batch_counts.json"
    if cpath.exists():
        per = {s["index"]: s for s in json.loads(cpath.read_text())["scenes"]}
        b_launch = 0.0
        for i in range(workloads.BATCH_SCENES):
            s_ = per[i]
            n_px = s_["width"] * s_["height"]
            b_launch += (path_bytes(s_["per_path"]) - 72) * n_px * s_["spp"] + 72 * n_px
        achieved = b_launch / tb_gpu / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic("batch_64scenes"),
                    "kernel": "k_trace_batch_units", "avg_launch_ms": round(tb_gpu * 1e3, 2),
                    "bytes_per_launch": round(b_launch)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_batch()
    return {"metric": "samples/s, 64-scene batch: one persistent launch vs one launch per scene",
            "value": round(paths / tb_wall, 1), "unit": "samples/s",
            "plain_one_launch_per_scene_samples_per_s": round(paths / pl_wall, 1),
            "speedup_vs_one_launch_per_scene": round(pl_wall / tb_wall, 3),
            "plain_fastest_shape": best_plain,
            "one_launch_per_scene_s_by_shape": {k: round(v, 3) for k, v in plain_s.items()},
            "speedup_vs_one_launch_per_scene_by_shape": {k: round(v / tb_wall, 3) for k, v in plain_s.items()},
            "plain_per_sample_launches_samples_per_s": round(paths / ps_wall, 1),
            "speedup_vs_per_sample_launches": round(ps_wall / tb_wall, 3),
            "batched_s": round(tb_wall, 3), "batched_gpu_s": round(tb_gpu, 3),
            "one_launch_per_scene_s": round(pl_wall, 3), "per_sample_launches_s": round(ps_wall, 3),
            "films_bit_identical": same and same_plain, "roofline": roofline, "cpu_baseline": cpu,
            "autotune": tune, "scenes": len(jobs), "triangles": tris, "paths": paths, "pixels": pixels,
            "setup_s": round(setup_s, 1), "setup_scene_generation_s": round(gen_s, 1),
            "setup_pack_bvh_upload_s": round(load_s, 1)}


def cpu_baseline_batch() -> dict:
    """The FP64 oracle path tracer on a bounded sample of the batch: a
    192x192 centre crop of three of its scenes at their own sample counts."""
    import oracle as O
    from lucent_amd import workloads

    threads = min(16, os.cpu_count() or 1)
    n, dt = 0, 0.0
    picks = (1, 13, 62)
    for i in picks:
        sc, w, h, spp = workloads.batch_scene(i)
        o = O.PtOracle(sc)
        f = o.frame(sc.camera, w, h, workloads.BATCH_DEPTH, 0.0)
        yy, xx = np.mgrid[h // 2 - 96:h // 2 + 96, w // 2 - 96:w // 2 + 96]
        xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1)
        t0 = time.perf_counter()
        o.trace_pixels(f, 0, spp, xy, threads=threads)
        dt += time.perf_counter() - t0
        n += xy.shape[0] * spp
    return {"value": round(n / dt, 1), "unit": "samples/s", "cores": threads, "kind": "port",
            "sample": f"192x192 centre crops of batch scenes {picks} at their own spp ({n} paths), FP64 C "
                      f"restatement of trace.frag on {threads} threads, {dt:.2f} s"}


def bench_copy(args) -> dict:
    """Calibration: a plain device copy of 1 GiB per step (for scaling the
    FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    n = 1 << 30
    src = torch.empty(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {"value": round(2 * n / dt / 1e9, 1), "unit": "GB/s (read+write)", "ms_per_step": round(dt * 1e3, 4),
            "dtype": "u8", "config": {"workload": "device copy of 1 GiB (counter calibration)", "bytes": n}}


def bench_env_sampler(args) -> dict:
    """The environment importance sampler (buildEnvSampler + Vose alias
    table, float64) built on the device from a map already in HBM, against
    the host build of the same tables: the path-trace HDRI (1024x512) and
    the differentiable-rendering map.  Device time between events around the
    seven kernels (median of 9 builds); host time of lucent_env_sampler from
    the same texels; tables compared bit for bit.  Beside them the parallel
    device build (FP64 tree sum and prefix-sum alias pairing: the same
    distribution, other pairs), its table checked by the probability with
    which it selects each cell."""
    import ctypes as C

    import torch

    from lucent_amd import _native, workloads
    from lucent_amd.runtime import Context

    ctx = Context(0)
    lib = _native.lib()
    res = {}
    diff_env = workloads.spheres_scene().env.map
    for name, (w, h, rgb) in {"hdri_1024x512": (1024, 512, workloads.synthetic_hdri(1024, 512, sigma_deg=3.0)),
                              f"diff_env_{diff_env[0]}x{diff_env[1]}": diff_env}.items():
        rgb = np.ascontiguousarray(np.asarray(rgb, np.float32).reshape(-1))
        rgba = np.zeros((h, w, 4), np.float32)
        rgba[..., :3] = rgb.reshape(h, w, 3)
        d_map = torch.from_numpy(rgba).to("cuda")
        dev_ms = []
        for _ in range(10):
            table, gw, gh, total, ms = ctx.env_sampler_gpu(d_map)
            dev_ms.append(ms)
        host_table = np.zeros(gw * gh * 4, np.float32)
        hgw, hgh, htotal = C.c_uint32(), C.c_uint32(), C.c_double()
        host_ms = []
        for _ in range(5):
            t0 = time.perf_counter()
            _native.check(lib.lucent_env_sampler(rgb.ctypes.data, w, h, 1024, host_table.ctypes.data, C.byref(hgw),
                                                 C.byref(hgh), C.byref(htotal)))
            host_ms.append((time.perf_counter() - t0) * 1e3)
        same = bool(np.array_equal(table.cpu().numpy().reshape(-1).view(np.uint32), host_table.view(np.uint32))
                    and total == htotal.value)
        par_ms = []
        for _ in range(10):
            ptable, _, _, _, ms = ctx.env_sampler_gpu(d_map, parallel=True)
            par_ms.append(ms)
        pt = ptable.cpu().numpy().astype(np.float64).reshape(-1, 4)
        implied = pt[:, 0].copy()
        np.add.at(implied, pt[:, 1].astype(np.int64), 1.0 - pt[:, 0])
        target = host_table.reshape(-1, 4)[:, 2].astype(np.float64)
        res[name] = {"grid": [gw, gh], "device_ms": round(float(np.median(dev_ms[1:])), 4),
                     "host_ms": round(float(np.median(host_ms)), 4), "bit_identical": same,
                     "parallel_device_ms": round(float(np.median(par_ms[1:])), 4),
                     "parallel_max_rel_error_of_selection_probability":
                         float(np.max(np.abs(implied - target) / np.maximum(target, 1e-30)))}
        log(f"env sampler {name}: device {res[name]['device_ms']} ms, host {res[name]['host_ms']} ms, "
            f"identical {same}; parallel device build {res[name]['parallel_device_ms']} ms")
    ctx.close()
    return res


def bench_frame_graph(args) -> dict:
    """Interactive frames on the knot scene: 1 sample, a 3-level denoise and
    the display pass per frame, issued as eager calls (9 kernel launches
    through 3 entry points) or as one replay of a captured frame graph, at
    640x480 and at a 128x96 thumbnail.  Both leave bit-identical films;
    frames/s over 200 frames."""
    import ctypes as C

    import torch

    from lucent_amd import _native, workloads
    from lucent_amd.renderer import FrameGraph, Renderer, RenderSettings
    from lucent_amd.runtime import Context

    N = 200
    ctx = Context(0)
    lib = _native.lib()
    res = {}
    for W, H in ((640, 480), (128, 96)):
        out = {}
        for mode in ("eager", "graph"):
            r = Renderer(ctx, W, H, RenderSettings(8, 0.0))
            r.load_scene(workloads.knot_scene())
            r.film.set_seed(1)
            r.reset()
            buf = torch.empty((H, W, 4), dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            dn, dp = (4.0, 64.0, 1.5, 3), _native.DisplayParamsStruct(1.0, 1.0, 0)
            g = None
            if mode == "eager":
                def frame():
                    r.trace_samples(1)
                    r.run_denoiser(*dn)
                    _native.check(lib.lucent_display(ctx.handle, r.film.handle, C.byref(dp), W, H, buf.data_ptr(),
                                                     None))
            else:
                g = FrameGraph(r, 1, dn, (1.0, 1.0, 0), buf.data_ptr())
                frame = g.launch
            for _ in range(5):
                frame()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(N):
                frame()
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / N
            out[mode] = {"ms_per_frame": round(dt * 1e3, 4), "frames_per_s": round(1 / dt, 1)}
            out[mode + "_film"] = r.film.read()["color"]
            if g is not None:
                g.close()
            r.close()
        same = bool(np.array_equal(out.pop("eager_film"), out.pop("graph_film")))
        out["speedup"] = round(out["eager"]["ms_per_frame"] / out["graph"]["ms_per_frame"], 3)
        out["film_bit_identical"] = same
        res[f"{W}x{H}"] = out
    # A batch of 16 small films (four scenes, 160x120, 1 spp each) per
    # iteration: eager lucent_trace_batch (job table uploaded, host
    # synchronised per call) against a replayed batch graph.
    from lucent_amd.renderer import BatchGraph, trace_batch
    scenes = [workloads.knot_scene(), workloads.spheres_scene()]
    settings = RenderSettings(8, 0.0)
    out = {}
    for mode in ("eager", "graph"):
        rs = []
        for i in range(16):
            r = Renderer(ctx, 160, 120, settings)
            r.load_scene(scenes[i % len(scenes)])
            r.film.set_seed(i)
            r.reset()
            rs.append(r)
        jobs = [(r, 1) for r in rs]
        g = BatchGraph(ctx, jobs, settings) if mode == "graph" else None
        step = g.launch if g is not None else (lambda: trace_batch(ctx, jobs, settings))
        for _ in range(5):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(N):
            step()
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / N
        out[mode] = {"ms_per_iteration": round(dt * 1e3, 4), "iterations_per_s": round(1 / dt, 1)}
        out[mode + "_films"] = [r.film.read()["color"] for r in rs]
        if g is not None:
            g.close()
        for r in rs:
            r.close()
    same = all(np.array_equal(a, b) for a, b in zip(out.pop("eager_films"), out.pop("graph_films")))
    out["speedup"] = round(out["eager"]["ms_per_iteration"] / out["graph"]["ms_per_iteration"], 3)
    out["films_bit_identical"] = same
    res["batch_16x160x120_1spp"] = out
    ctx.close()
    return {"metric": "interactive frames/s (1 spp + denoise + display), graph replay", "value":
            res["128x96"]["graph"]["frames_per_s"], "unit": "frames/s @128x96", **res}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="all",
                    choices=["all", "closest-hit", "path-trace", "batch", "denoise", "adjoint", "frame-graph", "copy",
                             "env-sampler"])
    ap.add_argument("--cpu-sample", type=int, default=1 << 19)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="batch: time only the batched launch (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path shards by "
              "rays/pixels/scenes with no exchange, see DESIGN.md, but no multi-GPU figure is produced)",
              file=sys.stderr)
        return 2
    extra = {}
    if args.workload == "closest-hit":
        res = bench_closest_hit(args)
    elif args.workload == "copy":
        res = bench_copy(args)
    elif args.workload == "adjoint":
        d = bench_adjoint(args)
        res = {"value": d.pop("value"), "unit": d.pop("unit"), "ms_per_step": d["forward_ms"] + d["backward_ms"],
               "dtype": "f32/f64", "config": {"workload": d.pop("metric")}, **d}
    elif args.workload == "batch":
        d = bench_batch(args)
        res = {"value": d.pop("value"), "unit": d.pop("unit"), "ms_per_step": d["batched_s"] * 1e3, "dtype": "f32",
               "config": {"workload": d.pop("metric")}, **d}
    elif args.workload == "env-sampler":
        d = bench_env_sampler(args)
        first = d["hdri_1024x512"]
        res = {"value": first["device_ms"], "unit": "ms per sampler build (1024x512 map)",
               "ms_per_step": first["device_ms"], "dtype": "f64", "higher_is_better": False,
               "config": {"workload": "environment sampler build from a map in HBM"}, **d}
    elif args.workload == "frame-graph":
        d = bench_frame_graph(args)
        res = {"value": d.pop("value"), "unit": d.pop("unit"), "ms_per_step": d["128x96"]["graph"]["ms_per_frame"],
               "dtype": "f32/f16", "config": {"workload": d.pop("metric")}, **d}
    elif args.workload == "denoise":
        d = bench_denoise(args)
        res = {"value": d["value"], "unit": d["unit"], "ms_per_step": d["value"], "dtype": "f32/f16",
               "higher_is_better": False, "config": {"workload": d["metric"]}, "roofline": d["roofline"],
               "cpu_baseline": d["cpu_baseline"]}
    else:
        res = bench_path_trace(args)
        if args.workload == "all":
            ch = bench_closest_hit(args)
            extra["closest_hit"] = {"metric": "Mrays/s closest-hit", "value": ch.pop("value"),
                                    "unit": ch.pop("unit"), **ch}
            extra["batch"] = bench_batch(args)
            extra["denoise"] = bench_denoise(args)
            extra["adjoint"] = bench_adjoint(args)
            extra["frame_graph"] = bench_frame_graph(args)
            extra["env_sampler"] = bench_env_sampler(args)
            cp = bench_copy(argparse.Namespace(steps=10, warmup=2))
            extra["streaming_copy"] = {"value": cp["value"], "unit": cp["unit"], "bytes_per_copy": 1 << 30,
                                       "note": "measured device-to-device copy bandwidth, beside the 8 TB/s peak"}
    line = {"metric": METRIC, "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": res.pop("higher_is_better", True), "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic", "config": res.pop("config"), "roofline": res.pop("roofline", None),
            "cpu_baseline": res.pop("cpu_baseline", None)}
    if "roofline" not in line or line["roofline"] is None:
        line["roofline"] = res.pop("roofline", None)
    line.update(res)
    line.update(extra)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
