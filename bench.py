"""Benchmark of the MI355X path-tracing hot path (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload closest-hit]

closest-hit: one step = one batched exact closest-hit launch of 2^20 rays
against the 81,920-triangle fBm icosphere (inputs resident in HBM).

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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline_closest_hit(bvh, positions, origins, dirs, sample: int) -> dict:
    """Times the JavaScript port of intersectBVH (oracle/js) with Node worker_threads on a
    contiguous sample of the benchmark rays; checks its hit ids against the GPU's."""
    workers = int(os.environ.get("LUCENT_CPU_WORKERS", "0")) or min(16, os.cpu_count() or 1)
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
    traffic = None
    tr_path = ROOT / "profiles/closest_hit_icosphere81920_2e20rays_traffic.json"
    if tr_path.exists():
        traffic = json.loads(tr_path.read_text()).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_ray": round(b_ray, 1), "avg_launch_ms": round(avg_launch_s * 1e3, 4)}

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
    """Plain progressive path trace: 1440x900, 64 spp per step (one launch per
    sample, as Renderer.traceSamples), depth 8, RR from bounce 2, clamp 0."""
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

    def step():
        r.reset()
        r.trace_samples(SPP, stream=stream.cuda_stream)

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
    step_ms = [s_.elapsed_time(e) for s_, e in zip(starts, ends)]
    paths = W * H * SPP
    value = paths * args.steps / wall
    launch_s = float(np.mean(step_ms)) / 1e3 / SPP
    img = r.film.read()["color"]
    mean_rgb = [float(x) for x in img[..., :3].reshape(-1, 3).mean(0)]
    r.close()
    ctx.close()
    return {"value": round(value, 1), "unit": "samples/s", "ms_per_step": round(wall / args.steps * 1e3, 3),
            "dtype": "f32",
            "config": {"workload": f"plain progressive path trace {W}x{H}, {SPP} spp, depth {workloads.PT_DEPTH}, "
                                   f"RR from bounce 2, clamp 0; 6 fBm torus knots + ground quad "
                                   f"({stats['tri_count']} tris), quad light, 1024x512 HDRI",
                       "width": W, "height": H, "spp": SPP, "tris": stats["tri_count"],
                       "scene_load_s": round(load_s, 3)},
            "avg_sample_launch_ms": round(launch_s * 1e3, 4), "image_mean_rgb": mean_rgb,
            "mpaths_per_s": round(value / 1e6, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="closest-hit", choices=["closest-hit", "path-trace"])
    ap.add_argument("--cpu-sample", type=int, default=1 << 19)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path shards by "
              "rays/pixels/scenes with no exchange, see DESIGN.md, but no multi-GPU figure is produced)",
              file=sys.stderr)
        return 2
    res = bench_closest_hit(args) if args.workload == "closest-hit" else bench_path_trace(args)
    line = {"metric": METRIC, "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic", "config": res.pop("config"), "roofline": res.pop("roofline", None),
            "cpu_baseline": res.pop("cpu_baseline", None)}
    if "roofline" not in line or line["roofline"] is None:
        line["roofline"] = res.pop("roofline", None)
    line.update(res)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
