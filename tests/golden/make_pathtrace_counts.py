"""Writes pathtrace_counts.json: per-path fetch counts of the 1440x900
path-trace benchmark, which define its algorithmic bytes per path.

The FP64 oracle traces a 64x64 crop at the centre of the frame, 16 samples
per pixel (seeds 0..15), counting in the reference shader's visiting order:
closest-hit rays with their 64-byte child-pair fetches and 48-byte triangle
tests, shadow rays with their 32-byte node fetches and triangle tests,
surface vertices, environment and light NEE samples and escaped rays.

    python tests/golden/make_pathtrace_counts.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
from lucent_amd import workloads  # noqa: E402


def main():
    sc = workloads.knot_scene()
    o = O.PtOracle(sc)
    W, H = workloads.PT_WIDTH, workloads.PT_HEIGHT
    f = o.frame(sc.camera, W, H, workloads.PT_DEPTH, 0.0)
    yy, xx = np.mgrid[H // 2 - 32:H // 2 + 32, W // 2 - 32:W // 2 + 32]
    xy = np.stack([xx.reshape(-1), yy.reshape(-1)], -1)
    _, _, c = o.trace_pixels(f, 0, 16, xy, counts=True)
    paths = c["paths"]
    per = {k: v / paths for k, v in c.items() if k != "paths"}
    out = {"workload": "knot_scene(seed 7) at 1440x900, depth 8, clamp 0; 64x64 centre crop, 16 spp",
           "paths": int(paths), "per_path": per}
    path = Path(__file__).resolve().parent / "pathtrace_counts.json"
    path.write_text(json.dumps(out, indent=2) + "\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
