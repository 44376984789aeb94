"""Writes adjoint_counts.json: per-path fetch counts of the differentiable-
rendering workload (16 displaced spheres, 512x512, depth 6), which define
the algorithmic bytes of its forward pass and of its path replay.

The FP64 oracle traces 2048 pixels drawn uniformly over the film, 8 samples
each, counting in the reference shader's visiting order.

    python tests/golden/make_adjoint_counts.py
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
    sc = workloads.spheres_scene()
    W, H = workloads.DIFF_WIDTH, workloads.DIFF_HEIGHT
    o = O.PtOracle(sc)
    f = o.frame(sc.camera, W, H, workloads.DIFF_DEPTH, 0.0)
    rng = np.random.default_rng(77)
    xy = np.stack([rng.integers(0, W, 2048), rng.integers(0, H, 2048)], -1)
    _, _, c = o.trace_pixels(f, 0, 8, xy, counts=True)
    paths = c.pop("paths")
    out = {"workload": "spheres_scene(seed 11) at 512x512, depth 6, clamp 0; 2048 uniform pixels x 8 spp",
           "paths": int(paths), "per_path": {k: v / paths for k, v in c.items()}}
    path = Path(__file__).resolve().parent / "adjoint_counts.json"
    path.write_text(json.dumps(out, indent=2) + "\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
