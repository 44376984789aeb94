"""Writes batch_counts.json: per-path fetch counts of every scene of the
batched 64-scene workload, which define its algorithmic bytes.

For each scene the FP64 oracle (its own packing and tree) traces 512 pixels
drawn uniformly over the film (a seeded draw per scene), 4 samples each,
depth 8, counting in the reference shader's visiting order (as
make_pathtrace_counts.py).  Takes a few minutes (the oracle builds every
scene's tree, 11.8 M triangles in all).

    python tests/golden/make_batch_counts.py
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
    scenes = []
    for i in range(workloads.BATCH_SCENES):
        sc, w, h, spp = workloads.batch_scene(i)
        o = O.PtOracle(sc)
        f = o.frame(sc.camera, w, h, workloads.BATCH_DEPTH, 0.0)
        rng = np.random.default_rng(1000 + i)
        xy = np.stack([rng.integers(0, w, 512), rng.integers(0, h, 512)], -1)
        _, _, c = o.trace_pixels(f, 0, 4, xy, counts=True)
        paths = c.pop("paths")
        per = {k: round(v / paths, 4) for k, v in c.items()}
        scenes.append({"index": i, "width": w, "height": h, "spp": spp, "tris": int(o.packed.tri_count),
                       "per_path": per})
        print(i, w, h, spp, o.packed.tri_count, per["closest_rays"], flush=True)
    out = {"workload": "batch_scene(0..63, seed 2024), depth 8, clamp 0; 512 uniform pixels x 4 spp per scene",
           "scenes": scenes}
    path = Path(__file__).resolve().parent / "batch_counts.json"
    path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
