"""Writes config1_counts.json: the per-ray fetch counts that define the
algorithmic bytes of the closest-hit benchmark.

Counts come from the oracle walking the reference shader's traversal order
(src/shaders/scene.glsl:321-376: root box untested, near child first) over
the benchmark's own rays: P = interior visits (each fetches one 64-byte child
pair), T = triangle tests (48 bytes each).  Run from the repo root:

    python tests/golden/make_config1_counts.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle as O  # noqa: E402
from lucent_amd import workloads  # noqa: E402
from lucent_amd.bvh import build_bvh  # noqa: E402


def main():
    w = workloads.config1()
    b = build_bvh(w["positions"])
    c = O.shader_order_counts(b, w["positions"], w["origins"], w["dirs"])
    t, tri, u, v = O.intersect_bvh_batch(b, w["positions"], w["origins"], w["dirs"])
    out = {
        "workload": "81,920-triangle fBm icosphere (subdivision 6), 2^20 sphere rays, seed 1",
        "triangles": int(w["positions"].size // 9),
        "rays": int(c["rays"]),
        "bvh_nodes": int(b.node_count),
        "bvh_depth": int(b.depth()),
        "P_interior_visits_per_ray": c["P"],
        "T_triangle_tests_per_ray": c["T"],
        "leaf_visits_per_ray": c["leaf_visits"],
        "hit_fraction": float((tri != 0xFFFFFFFF).mean()),
        "hit_id_checksum": int(tri.astype("uint64").sum() % (1 << 61)),
        "t_sum": float(t[tri != 0xFFFFFFFF].sum()),
    }
    path = Path(__file__).resolve().parent / "config1_counts.json"
    path.write_text(json.dumps(out, indent=2) + "\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
