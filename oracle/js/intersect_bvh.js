// TEST INFRASTRUCTURE — the CPU baseline port, not the product.
//
// A JavaScript restatement of Lucent's CPU closest-hit traversal with the same
// algorithm, float64 arithmetic and per-test allocation shape as the
// reference TypeScript: intersectBVH (src/math/bvh.ts:265-320),
// AABB.intersect (src/math/aabb.ts:84-98), intersectTriangle
// (src/math/ray.ts:22-45) and the vec3 helpers (src/math/vec3.ts:15-25).
// Runs on Node 12 (no TypeScript syntax).
'use strict';

const sub = (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]];
const dot = (a, b) => a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];

function intersectTriangle(ray, v0, v1, v2, tMin, tMax) {
  const e1 = sub(v1, v0);
  const e2 = sub(v2, v0);
  const p = cross(ray.dir, e2);
  const det = dot(e1, p);
  if (Math.abs(det) < 1e-12) return null;
  const invDet = 1 / det;
  const s = sub(ray.origin, v0);
  const u = dot(s, p) * invDet;
  if (u < 0 || u > 1) return null;
  const q = cross(s, e1);
  const v = dot(ray.dir, q) * invDet;
  if (v < 0 || u + v > 1) return null;
  const t = dot(e2, q) * invDet;
  if (t <= tMin || t >= tMax) return null;
  return { t, u, v };
}

class AABB {
  constructor() {
    this.min = [Infinity, Infinity, Infinity];
    this.max = [-Infinity, -Infinity, -Infinity];
  }
  intersect(origin, invDir, tMax) {
    let tmin = 0;
    let tmax = tMax;
    for (let i = 0; i < 3; i++) {
      const o = origin[i];
      const inv = invDir[i];
      let t0 = (this.min[i] - o) * inv;
      let t1 = (this.max[i] - o) * inv;
      if (t0 > t1) { const tt = t0; t0 = t1; t1 = tt; }
      if (t0 > tmin) tmin = t0;
      if (t1 < tmax) tmax = t1;
      if (tmax < tmin) return Infinity;
    }
    return tmin;
  }
}

const NODE_STRIDE = 8;

function intersectBVH(bvh, positions, ray, tMin, tMax) {
  if (tMin === undefined) tMin = 1e-5;
  if (tMax === undefined) tMax = Infinity;
  const nodes = bvh.nodes;
  const triIndices = bvh.triIndices;
  if (triIndices.length === 0) return null;
  const invDir = [1 / ray.dir[0], 1 / ray.dir[1], 1 / ray.dir[2]];
  const stack = new Int32Array(64);
  let sp = 0;
  stack[sp++] = 0;
  let best = null;
  let closest = tMax;
  const box = new AABB();
  while (sp > 0) {
    const n = stack[--sp];
    const o = n * NODE_STRIDE;
    box.min[0] = nodes[o]; box.min[1] = nodes[o + 1]; box.min[2] = nodes[o + 2];
    box.max[0] = nodes[o + 4]; box.max[1] = nodes[o + 5]; box.max[2] = nodes[o + 6];
    if (box.intersect(ray.origin, invDir, closest) === Infinity) continue;
    const leftFirst = nodes[o + 3];
    const count = nodes[o + 7];
    if (count > 0) {
      for (let i = 0; i < count; i++) {
        const tri = triIndices[leftFirst + i];
        const p = tri * 9;
        const hit = intersectTriangle(
          ray,
          [positions[p], positions[p + 1], positions[p + 2]],
          [positions[p + 3], positions[p + 4], positions[p + 5]],
          [positions[p + 6], positions[p + 7], positions[p + 8]],
          tMin,
          closest,
        );
        if (hit) {
          closest = hit.t;
          best = { t: hit.t, triIndex: tri, u: hit.u, v: hit.v };
        }
      }
    } else {
      stack[sp++] = leftFirst + 1;
      stack[sp++] = leftFirst;
    }
  }
  return best;
}

module.exports = { intersectBVH, intersectTriangle, AABB };
