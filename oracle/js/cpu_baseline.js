// TEST INFRASTRUCTURE — times the JavaScript port of intersectBVH on the host
// cores with worker_threads (the CPU baseline of the closest-hit benchmark).
//
// usage: node cpu_baseline.js <dir> <workers> <repeats>
//   <dir> holds nodes.f32, tri_indices.u32, positions.f32, origins.f32,
//   dirs.f32 (raw little-endian).  Rays are split into contiguous slices,
//   one per worker; BVH, positions and rays are shared via SharedArrayBuffer.
//   One warm-up pass, then `repeats` timed passes; prints one JSON line with
//   the median wall time from dispatch until every worker has returned, plus
//   the hit ids of the last pass written to <dir>/ids.u32.
'use strict';
const fs = require('fs');
const os = require('os');
const path = require('path');
const { Worker, isMainThread, parentPort, workerData } = require('worker_threads');
const { intersectBVH } = require('./intersect_bvh');

function shared(file, Ctor) {
  const buf = fs.readFileSync(file);
  const sab = new SharedArrayBuffer(buf.length);
  new Uint8Array(sab).set(buf);
  return new Ctor(sab);
}

if (isMainThread) {
  const dir = process.argv[2];
  const workers = parseInt(process.argv[3] || String(os.cpus().length), 10);
  const repeats = parseInt(process.argv[4] || '5', 10);
  const data = {
    nodes: shared(path.join(dir, 'nodes.f32'), Float32Array),
    tri: shared(path.join(dir, 'tri_indices.u32'), Uint32Array),
    pos: shared(path.join(dir, 'positions.f32'), Float32Array),
    org: shared(path.join(dir, 'origins.f32'), Float32Array),
    dirs: shared(path.join(dir, 'dirs.f32'), Float32Array),
  };
  const n = data.org.length / 3;
  const ids = new Uint32Array(new SharedArrayBuffer(n * 4));
  const pool = [];
  for (let w = 0; w < workers; w++) {
    const a = Math.floor((n * w) / workers);
    const b = Math.floor((n * (w + 1)) / workers);
    pool.push(new Worker(__filename, { workerData: Object.assign({ a, b, ids: ids }, data) }));
  }
  const pass = () => new Promise((resolve) => {
    let left = pool.length;
    const t0 = process.hrtime.bigint();
    for (const wk of pool) {
      wk.once('message', () => {
        left -= 1;
        if (left === 0) resolve(Number(process.hrtime.bigint() - t0) / 1e9);
      });
      wk.postMessage('go');
    }
  });
  (async () => {
    await pass();
    const times = [];
    for (let r = 0; r < repeats; r++) times.push(await pass());
    times.sort((x, y) => x - y);
    const med = times[Math.floor(times.length / 2)];
    fs.writeFileSync(path.join(dir, 'ids.u32'), Buffer.from(ids.buffer));
    let hits = 0;
    for (let i = 0; i < n; i++) if (ids[i] !== 0xFFFFFFFF) hits++;
    process.stdout.write(JSON.stringify({ rays: n, workers, repeats, median_s: med, times_s: times,
      mrays_per_s: n / med / 1e6, hits, node: process.version, cpus: os.cpus().length }) + '\n');
    for (const wk of pool) wk.terminate();
  })();
} else {
  const d = workerData;
  const bvh = { nodes: d.nodes, triIndices: d.tri };
  parentPort.on('message', () => {
    for (let i = d.a; i < d.b; i++) {
      const ray = { origin: [d.org[i * 3], d.org[i * 3 + 1], d.org[i * 3 + 2]],
        dir: [d.dirs[i * 3], d.dirs[i * 3 + 1], d.dirs[i * 3 + 2]] };
      const hit = intersectBVH(bvh, d.pos, ray, 1e-5, Infinity);
      d.ids[i] = hit ? hit.triIndex : 0xFFFFFFFF;
    }
    parentPort.postMessage('done');
  });
}
