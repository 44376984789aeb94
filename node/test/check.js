// Test driver for the N-API facade (run by tests/test_node_addon.py).
//   node check.js guard <dir>   the async guard of every synchronous LucentRenderer member (CPU)
//   node check.js bvh <dir>     buildBVH over <dir>/positions.f32 -> nodes.f32, tri.u32
//   node check.js hits <dir>    GpuTraversal over positions + bvh, rays org.f64/dir.f64 -> t/u/v.f64, tri.u32
//   node check.js hdr <dir>     parseHDR(in.hdr) -> rgb.f32
//   node check.js gltf <dir>    loadGLTF(in.gltf, files of <dir>) -> positions.f32 and a JSON summary
//   node check.js async <dir>   the render scene through the queued async calls vs the sync ones
//   node check.js render <dir>  LucentRenderer over the packed scene in <dir> -> color.f32, albedo.f32,
//                               den.f32, then a camera move -> moved.f32, shot.u8
'use strict';
const fs = require('fs');
const path = require('path');
const lucent = require('../lucent.js');

const [mode, dir] = process.argv.slice(2);
const f32 = (n) => { const b = fs.readFileSync(path.join(dir, n)); return new Float32Array(b.buffer, b.byteOffset, b.length / 4); };
const f64 = (n) => { const b = fs.readFileSync(path.join(dir, n)); return new Float64Array(b.buffer, b.byteOffset, b.length / 8); };
const u32 = (n) => { const b = fs.readFileSync(path.join(dir, n)); return new Uint32Array(b.buffer, b.byteOffset, b.length / 4); };
const i32 = (n) => { const b = fs.readFileSync(path.join(dir, n)); return new Int32Array(b.buffer, b.byteOffset, b.length / 4); };
const write = (n, ta) => fs.writeFileSync(path.join(dir, n), Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength));

const same = (a, b) => a.length === b.length && Buffer.from(a.buffer, a.byteOffset, a.byteLength)
  .equals(Buffer.from(b.buffer, b.byteOffset, b.byteLength));

if (mode === 'guard') {
  // CPU only: a renderer with an asynchronous call in flight refuses every
  // synchronous member that touches its context, scene or film.
  const r = Object.create(lucent.LucentRenderer.prototype);
  r._inflight = 1;
  r.settings = { temporal: true };
  const calls = {
    loadScene: () => r.loadScene({}, {}), uploadLighting: () => r.uploadLighting({}),
    uploadCamera: () => r.uploadCamera({}), reset: () => r.reset(), traceSamples: () => r.traceSamples(1),
    render: () => r.render(1), screenshot: () => r.screenshot(), benchmark: () => r.benchmark(1),
    runDenoiser: () => r.runDenoiser(), read: () => r.read(), stats: () => r.stats(), dispose: () => r.dispose(),
  };
  const out = {};
  for (const [name, fn] of Object.entries(calls)) {
    try { fn(); out[name] = 'ran'; } catch (e) { out[name] = e.message; }
  }
  process.stdout.write(JSON.stringify(out) + '\n');
} else if (mode === 'bvh') {
  const bvh = lucent.buildBVH(f32('positions.f32'), { maxLeafSize: 4, bins: 12 });
  write('nodes.f32', bvh.nodes);
  write('tri.u32', bvh.triIndices);
  let err = '';
  try { lucent.buildBVH(new Float32Array(10)); } catch (e) { err = e.message; }
  lucent.buildBVHAsync(f32('positions.f32'), { maxLeafSize: 4, bins: 12 }).then((abvh) => {
    const asyncSame = abvh.nodeCount === bvh.nodeCount && same(abvh.nodes, bvh.nodes) && same(abvh.triIndices, bvh.triIndices);
    return lucent.buildBVHAsync(new Float32Array(10)).then(() => 'resolved', (e) => e.message).then((asyncErr) => {
      process.stdout.write(JSON.stringify({ nodeCount: bvh.nodeCount, bounds: bvh.bounds, error: err, asyncSame, asyncErr }) + '\n');
    });
  });
} else if (mode === 'hdr') {
  const img = lucent.parseHDR(fs.readFileSync(path.join(dir, 'in.hdr')));
  write('rgb.f32', img.data);
  let err = '';
  try { lucent.parseHDR(Buffer.from('P6\n1 1\n255\n')); } catch (e) { err = e.message; }
  process.stdout.write(JSON.stringify({ width: img.width, height: img.height, error: err }) + '\n');
} else if (mode === 'gltf') {
  const fetched = [];
  const resolve = (uri) => { fetched.push(uri); return Promise.resolve(fs.readFileSync(path.join(dir, uri))); };
  lucent.loadGLTF(fs.readFileSync(path.join(dir, 'in.gltf')), resolve).then((a) => {
    write('positions.f32', a.meshes[0].mesh.positions);
    return lucent.loadGLTF(Buffer.from('{"asset":{"version":"1.0"}}'), resolve).then(() => '', (e) => e.message)
      .then((err) => {
        process.stdout.write(JSON.stringify({
          fetched, triangleCount: a.triangleCount, bounds: a.bounds, warnings: a.warnings,
          material: a.materials[a.meshes[0].material], hasUvs: !!a.meshes[0].mesh.uvs,
          images: a.images.map((im) => ({ mimeType: im.mimeType, name: im.name, data: Array.from(im.data) })), error: err,
        }) + '\n');
      });
  });
} else if (mode === 'hits') {
  const bvh = { nodes: f32('nodes.f32'), triIndices: u32('tri.u32') };
  const tr = new lucent.GpuTraversal(bvh, f32('positions.f32'));
  const r = tr.intersectBatch(f64('org.f64'), f64('dir.f64'));
  write('t.f64', r.t); write('u.f64', r.u); write('v.f64', r.v); write('hit.u32', r.triIndex);
  const one = tr.intersect({ origin: Array.from(f64('org.f64').subarray(0, 3)), dir: Array.from(f64('dir.f64').subarray(0, 3)) });
  process.stdout.write(JSON.stringify({ first: one }) + '\n');
} else if (mode === 'async') {
  // The same scene as 'render': sync calls on one renderer, queued async calls on another.
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const packed = {
    triGeometry: f32('geo.f32'), triAttributes: f32('attr.f32'), triCount: meta.triCount,
    bvhNodes: f32('nodes.f32'), bvhNodeCount: meta.nodeCount, materials: f32('mat.f32'),
    materialCount: meta.materialCount, spheres: f32('sph.f32'), sphereMaterials: i32('sphm.i32'),
    sphereLights: i32('sphl.i32'), sphereCount: meta.sphereCount, lights: f32('lights.f32'),
    lightCount: meta.lightCount, groundMaterial: meta.groundMaterial, cpu: { bvh: { triIndices: u32('tri.u32') } },
  };
  const scene = Object.assign({}, meta.scene);
  scene.env = Object.assign({}, meta.scene.env, { map: { width: meta.envW, height: meta.envH, data: f32('env.f32') } });
  const mk = () => {
    const r = new lucent.LucentRenderer({ width: meta.width, height: meta.height });
    r.settings = { maxBounces: 8, fireflyClamp: 0, temporal: true, denoise: true, exposure: 1, tonemap: 'aces' };
    r.loadScene(packed, scene);
    lucent.addon.resetFilm(r.ctx, r.film, meta.seed);
    return r;
  };
  const params = { sigmaL: 4, sigmaN: 64, sigmaP: 1.5, iterations: 3 };
  const a = mk();
  a.traceSamples(meta.spp);
  const da = a.runDenoiser(params);
  const b = mk();
  const p1 = b.traceSamplesAsync(1);
  const p2 = b.traceSamplesAsync(meta.spp - 1);
  let guard = '';
  try { b.traceSamples(1); } catch (e) { guard = e.message; }
  const p3 = b.runDenoiserAsync(params);
  Promise.all([p1, p2, p3]).then(([, , db]) => {
    const filmSame = same(a.read().color, b.read().color);
    const denSame = same(da.illumVar, db.illumVar) && same(da.radiance, db.radiance);
    process.stdout.write(JSON.stringify({ filmSame, denSame, guard, samples: b.stats().samples }) + '\n');
  });
} else if (mode === 'render') {
  const meta = JSON.parse(fs.readFileSync(path.join(dir, 'meta.json'), 'utf8'));
  const packed = {
    triGeometry: f32('geo.f32'), triAttributes: f32('attr.f32'), triCount: meta.triCount,
    bvhNodes: f32('nodes.f32'), bvhNodeCount: meta.nodeCount, materials: f32('mat.f32'),
    materialCount: meta.materialCount, spheres: f32('sph.f32'), sphereMaterials: i32('sphm.i32'),
    sphereLights: i32('sphl.i32'), sphereCount: meta.sphereCount, lights: f32('lights.f32'),
    lightCount: meta.lightCount, groundMaterial: meta.groundMaterial, cpu: { bvh: { triIndices: u32('tri.u32') } },
  };
  const scene = Object.assign({}, meta.scene);
  scene.env = Object.assign({}, meta.scene.env, { map: { width: meta.envW, height: meta.envH, data: f32('env.f32') } });
  const r = new lucent.LucentRenderer({ width: meta.width, height: meta.height });
  r.settings = { maxBounces: 8, fireflyClamp: 0, temporal: true, denoise: true, exposure: 1, tonemap: 'aces' };
  r.loadScene(packed, scene);
  lucent.addon.resetFilm(r.ctx, r.film, meta.seed);
  r.traceSamples(meta.spp);
  const out = r.read();
  write('color.f32', out.color);
  write('albedo.f32', out.albedo);
  const d = r.runDenoiser({ sigmaL: 4, sigmaN: 64, sigmaP: 1.5, iterations: 3 });
  write('den.f32', d.illumVar);
  const before = r.stats();
  // A camera move with temporal reuse: one reprojecting sample, one plain sample, then the displayed image.
  r.uploadCamera(meta.camera2);
  r.traceSamples(2);
  write('moved.f32', r.read().color);
  write('shot.u8', r.screenshot().data);
  process.stdout.write(JSON.stringify(Object.assign(before, { movedSamples: r.stats().samples })) + '\n');
}
