// @ts-check
'use strict';
/**
 * Node host facade of the MI355X path-tracing library (CommonJS, runs on
 * Node 12).  It mirrors the reference's TypeScript entry points on this
 * path so that a caller swaps implementations, not call sites:
 *   buildBVH(positions, options)          src/math/bvh.ts:52
 *   buildBVHAsync(positions, options)     src/math/bvh-async.ts:8
 *   parseHDR(buffer)                      src/loaders/hdr.ts:7
 *   loadGLTF(source, resolve)             src/loaders/gltf.ts:188
 *   GpuTraversal#intersect / pickDistance src/math/bvh.ts:265, src/scene/scene.ts:327
 *   LucentRenderer                        src/renderer.ts:171 (loadScene, uploadLighting,
 *                                         uploadCamera with temporal reprojection, reset,
 *                                         traceSamples, render, screenshot, benchmark,
 *                                         runDenoiser, stats, dispose)
 * Types are in lucent.d.ts.
 */
const path = require('path');

/** @type {any} */
const addon = require(path.join(__dirname, 'build', 'lucent_napi.node'));

const QUERY_ANY = 1;
/** Samples a reprojected pixel may carry (HISTORY_CAP, renderer.ts:139). */
const HISTORY_CAP = 8;
const TONEMAPS = { aces: 0, reinhard: 1, none: 2 };
/** atrousLevels / denoiseWeight (renderer.ts:142-144). */
const denoiseWeight = (spp) => Math.min(1, Math.max(0, (1024 - spp) / (1024 - 128)));
const PREC_GLSL_FP32 = 2;
const MISS = 0xffffffff;

/**
 * buildBVH (src/math/bvh.ts:52): binned SAH, same node layout and leaf order.
 * @param {Float32Array} positions 9 floats per triangle
 * @param {{maxLeafSize?: number, bins?: number}} [options]
 */
function buildBVH(positions, options) {
  return addon.buildBVH(positions, options || {});
}

/**
 * buildBVHAsync (src/math/bvh-async.ts:8-30): the same build on a worker
 * thread; the Promise rejects with the same errors.
 * @param {Float32Array} positions 9 floats per triangle
 * @param {{maxLeafSize?: number, bins?: number}} [options]
 */
function buildBVHAsync(positions, options) {
  try {
    return addon.buildBVHAsync(positions, options || {});
  } catch (e) {
    return Promise.reject(e);
  }
}

/**
 * parseHDR (src/loaders/hdr.ts:7): a Radiance .hdr file -> HDRImage
 * {width, height, data} (RGB floats, rows top to bottom).  Throws Error
 * with parseHDR's messages.
 * @param {ArrayBuffer|Uint8Array} buffer
 */
function parseHDR(buffer) {
  const bytes = buffer instanceof Uint8Array ? buffer : new Uint8Array(buffer);
  return addon.decodeHDR(bytes);
}

const ALPHA_MODES = ['opaque', 'mask', 'blend'];

/** A Material (material.ts:17-57) from its 40 texel values (material.ts:132-149 order). */
function materialFromTexels(t) {
  return {
    albedo: [t[0], t[1], t[2]], metallic: t[3], roughness: t[4], ior: t[5], transmission: t[6],
    checkerScale: t[7], emission: [t[8], t[9], t[10]], alpha: t[11], albedo2: [t[12], t[13], t[14]],
    alphaMode: ALPHA_MODES[t[15]], baseColorMap: t[16], metalRoughMap: t[17], normalMap: t[18],
    emissiveMap: t[19], normalScale: t[20], clearcoat: t[21], clearcoatRoughness: t[22], alphaCutoff: t[23],
    sheenColor: [t[24], t[25], t[26]], sheenRoughness: t[27], attenuationColor: [t[28], t[29], t[30]],
    attenuationDistance: t[31] > 0 ? t[31] : Infinity, thinWalled: t[35] !== 0,
    uvTransform: [t[32], t[33], t[34], t[36], t[37], t[38]],
  };
}

/**
 * loadGLTF (src/loaders/gltf.ts:188): a .glb or .gltf file -> Promise of
 * GLTFAsset {meshes: [{mesh, material}], materials, images, bounds,
 * triangleCount, warnings}.  The native loader lists the external files it
 * needs (buffers, then the images some material samples); `resolve(uri)`
 * fetches each (an ArrayBuffer or a Promise of one).  Rejects with the
 * reference's messages.
 * @param {ArrayBuffer|Uint8Array} source
 * @param {(uri: string) => Promise<ArrayBuffer>|ArrayBuffer} resolve
 */
async function loadGLTF(source, resolve) {
  const bytes = source instanceof Uint8Array ? source : new Uint8Array(source);
  const doc = addon.gltfParse(bytes);
  for (let i = 0; i < doc.uris.length; i++) {
    const data = await resolve(doc.uris[i]);
    addon.gltfSetResource(doc.handle, i, data instanceof Uint8Array ? data : new Uint8Array(data));
  }
  const a = addon.gltfBuild(doc.handle);
  return {
    meshes: a.meshes.map((m) => ({
      mesh: m.uvs ? { positions: m.positions, normals: m.normals, uvs: m.uvs }
        : { positions: m.positions, normals: m.normals },
      material: m.material,
    })),
    materials: a.materials.map(materialFromTexels),
    images: a.images,
    bounds: a.bounds,
    triangleCount: a.triangleCount,
    warnings: a.warnings,
  };
}

/** Batched closest-hit queries against one uploaded BVH (intersectBVH on the GPU). */
class GpuTraversal {
  /**
   * @param {{nodes: Float32Array, triIndices: Uint32Array}} bvh
   * @param {Float32Array} positions
   * @param {number} [device]
   */
  constructor(bvh, positions, device) {
    this.ctx = addon.createContext(device || 0);
    this.scene = addon.sceneFromBVH(this.ctx, positions, bvh);
  }

  /**
   * Exact FP64 hit records for many rays (intersectBVH semantics, bit for bit).
   * @param {Float32Array|Float64Array} origins xyz per ray
   * @param {Float32Array|Float64Array} dirs xyz per ray
   * @param {number} [tMin]
   * @param {number} [tMax]
   * @param {{anyHit?: boolean, glsl?: boolean}} [opts]
   */
  intersectBatch(origins, dirs, tMin, tMax, opts) {
    const o = opts || {};
    const flags = (o.anyHit ? QUERY_ANY : 0) | (o.glsl ? PREC_GLSL_FP32 : 0);
    return addon.intersect(this.ctx, this.scene, origins, dirs, tMin === undefined ? 1e-5 : tMin,
      tMax === undefined ? Infinity : tMax, flags);
  }

  /**
   * intersectBVH(bvh, positions, ray, tMin, tMax) for one ray: a BVHHit or null.
   * @param {{origin: number[], dir: number[]}} ray
   * @param {number} [tMin]
   * @param {number} [tMax]
   */
  intersect(ray, tMin, tMax) {
    const r = this.intersectBatch(Float64Array.from(ray.origin), Float64Array.from(ray.dir), tMin, tMax);
    if (r.triIndex[0] === MISS) return null;
    return { t: r.t[0], triIndex: r.triIndex[0], u: r.u[0], v: r.v[0] };
  }
}

/** Camera basis of Renderer (src/renderer.ts:156-161), for pickFocus. */
function cameraBasis(cam) {
  const sub = (a, b) => [a[0] - b[0], a[1] - b[1], a[2] - b[2]];
  const cross = (a, b) => [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]];
  const norm = (a) => { const l = Math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); return l === 0 ? [0, 0, 0] : [a[0] / l, a[1] / l, a[2] / l]; };
  const forward = norm(sub(cam.target, cam.position));
  const worldUp = Math.abs(forward[1]) > 0.999 ? [0, 0, 1] : [0, 1, 0];
  const right = norm(cross(forward, worldUp));
  return { forward, right, up: cross(right, forward) };
}

/**
 * The progressive renderer (Renderer, src/renderer.ts:171) on the GPU:
 * accumulation film, one launch per sample, SVGF-style denoiser.
 */
class LucentRenderer {
  /** @param {{width: number, height: number, device?: number}} opts */
  constructor(opts) {
    this.ctx = addon.createContext(opts.device || 0);
    this.width = opts.width;
    this.height = opts.height;
    this.film = addon.createFilm(this.ctx, opts.width, opts.height);
    /** RenderSettings (renderer.ts:42-84): integrator and presentation parts. */
    this.settings = { maxBounces: 8, fireflyClamp: 8, temporal: true, denoise: true, exposure: 1, tonemap: 'aces' };
    this.scene = null;
    this.camera = null;
    this.seed = 0;
    /** The camera the film's samples were traced with, and a pending reprojection's previous camera. */
    this.historyCamera = null;
    this.reprojectFrom = null;
    this.denoisedFrame = -1;
    /** Asynchronous calls run one after another (a context is not re-entrant). */
    this._queue = Promise.resolve();
    this._inflight = 0;
  }

  /** Every synchronous member that touches the context, scene or film
   * refuses to run while an asynchronous call may be using them. */
  _sync() {
    if (this._inflight > 0) throw new Error('an asynchronous call is still running on this renderer');
  }

  _enqueue(start) {
    this._inflight++;
    const run = this._queue.then(start);
    this._queue = run.then(() => undefined, () => undefined);
    return run.then((v) => { this._inflight--; return v; }, (e) => { this._inflight--; throw e; });
  }

  /**
   * loadScene (renderer.ts:279): `packed` is the reference's PackedScene
   * (packScene output), `scene` the Scene it came from (camera, lighting, ground).
   */
  loadScene(packed, scene) {
    this._sync();
    const p = Object.assign({}, packed, {
      groundMaterial: scene.ground && scene.ground.enabled ? packed.groundMaterial : -1,
      groundHeight: scene.ground ? scene.ground.height : 0,
    });
    this.scene = addon.uploadScene(this.ctx, p);
    this.uploadLighting(scene);
    this.uploadCamera(scene.camera);
  }

  /** uploadLighting (renderer.ts:355): {sun, sky, env}; env.map = {width, height, data} or null. */
  uploadLighting(scene) {
    this._sync();
    addon.setLighting(this.ctx, this.scene, { sun: scene.sun, sky: scene.sky, env: scene.env });
    this.reset();
  }

  /**
   * uploadCamera + invalidateView (renderer.ts:424-451): with temporal reuse
   * and samples in the film the next sample reprojects the history,
   * otherwise the film starts over.  `moving` only drives the reference's
   * preview scaling.
   */
  uploadCamera(cam, moving) {
    this._sync();
    this.camera = cam;
    if (this.settings.temporal && this.historyCamera && addon.filmInfo(this.film).samples > 0) {
      this.reprojectFrom = this.historyCamera;
      this.denoisedFrame = -1;
    } else {
      this.reset();
    }
  }

  /** reset (renderer.ts:463): the RNG stream keeps advancing. */
  reset() {
    this._sync();
    addon.resetFilm(this.ctx, this.film);
    this.historyCamera = null;
    this.reprojectFrom = null;
    this.denoisedFrame = -1;
  }

  /** traceSamples (renderer.ts:702): `count` progressive samples, the first reprojecting after a camera move. */
  traceSamples(count) {
    this._sync();
    if (count <= 0) return;
    if (this.reprojectFrom) {
      addon.traceReprojected(this.ctx, this.scene, this.film, this.camera, this.film, this.reprojectFrom, HISTORY_CAP,
        this.settings);
      this.reprojectFrom = null;
      count -= 1;
    }
    if (count > 0) addon.traceSamples(this.ctx, this.scene, this.film, this.camera, this.settings, count);
    this.historyCamera = this.camera;
  }

  /**
   * render (renderer.ts:552) without the GPU budget scheduler: traces
   * `samples` samples and denoises when the denoise weight is above 0.
   */
  render(samples) {
    this._sync();
    this.traceSamples(samples === undefined ? 1 : samples);
    const spp = addon.filmInfo(this.film).samples;
    if (this.settings.denoise && spp > 0 && denoiseWeight(spp) > 0 && this.denoisedFrame !== spp) {
      this.runDenoiser();
      this.denoisedFrame = spp;
    }
    return true;
  }

  /**
   * screenshot (renderer.ts:857) without the PNG encoding: the displayed
   * image (display.frag) as RGBA8, rows bottom to top.
   * @param {number} [viewWidth] defaults to the film size; larger views are Catmull-Rom upsampled
   * @param {number} [viewHeight]
   */
  screenshot(viewWidth, viewHeight) {
    this._sync();
    const spp = addon.filmInfo(this.film).samples;
    const w = this.settings.denoise && spp > 0 && this.denoisedFrame === spp ? denoiseWeight(spp) : 0;
    const vw = viewWidth || this.width, vh = viewHeight || this.height;
    const data = addon.screenshot(this.ctx, this.film, { denoiseWeight: w, exposure: this.settings.exposure,
      tonemap: TONEMAPS[this.settings.tonemap] }, vw, vh);
    return { width: vw, height: vh, data };
  }

  /** benchmark (renderer.ts:745): warm-up sample, then `samples` timed. */
  benchmark(samples) {
    this._sync();
    this.reset();
    this.traceSamples(1);
    let t0 = process.hrtime.bigint();
    this.traceSamples(samples);
    const traceMs = Number(process.hrtime.bigint() - t0) / 1e6;
    t0 = process.hrtime.bigint();
    this.runDenoiser();
    const denoiseMs = Number(process.hrtime.bigint() - t0) / 1e6;
    return { traceMs, mpathsPerSec: (samples * this.width * this.height) / (traceMs / 1000) / 1e6, denoiseMs,
      width: this.width, height: this.height };
  }

  /** runDenoiser (renderer.ts:788): sigmaL 4, sigmaN 64, sigmaP 1.5 by default. */
  runDenoiser(params) {
    this._sync();
    return addon.denoise(this.ctx, this.film, this.camera, params || { sigmaL: 4, sigmaN: 64, sigmaP: 1.5, iterations: 0 });
  }

  /**
   * traceSamples on a worker thread: resolves once the samples are in the
   * film.  Calls queue behind one another.
   * @param {number} count
   */
  traceSamplesAsync(count) {
    return this._enqueue(() => {
      if (count <= 0) return undefined;
      if (this.reprojectFrom) {
        addon.traceReprojected(this.ctx, this.scene, this.film, this.camera, this.film, this.reprojectFrom, HISTORY_CAP,
          this.settings);
        this.reprojectFrom = null;
        count -= 1;
      }
      this.historyCamera = this.camera;
      return count > 0 ? addon.traceSamplesAsync(this.ctx, this.scene, this.film, this.camera, this.settings, count)
        : undefined;
    });
  }

  /** runDenoiser on a worker thread: resolves to {illumVar, radiance}. */
  runDenoiserAsync(params) {
    return this._enqueue(() => addon.denoiseAsync(this.ctx, this.film, this.camera,
      params || { sigmaL: 4, sigmaN: 64, sigmaP: 1.5, iterations: 0 }));
  }

  /** The accumulation targets (rows bottom to top, like readPixels). */
  read() {
    this._sync();
    return addon.readFilm(this.ctx, this.film);
  }

  /** stats (renderer.ts:832), the fields this path knows. */
  stats() {
    this._sync();
    const f = addon.filmInfo(this.film);
    return { samples: f.samples, width: f.width, height: f.height, seed: f.seed };
  }

  /**
   * pickFocus (renderer.ts:499) for a pixel of the film: the focus distance
   * along the view axis of the nearest triangle hit, or null.
   */
  pickFocus(traversal, u, v) {
    const cam = this.camera;
    const { forward, right, up } = cameraBasis(cam);
    const tanHalf = Math.tan((cam.fov * Math.PI) / 360);
    const aspect = this.width / this.height;
    const d = [0, 1, 2].map((i) => forward[i] + right[i] * u * aspect * tanHalf + up[i] * v * tanHalf);
    const l = Math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const dir = [d[0] / l, d[1] / l, d[2] / l];
    const hit = traversal.intersect({ origin: cam.position, dir });
    if (!hit) return null;
    return hit.t * (dir[0] * forward[0] + dir[1] * forward[1] + dir[2] * forward[2]);
  }

  dispose() {
    this._sync();
    this.scene = null;
    this.film = null;
    this.ctx = null;
  }
}

module.exports = { buildBVH, buildBVHAsync, parseHDR, loadGLTF, GpuTraversal, LucentRenderer, cameraBasis, version: addon.version, addon, MISS };
