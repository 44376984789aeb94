// Types of node/lucent.js (the Node host facade over the N-API addon).
export interface BVH {
  nodes: Float32Array;
  nodeCount: number;
  triIndices: Uint32Array;
  bounds: { min: [number, number, number]; max: [number, number, number] };
}
export interface BVHOptions { maxLeafSize?: number; bins?: number; }
export interface BVHHit { t: number; triIndex: number; u: number; v: number; }
export interface HitBatch { t: Float64Array; triIndex: Uint32Array; u: Float64Array; v: Float64Array; }
export interface Ray { origin: number[]; dir: number[]; }
export declare const MISS: number;
export declare function version(): string;
export declare function buildBVH(positions: Float32Array, options?: BVHOptions): BVH;
export declare class GpuTraversal {
  constructor(bvh: BVH, positions: Float32Array, device?: number);
  intersectBatch(origins: Float32Array | Float64Array, dirs: Float32Array | Float64Array, tMin?: number,
    tMax?: number, opts?: { anyHit?: boolean; glsl?: boolean }): HitBatch;
  intersect(ray: Ray, tMin?: number, tMax?: number): BVHHit | null;
}
export interface Camera { position: number[]; target: number[]; fov: number; aperture: number; focusDistance: number; }
export interface FilmTargets { color: Float32Array; albedo: Float32Array; normal: Float32Array; history: Float32Array; }
export declare function buildBVHAsync(positions: Float32Array, options?: BVHOptions): Promise<BVH>;
export declare function parseHDR(buffer: ArrayBuffer | Uint8Array): { width: number; height: number; data: Float32Array };
export interface GLTFImage { data: Uint8Array; mimeType: string; name: string; }
export interface GLTFAsset {
  meshes: { mesh: { positions: Float32Array; normals: Float32Array; uvs?: Float32Array }; material: number }[];
  materials: Record<string, unknown>[];
  images: GLTFImage[];
  bounds: { min: number[]; max: number[] };
  triangleCount: number;
  warnings: string[];
}
export declare function loadGLTF(source: ArrayBuffer | Uint8Array,
  resolve: (uri: string) => Promise<ArrayBuffer | Uint8Array> | ArrayBuffer | Uint8Array): Promise<GLTFAsset>;

export declare class LucentRenderer {
  constructor(opts: { width: number; height: number; device?: number });
  settings: { maxBounces: number; fireflyClamp: number; temporal: boolean; denoise: boolean; exposure: number;
    tonemap: 'aces' | 'reinhard' | 'none';
    /* launch structure (every one gives the same film): */
    perSampleLaunch?: boolean; wavefront?: boolean; inlineShadows?: boolean; occupancy?: number; wavesPerCU?: number;
    /* per-sample hand-out shape (0 = default; tile 0 x 0 sizes each film's jobs for jobPaths paths): */
    tileWidth?: number; tileHeight?: number; shadeBatch?: number; traceTurns?: number; jobPaths?: number };
  loadScene(packed: object, scene: object): void;
  uploadLighting(scene: object): void;
  uploadCamera(cam: Camera, moving?: boolean): void;
  reset(): void;
  traceSamples(count: number): void;
  traceSamplesAsync(count: number): Promise<void>;
  runDenoiserAsync(params?: { sigmaL: number; sigmaN: number; sigmaP: number; iterations: number }):
    Promise<{ illumVar: Float32Array; radiance: Float32Array }>;
  render(samples?: number): boolean;
  screenshot(viewWidth?: number, viewHeight?: number): { width: number; height: number; data: Uint8Array };
  benchmark(samples: number): { traceMs: number; mpathsPerSec: number; denoiseMs: number; width: number; height: number };
  runDenoiser(params?: { sigmaL: number; sigmaN: number; sigmaP: number; iterations: number }):
    { illumVar: Float32Array; radiance: Float32Array };
  read(): FilmTargets;
  stats(): { samples: number; width: number; height: number; seed: number };
  pickFocus(traversal: GpuTraversal, u: number, v: number): number | null;
  dispose(): void;
}
