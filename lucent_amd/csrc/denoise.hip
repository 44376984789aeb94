// SVGF-style denoiser on gfx950 (Renderer.runDenoiser, reference
// src/renderer.ts:788-830): a variance pass (variance.frag) that demodulates
// the running mean by the guide albedo and estimates its variance, then
// à-trous iterations (atrous.frag) with step widths 1, 2, 4, 8, 16, a 5x5 B3
// kernel and normal / plane-distance / luminance edge stops.
//
// Guides are read from the film's FP16 targets; every tap is accumulated in
// float32.  The variance pass stores a 16-byte guide per pixel (unit normal,
// depth over the length of the unnormalised camera ray) and a compact copy
// of the variance.  Each à-trous iteration with a step S <= 16 runs on tiles
// of 64 columns by R rows spaced S apart: every tap of the tile's outputs
// lies on the same dilated grid, so a block stages (64+4S) x (R+4) pixels of
// illumination and guides in LDS once and its 64 R outputs take their 24
// taps from there.  Larger steps read the taps from global memory.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "denoise.hpp"

namespace lucent {
namespace {

constexpr int kTx = 16, kTy = 8;

struct Cam {
  float pos[3], right[3], up[3], fwd[3];
  float tan_half_fov;
};

__device__ inline float lum3(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

__device__ inline float4 load_half4(const uint16_t* p, size_t i) {
  const ushort4 h = reinterpret_cast<const ushort4*>(p)[i];
  return make_float4(__half2float(__ushort_as_half(h.x)), __half2float(__ushort_as_half(h.y)),
                     __half2float(__ushort_as_half(h.z)), __half2float(__ushort_as_half(h.w)));
}

// unitNormal (denoise.glsl:16-19).
__device__ inline void unit_normal(float x, float y, float z, float& ox, float& oy, float& oz) {
  const float l = sqrtf(x * x + y * y + z * z);
  if (l > 1e-6f) {
    ox = x / l;
    oy = y / l;
    oz = z / l;
  } else {
    ox = 0.0f;
    oy = 0.0f;
    oz = 1.0f;
  }
}

// primaryDir (atrous.frag:22-26) before its normalisation: the camera ray
// through the pixel centre as fwd + right*sx + up*sy.
// (The quotients use the hardware reciprocal, 1 ulp.)
__device__ inline void ray_vector(const Cam& c, int px, int py, int w, int h, float& vx, float& vy, float& vz) {
  const float iw = __builtin_amdgcn_rcpf(float(w)), ih = __builtin_amdgcn_rcpf(float(h));
  const float ux = (float(px) + 0.5f) * iw * 2.0f - 1.0f;
  const float uy = (float(py) + 0.5f) * ih * 2.0f - 1.0f;
  const float aspect = float(w) * ih;
  const float sx = ux * aspect * c.tan_half_fov, sy = uy * c.tan_half_fov;
  vx = c.fwd[0] + c.right[0] * sx + c.up[0] * sy;
  vy = c.fwd[1] + c.right[1] * sx + c.up[1] * sy;
  vz = c.fwd[2] + c.right[2] * sx + c.up[2] * sy;
}

// variance.frag:15-55.  Besides the demodulated illumination and its
// variance (out), writes per pixel the à-trous guide (unit normal, tau) with
// tau = depth / |v|, v the unnormalised camera ray through the pixel centre,
// so that the shading point is cam.pos + v * tau; an uncovered pixel's guide is
// all zero.  var_out is a compact copy of the variance channel.
__global__ __launch_bounds__(kTx* kTy) void k_variance(const float* __restrict__ color,
                                                       const uint16_t* __restrict__ albedo,
                                                       const uint16_t* __restrict__ normal,
                                                       const float* __restrict__ history, float* __restrict__ out,
                                                       float* __restrict__ guides, float* __restrict__ var_out, int w,
                                                       int h, Cam cam) {
  const int x = blockIdx.x * kTx + threadIdx.x % kTx;
  const int y = blockIdx.y * kTy + threadIdx.x / kTx;
  if (x >= w || y >= h) return;
  const size_t i = size_t(y) * w + x;
  const float4 c = reinterpret_cast<const float4*>(color)[i];
  const float4 a = load_half4(albedo, i);
  float4* o = reinterpret_cast<float4*>(out);
  float4* g = reinterpret_cast<float4*>(guides) + i;
  if (a.w < 0.5f) {
    *g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    o[i] = make_float4(c.x, c.y, c.z, 0.0f);
    var_out[i] = 0.0f;
    return;
  }
  const float4 np = load_half4(normal, i);
  float nx, ny, nz;
  unit_normal(np.x, np.y, np.z, nx, ny, nz);
  {
    float vx, vy, vz;
    ray_vector(cam, x, y, w, h, vx, vy, vz);
    *g = make_float4(nx, ny, nz, np.w / sqrtf(vx * vx + vy * vy + vz * vz));
  }
  const float ir = c.x / fmaxf(a.x, 1e-3f), ig = c.y / fmaxf(a.y, 1e-3f), ib = c.z / fmaxf(a.z, 1e-3f);
  const float samples = fmaxf(history[i], 1.0f);
  const float la = fmaxf(lum3(a.x, a.y, a.z), 1e-3f);
  float variance;
  if (samples >= 4.0f) {
    const float l = lum3(c.x, c.y, c.z);
    variance = fmaxf(0.0f, c.w - l * l) / (la * la) / samples;
  } else {
    float s1 = 0.0f, s2 = 0.0f, ws = 0.0f;
    for (int dy = -3; dy <= 3; ++dy)
      for (int dx = -3; dx <= 3; ++dx) {
        const int qx = x + dx, qy = y + dy;
        if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
        const size_t q = size_t(qy) * w + qx;
        const float4 aq = load_half4(albedo, q);
        if (aq.w < 0.5f) continue;
        const float4 nq = load_half4(normal, q);
        float ux, uy, uz;
        unit_normal(nq.x, nq.y, nq.z, ux, uy, uz);
        const float wn = powf(fmaxf(0.0f, nx * ux + ny * uy + nz * uz), 32.0f);
        const float wz = fabsf(np.w - nq.w) <= 0.1f * np.w ? 1.0f : 0.0f;  // step(|Δz|, 0.1 z)
        const float wgt = wn * wz;
        const float4 cq = reinterpret_cast<const float4*>(color)[q];
        const float l = lum3(cq.x / fmaxf(aq.x, 1e-3f), cq.y / fmaxf(aq.y, 1e-3f), cq.z / fmaxf(aq.z, 1e-3f));
        s1 += wgt * l;
        s2 += wgt * l * l;
        ws += wgt;
      }
    const float m1 = s1 / fmaxf(ws, 1e-6f);
    variance = fmaxf(0.0f, s2 / fmaxf(ws, 1e-6f) - m1 * m1) / samples;
  }
  o[i] = make_float4(ir, ig, ib, variance);
  var_out[i] = variance;
}

// ---------------------------------------------------------------------------
// One à-trous iteration (atrous.frag:30-82).  Per tap q of centre p the
// reference's weight is
//   k * max(0, n_p·n_q)^sigma_n * exp(-|n_p·(Q-P)| / (reach + 1e-6))
//                              * exp(-|l_p - l_q| / denom),
// evaluated here as one exp2 of the summed base-2 exponents (log2 k is a
// constant per tap).  With Q = C + v_q tau_q and v_q linear in the pixel
// coordinates, n_p·(Q-P) = tau_q (a0 + ox ax + oy ay) - tau_p a0, so with
// a0 + oy ay precomputed per row a tap spends two operations on the plane
// term.  An uncovered or
// outside neighbour has a zero normal, so its normal weight, and with it the
// whole weight, is exp2(-inf) = 0 (for sigma_n = 0 coverage is tested
// explicitly).

struct Centre {
  float nx, ny, nz;  // unit normal
  float ax;          // n_p·v_q = row[oy+2] + ox*ax
  float row[5];      // a0 + oy*ay
  float d;           // tau_p * a0
  float lp;          // luminance of the centre illumination
  float il;          // log2(e) / (sigma_l sqrt(blurred variance) + 1e-6)
  float irl[5];      // log2(e) / (reach + 1e-6) for |offset|^2 = 1, 2, 4, 5, 8
  float sr, sg, sb, svar, wsum;
};

constexpr float kLog2e = 1.4426950408889634f;

// The hardware reciprocal (1 ulp): the edge stops and the final
// normalisation need far less than correctly rounded division, which costs
// ten instructions per quotient here.
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline void centre_begin(Centre& c, float4 cp, float4 gp, float var_blur, const Cam& cam, int x, int y,
                                    int w, int h, int step, float sigma_l, float sigma_p) {
  c.nx = gp.x;
  c.ny = gp.y;
  c.nz = gp.z;
  float vx, vy, vz;
  ray_vector(cam, x, y, w, h, vx, vy, vz);
  const float a0 = c.nx * vx + c.ny * vy + c.nz * vz;
  const float inv_h = rcp(float(h));
  const float dsx = 2.0f * inv_h * cam.tan_half_fov * float(step);  // 2/w * (w/h)
  const float dsy = 2.0f * inv_h * cam.tan_half_fov * float(step);
  c.ax = (c.nx * cam.right[0] + c.ny * cam.right[1] + c.nz * cam.right[2]) * dsx;
  const float ay = (c.nx * cam.up[0] + c.ny * cam.up[1] + c.nz * cam.up[2]) * dsy;
  c.row[0] = a0 - 2.0f * ay;
  c.row[1] = a0 - ay;
  c.row[2] = a0;
  c.row[3] = a0 + ay;
  c.row[4] = a0 + 2.0f * ay;
  c.d = gp.w * a0;
  // footprint = 2 tan(fov/2) depth / H with depth = tau_p |v_p|.  The
  // square roots here are the hardware's (1 ulp, as GLSL's sqrt allows).
  const float depth = gp.w * __builtin_amdgcn_sqrtf(vx * vx + vy * vy + vz * vz);
  const float footprint = 2.0f * cam.tan_half_fov * depth * inv_h;
  const float base = sigma_p * footprint * float(step);
  c.irl[0] = kLog2e * rcp(base * 1.0f + 1e-6f);
  c.irl[1] = kLog2e * rcp(base * 1.41421356f + 1e-6f);
  c.irl[2] = kLog2e * rcp(base * 2.0f + 1e-6f);
  c.irl[3] = kLog2e * rcp(base * 2.23606798f + 1e-6f);
  c.irl[4] = kLog2e * rcp(base * 2.82842712f + 1e-6f);
  c.il = kLog2e * rcp(sigma_l * __builtin_amdgcn_sqrtf(fmaxf(var_blur, 0.0f)) + 1e-6f);
  c.lp = lum3(cp.x, cp.y, cp.z);
  constexpr float h0 = 0.375f;
  c.sr = cp.x * (h0 * h0);
  c.sg = cp.y * (h0 * h0);
  c.sb = cp.z * (h0 * h0);
  c.svar = cp.w * (h0 * h0) * (h0 * h0);
  c.wsum = h0 * h0;
}

__host__ __device__ constexpr float log2_tap(int o) { return o == 0 ? -1.4150374992788437f : o == 1 ? -2.0f : -4.0f; }
__host__ __device__ constexpr int iabs(int v) { return v < 0 ? -v : v; }
__host__ __device__ constexpr int reach_slot(int ox, int oy) {
  return ox * ox + oy * oy == 1 ? 0 : ox * ox + oy * oy == 2 ? 1 : ox * ox + oy * oy == 4 ? 2 : ox * ox + oy * oy == 5 ? 3 : 4;
}

template <int OX, int OY, bool kNormalStop>
__device__ __forceinline__ void tap(Centre& c, float4 cq, float4 gq, float lq, float sigma_n) {
  constexpr float log2k = log2_tap(iabs(OX)) + log2_tap(iabs(OY));
  const float dot = c.nx * gq.x + c.ny * gq.y + c.nz * gq.z;
  float e;
  if constexpr (kNormalStop) {
    // max(0, n_p·n_q) is not taken: a negative dot gives a NaN exponent,
    // and the clamp of the weight to [0, 1] (the exp's output modifier) maps
    // NaN to 0, the weight that max(0, .)^sigma_n gives.
    e = sigma_n * __builtin_amdgcn_logf(dot) + log2k;
  } else {
    const bool covered = gq.x != 0.0f || gq.y != 0.0f || gq.z != 0.0f;
    e = covered ? log2k : -INFINITY;
  }
  float nv = c.row[OY + 2];
  if constexpr (OX == 1) nv += c.ax;
  if constexpr (OX == -1) nv -= c.ax;
  if constexpr (OX == 2 || OX == -2) nv = fmaf(float(OX), c.ax, nv);
  const float plane = fabsf(fmaf(gq.w, nv, -c.d));
  const float dl = fabsf(lq - c.lp);
  e -= plane * c.irl[reach_slot(OX, OY)];
  e -= dl * c.il;
  // A tap's weight is at most its kernel weight (<= 1/4), so the clamp
  // changes no finite weight.
  const float wgt = __builtin_amdgcn_fmed3f(__builtin_amdgcn_exp2f(e), 0.0f, 1.0f);
  c.sr += cq.x * wgt;
  c.sg += cq.y * wgt;
  c.sb += cq.z * wgt;
  c.svar += cq.w * (wgt * wgt);
  c.wsum += wgt;
}

__device__ inline void centre_store(const Centre& c, size_t i, float4* out, float* var_out, const uint16_t* albedo,
                                    float4* remodulated) {
  const float inv = rcp(c.wsum);
  const float4 r = make_float4(c.sr * inv, c.sg * inv, c.sb * inv, c.svar * (inv * inv));
  out[i] = r;
  var_out[i] = r.w;
  if (remodulated) {
    // display.frag:32-40 with weight 1: illumination times max(albedo, 1e-3).
    const float4 a = load_half4(albedo, i);
    remodulated[i] = make_float4(r.x * fmaxf(a.x, 1e-3f), r.y * fmaxf(a.y, 1e-3f), r.z * fmaxf(a.z, 1e-3f), 1.0f);
  }
}

__device__ inline void passthrough(float4 cp, size_t i, float4* out, float* var_out, const uint16_t* albedo,
                                   float4* remodulated) {
  out[i] = cp;
  var_out[i] = cp.w;
  if (remodulated) {
    const float4 a = load_half4(albedo, i);
    remodulated[i] = make_float4(cp.x * fmaxf(a.x, 1e-3f), cp.y * fmaxf(a.y, 1e-3f), cp.z * fmaxf(a.z, 1e-3f), 1.0f);
  }
}

// 3x3 binomial blur of the variance over the in-image neighbours
// (atrous.frag:43-54); v[dy+1][dx+1] are the neighbours' variances (zero
// outside the image).
__device__ inline float blur_variance(const float v[3][3], int x, int y, int w, int h) {
  if (x > 0 && y > 0 && x + 1 < w && y + 1 < h) {
    // Inside the image the weights sum to 1: the same sum as below, whose
    // normalisation is then a multiplication by rcp(1) = 1.
    constexpr float k[3] = {0.25f, 0.5f, 0.25f};
    float var = 0.0f;
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) var += v[dy][dx] * (k[dx] * k[dy]);
    return var;
  }
  const float kx[3] = {x > 0 ? 0.25f : 0.0f, 0.5f, x + 1 < w ? 0.25f : 0.0f};
  const float ky[3] = {y > 0 ? 0.25f : 0.0f, 0.5f, y + 1 < h ? 0.25f : 0.0f};
  float var = 0.0f;
#pragma unroll
  for (int dy = 0; dy < 3; ++dy)
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) var += v[dy][dx] * (kx[dx] * ky[dy]);
  return var * rcp((kx[0] + kx[1] + kx[2]) * (ky[0] + ky[1] + ky[2]));
}

// Tile of the LDS kernel: 64 consecutive columns times R rows spaced by the
// step (rows y0, y0+S, ..., y0+(R-1)S), so that every tap of every output
// lies on the tile's own dilated grid: R+4 staged rows of 64+4S columns.
// 4 waves; lane = column, wave = row.  Four rows stage 2x (step 1) to 4x
// (step 16) the pixels they filter but leave LDS for 4-6 blocks per CU,
// which measured faster than taller tiles with less halo.
constexpr int kTileW = 64, kWaves = 4;
constexpr uint32_t kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

#ifndef LUCENT_ATROUS_ROWS
#define LUCENT_ATROUS_ROWS 4
#endif
__host__ __device__ constexpr int tile_rows(int) { return LUCENT_ATROUS_ROWS; }

template <int S>
struct AtrousTile {
  static constexpr int R = tile_rows(S);
  static constexpr int cols = kTileW + 4 * S;
  static constexpr int rows = R + 4;
  static constexpr int vcols = kTileW + 2;
};

// Tile t in the XCD-aware order: column tiles fastest, then the row bands
// of one residue class mod S, then the classes.
template <int S>
__device__ inline void tile_origin(uint32_t t, int w, int h, int& x0, int& y0) {
  using T = AtrousTile<S>;
  const uint32_t cols_n = (uint32_t(w) + kTileW - 1) / kTileW;
  const uint32_t bands = (uint32_t((h + S - 1) / S) + T::R - 1) / T::R;
  const uint32_t col = t % cols_n, rest = t / cols_n, band = rest % bands, cls = rest / bands;
  x0 = int(col) * kTileW;
  y0 = int(cls) + int(band) * T::R * S;
}

// A thread's share of one tile's staged pixels, held in registers between
// the global loads and the LDS stores.  Each wave stages whole rows (staged
// rows wave, wave+4, ...; variance rows likewise), a lane one column of each
// in runs of 64: the row's index, bounds and base address are the wave's
// (scalar), and a lane adds only its column.
template <int S>
struct TileRegs {
  using T = AtrousTile<S>;
  static constexpr int kRows = (T::rows + kWaves - 1) / kWaves;       // staged rows per wave
  static constexpr int kRuns = (T::cols + 63) / 64;                   // 64-column runs per row
  static constexpr int kVRows = (T::R * 3 + kWaves - 1) / kWaves;     // variance rows per wave
  static constexpr int kVRuns = (T::vcols + 63) / 64;
  float4 c[kRows][kRuns], g[kRows][kRuns];
  float v[kVRows][kVRuns];
};

template <int S>
__device__ __forceinline__ void tile_load(TileRegs<S>& r, const float4* __restrict__ in,
                                          const float4* __restrict__ guides, const float* __restrict__ var_in, int w,
                                          int h, int x0, int y0, int wave, int lane) {
  using T = AtrousTile<S>;
  using TR = TileRegs<S>;
#pragma unroll
  for (int k = 0; k < TR::kRows; ++k) {
    const int j = wave + k * kWaves;
    const int gy = y0 + (j - 2) * S;
    const bool row_ok = j < T::rows && gy >= 0 && gy < h;
    const size_t row = size_t(row_ok ? gy : 0) * size_t(w);
    const float4* crow = in + row;
    const float4* grow = guides + row;
#pragma unroll
    for (int m = 0; m < TR::kRuns; ++m) {
      const int cc = lane + 64 * m;
      const int gx = x0 - 2 * S + cc;
      float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g = c;
      if (row_ok && cc < T::cols && uint32_t(gx) < uint32_t(w)) {
        c = crow[gx];
        g = grow[gx];
      }
      r.c[k][m] = c;
      r.g[k][m] = g;
    }
  }
#pragma unroll
  for (int k = 0; k < TR::kVRows; ++k) {
    const int j = wave + k * kWaves;
    const int gy = y0 + (j / 3) * S + (j % 3) - 1;
    const bool row_ok = j < T::R * 3 && gy >= 0 && gy < h;
    const float* vrow = var_in + size_t(row_ok ? gy : 0) * size_t(w);
#pragma unroll
    for (int m = 0; m < TR::kVRuns; ++m) {
      const int cc = lane + 64 * m;
      const int gx = x0 - 1 + cc;
      r.v[k][m] = row_ok && cc < T::vcols && uint32_t(gx) < uint32_t(w) ? vrow[gx] : 0.0f;
    }
  }
}

template <int S>
__device__ __forceinline__ void tile_store(const TileRegs<S>& r, float4* sc, float4* sg, float* sl, float* sv,
                                           int wave, int lane) {
  using T = AtrousTile<S>;
  using TR = TileRegs<S>;
#pragma unroll
  for (int k = 0; k < TR::kRows; ++k) {
    const int j = wave + k * kWaves;
#pragma unroll
    for (int m = 0; m < TR::kRuns; ++m) {
      const int cc = lane + 64 * m;
      if (j < T::rows && cc < T::cols) {
        const int e = j * T::cols + cc;
        sc[e] = r.c[k][m];
        sg[e] = r.g[k][m];
        sl[e] = lum3(r.c[k][m].x, r.c[k][m].y, r.c[k][m].z);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < TR::kVRows; ++k) {
    const int j = wave + k * kWaves;
#pragma unroll
    for (int m = 0; m < TR::kVRuns; ++m) {
      const int cc = lane + 64 * m;
      if (j < T::R * 3 && cc < T::vcols) sv[j * T::vcols + cc] = r.v[k][m];
    }
  }
}

#ifndef LUCENT_ATROUS_TILES_PER_BLOCK
#define LUCENT_ATROUS_TILES_PER_BLOCK 1
#endif
constexpr uint32_t kTilesPerBlock = LUCENT_ATROUS_TILES_PER_BLOCK;

// Each block filters a run of tiles (one by default, kTilesPerBlock), staging
// each (global loads into registers, then LDS) before it filters it.  The
// staging latency hides behind other blocks' taps: without a register copy
// of the next tile held across the taps the kernels need 68-82 VGPRs, so 6
// waves per SIMD fit where LDS allows (a prefetching version needed 110-139
// and ran 3-4 waves, 3.5 % slower).  Tiles are dealt in the XCD-aware order:
// workgroups go to the 8 XCDs round robin (b mod 8); XCD k's blocks take
// interleaved tiles of a contiguous run of the order, so the halos that
// neighbouring tiles share stay in its own L2.
#ifndef LUCENT_ATROUS_WAVES_PER_EU
#define LUCENT_ATROUS_WAVES_PER_EU 1
#endif
template <int S, bool kNormalStop>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(LUCENT_ATROUS_WAVES_PER_EU))) void k_atrous_tile(const float4* __restrict__ in,
                                                             const float4* __restrict__ guides,
                                                             const float* __restrict__ var_in,
                                                             float4* __restrict__ out, float* __restrict__ var_out,
                                                             const uint16_t* __restrict__ albedo,
                                                             float4* __restrict__ remodulated, int w, int h, Cam cam,
                                                             float sigma_l, float sigma_n, float sigma_p,
                                                             uint32_t n_tiles) {
  using T = AtrousTile<S>;
  __shared__ float4 sc[T::rows * T::cols];
  __shared__ float4 sg[T::rows * T::cols];
  __shared__ float sl[T::rows * T::cols];  // luminance of the staged illumination
  __shared__ float sv[T::R * 3 * T::vcols];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // This block's tiles: XCD k (the blocks b = k mod 8, `mine` of them)
  // owns a contiguous run [first, end) of the order, in proportion to its
  // blocks, dealt to its blocks round robin.
  const uint32_t nb = gridDim.x, b = blockIdx.x, xcd = b % kXcds;
  uint32_t before = 0;
  for (uint32_t k = 0; k < xcd; ++k) before += (nb - k + kXcds - 1) / kXcds;
  const uint32_t mine = (nb - xcd + kXcds - 1) / kXcds;
  const uint32_t first = uint32_t(uint64_t(n_tiles) * before / nb);
  const uint32_t end = uint32_t(uint64_t(n_tiles) * (before + mine) / nb);
  const uint32_t blocks_in_xcd = mine;
  uint32_t t = first + b / kXcds;
  if (t >= end) return;
  int x0, y0;
  for (;;) {
    tile_origin<S>(t, w, h, x0, y0);
    {
      TileRegs<S> regs;
      tile_load<S>(regs, in, guides, var_in, w, h, x0, y0, wave, lane);
      tile_store<S>(regs, sc, sg, sl, sv, wave, lane);
    }
    __syncthreads();
    const int cx0 = x0, cy0 = y0;
    const uint32_t next = t + blocks_in_xcd;
    const int x = cx0 + lane;
    for (int r = wave; r < T::R; r += kWaves) {
      const int y = cy0 + r * S;
      if (x >= w || y >= h) continue;
      const size_t i = size_t(y) * w + x;
      const int ci = (r + 2) * T::cols + lane + 2 * S;
      const float4 cp = sc[ci];
      const float4 gp = sg[ci];
      if (gp.x == 0.0f && gp.y == 0.0f && gp.z == 0.0f) {
        passthrough(cp, i, out, var_out, albedo, remodulated);
        continue;
      }
      float v[3][3];
#pragma unroll
      for (int dy = 0; dy < 3; ++dy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) v[dy][dx] = sv[(r * 3 + dy) * T::vcols + lane + dx];
      Centre c;
      centre_begin(c, cp, gp, blur_variance(v, x, y, w, h), cam, x, y, w, h, S, sigma_l, sigma_p);
      const float4* pc = sc + ci;
      const float4* pg = sg + ci;
      const float* pl = sl + ci;
#define LUCENT_TAP(OX, OY) \
  tap<OX, OY, kNormalStop>(c, pc[OY * T::cols + OX * S], pg[OY * T::cols + OX * S], pl[OY * T::cols + OX * S], sigma_n)
#define LUCENT_ROW(OY) \
  LUCENT_TAP(-2, OY);  \
  LUCENT_TAP(-1, OY);  \
  LUCENT_TAP(0, OY);   \
  LUCENT_TAP(1, OY);   \
  LUCENT_TAP(2, OY)
      LUCENT_ROW(-2);
      LUCENT_ROW(-1);
      LUCENT_TAP(-2, 0);
      LUCENT_TAP(-1, 0);
      LUCENT_TAP(1, 0);
      LUCENT_TAP(2, 0);
      LUCENT_ROW(1);
      LUCENT_ROW(2);
#undef LUCENT_ROW
#undef LUCENT_TAP
      centre_store(c, i, out, var_out, albedo, remodulated);
    }
    if (next >= end) break;
    t = next;
    __syncthreads();  // every wave is done with this tile's LDS
  }
}

// Steps above the tile kernel's (iterations beyond 5): one lane per pixel,
// taps read from global memory, outside taps skipped.
template <bool kNormalStop>
__global__ __launch_bounds__(kTx* kTy) void k_atrous_direct(const float4* __restrict__ in,
                                                            const float4* __restrict__ guides,
                                                            const float* __restrict__ var_in,
                                                            float4* __restrict__ out, float* __restrict__ var_out,
                                                            const uint16_t* __restrict__ albedo,
                                                            float4* __restrict__ remodulated, int w, int h, int step,
                                                            Cam cam, float sigma_l, float sigma_n, float sigma_p) {
  const int x = blockIdx.x * kTx + threadIdx.x % kTx;
  const int y = blockIdx.y * kTy + threadIdx.x / kTx;
  if (x >= w || y >= h) return;
  const size_t i = size_t(y) * w + x;
  const float4 cp = in[i];
  const float4 gp = guides[i];
  if (gp.x == 0.0f && gp.y == 0.0f && gp.z == 0.0f) {
    passthrough(cp, i, out, var_out, albedo, remodulated);
    return;
  }
  float v[3][3];
  for (int dy = 0; dy < 3; ++dy)
    for (int dx = 0; dx < 3; ++dx) {
      const int qx = x + dx - 1, qy = y + dy - 1;
      v[dy][dx] = qx >= 0 && qx < w && qy >= 0 && qy < h ? var_in[size_t(qy) * w + qx] : 0.0f;
    }
  Centre c;
  centre_begin(c, cp, gp, blur_variance(v, x, y, w, h), cam, x, y, w, h, step, sigma_l, sigma_p);
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#define LUCENT_TAP(OX, OY)                                                                        \
  {                                                                                               \
    const int qx = x + OX * step, qy = y + OY * step;                                             \
    const bool inside = qx >= 0 && qx < w && qy >= 0 && qy < h;                                   \
    const size_t q = inside ? size_t(qy) * w + qx : i;                                            \
    const float4 cq = inside ? in[q] : zero;                                                      \
    tap<OX, OY, kNormalStop>(c, cq, inside ? guides[q] : zero, lum3(cq.x, cq.y, cq.z), sigma_n);  \
  }
#define LUCENT_ROW(OY) \
  LUCENT_TAP(-2, OY);  \
  LUCENT_TAP(-1, OY);  \
  LUCENT_TAP(0, OY);   \
  LUCENT_TAP(1, OY);   \
  LUCENT_TAP(2, OY)
  LUCENT_ROW(-2);
  LUCENT_ROW(-1);
  LUCENT_TAP(-2, 0);
  LUCENT_TAP(-1, 0);
  LUCENT_TAP(1, 0);
  LUCENT_TAP(2, 0);
  LUCENT_ROW(1);
  LUCENT_ROW(2);
#undef LUCENT_ROW
#undef LUCENT_TAP
  centre_store(c, i, out, var_out, albedo, remodulated);
}

struct PassArgs {
  const float4* in;
  const float4* guides;
  const float* var_in;
  float4* out;
  float* var_out;
  const uint16_t* albedo;
  float4* remodulated;
  int w, h, step;
  Cam cam;
  float sigma_l, sigma_n, sigma_p;
  int compute_units;
};

template <int S, bool kNormalStop>
void launch_tile(const PassArgs& p, hipStream_t stream) {
  const int rows_per_class = (p.h + S - 1) / S;
  const int R = AtrousTile<S>::R;
  const uint32_t n_tiles = uint32_t((p.w + kTileW - 1) / kTileW) * S * ((rows_per_class + R - 1) / R);
  // One tile per block: the hardware deals blocks to CUs as others finish,
  // which balances the CUs better than persistent blocks with a fixed run of
  // tiles each (measured at 1080p: 0.218 against 0.227 ms per denoise).
  auto kern = k_atrous_tile<S, kNormalStop>;
  const uint32_t grid = (n_tiles + kTilesPerBlock - 1) / kTilesPerBlock;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kWaves * 64), 0, stream, p.in, p.guides, p.var_in, p.out, p.var_out,
                     p.albedo, p.remodulated, p.w, p.h, p.cam, p.sigma_l, p.sigma_n, p.sigma_p, n_tiles);
}

template <bool kNormalStop>
void launch_pass(const PassArgs& p, hipStream_t stream) {
  switch (p.step) {
    case 1: return launch_tile<1, kNormalStop>(p, stream);
    case 2: return launch_tile<2, kNormalStop>(p, stream);
    case 4: return launch_tile<4, kNormalStop>(p, stream);
    case 8: return launch_tile<8, kNormalStop>(p, stream);
    case 16: return launch_tile<16, kNormalStop>(p, stream);
    default: {
      const dim3 grid((p.w + kTx - 1) / kTx, (p.h + kTy - 1) / kTy);
      hipLaunchKernelGGL((k_atrous_direct<kNormalStop>), grid, dim3(kTx * kTy), 0, stream, p.in, p.guides, p.var_in,
                         p.out, p.var_out, p.albedo, p.remodulated, p.w, p.h, p.step, p.cam, p.sigma_l, p.sigma_n,
                         p.sigma_p);
    }
  }
}

}  // namespace

Status launch_denoise(const DenoiseArgs& a, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const dim3 grid((a.width + kTx - 1) / kTx, (a.height + kTy - 1) / kTy);
  const dim3 block(kTx * kTy);
  Cam cam;
  for (int k = 0; k < 3; ++k) {
    cam.pos[k] = a.cam_pos[k];
    cam.right[k] = a.cam_right[k];
    cam.up[k] = a.cam_up[k];
    cam.fwd[k] = a.cam_fwd[k];
  }
  cam.tan_half_fov = a.tan_half_fov;
  hipLaunchKernelGGL(k_variance, grid, block, 0, stream, a.color, a.albedo, a.normal, a.history, a.ping, a.guides,
                     a.var_a, a.width, a.height, cam);
  PassArgs p;
  p.guides = reinterpret_cast<const float4*>(a.guides);
  p.albedo = a.albedo;
  p.w = a.width;
  p.h = a.height;
  p.cam = cam;
  p.sigma_l = a.sigma_l;
  p.sigma_n = a.sigma_n;
  p.sigma_p = a.sigma_p;
  p.compute_units = a.compute_units;
  float* read = a.ping;
  float* write = a.pong;
  float* var_read = a.var_a;
  float* var_write = a.var_b;
  for (uint32_t it = 0; it < a.iterations; ++it) {
    p.in = reinterpret_cast<const float4*>(read);
    p.out = reinterpret_cast<float4*>(write);
    p.var_in = var_read;
    p.var_out = var_write;
    p.step = int(1u << it);
    p.remodulated = it + 1 == a.iterations ? reinterpret_cast<float4*>(a.remodulated) : nullptr;
    if (a.sigma_n == 0.0f)
      launch_pass<false>(p, stream);
    else
      launch_pass<true>(p, stream);
    std::swap(read, write);
    std::swap(var_read, var_write);
  }
  *a.result = read;
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "denoise launch");
  return LUCENT_OK;
}

}  // namespace lucent
