// SVGF-style denoiser on gfx950 (Renderer.runDenoiser, reference
// src/renderer.ts:788-830): a variance pass (variance.frag) that demodulates
// the running mean by the guide albedo and estimates its variance, then
// à-trous iterations (atrous.frag) with step widths 1, 2, 4, 8, 16, a 5x5 B3
// kernel and normal / plane-distance / luminance edge stops.
//
// Guides are read from the film's FP16 targets; every tap is accumulated in
// float32.  One lane per pixel, 16x8 pixel tiles per block (two wave64s), so
// a block's taps at one offset form two 8-row spans that the L2 serves for
// the neighbouring blocks too.  The variance pass also stores, per pixel, the
// unit normal, the coverage flag and the world position (camera ray through
// the pixel centre times the stored depth): the à-trous taps then read these
// 32 bytes instead of renormalising the neighbour's normal and rebuilding
// its position at every tap, and evaluate their edge-stopping exponentials
// with the hardware log2/exp2 and per-pixel reciprocals.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

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

// primaryDir (atrous.frag:22-26) — the pixel centre's camera ray.
__device__ inline void primary_dir(const Cam& c, int px, int py, int w, int h, float& dx, float& dy, float& dz) {
  const float ux = (float(px) + 0.5f) / float(w) * 2.0f - 1.0f;
  const float uy = (float(py) + 0.5f) / float(h) * 2.0f - 1.0f;
  const float aspect = float(w) / float(h);
  const float sx = ux * aspect * c.tan_half_fov, sy = uy * c.tan_half_fov;
  float x = c.fwd[0] + c.right[0] * sx + c.up[0] * sy;
  float y = c.fwd[1] + c.right[1] * sx + c.up[1] * sy;
  float z = c.fwd[2] + c.right[2] * sx + c.up[2] * sy;
  const float inv = rsqrtf(x * x + y * y + z * z);
  dx = x * inv;
  dy = y * inv;
  dz = z * inv;
}

// variance.frag:15-55.
__global__ __launch_bounds__(kTx* kTy) void k_variance(const float* __restrict__ color,
                                                       const uint16_t* __restrict__ albedo,
                                                       const uint16_t* __restrict__ normal,
                                                       const float* __restrict__ history, float* __restrict__ out,
                                                       float* __restrict__ guides, int w, int h, Cam cam) {
  const int x = blockIdx.x * kTx + threadIdx.x % kTx;
  const int y = blockIdx.y * kTy + threadIdx.x / kTx;
  if (x >= w || y >= h) return;
  const size_t i = size_t(y) * w + x;
  const float4 c = reinterpret_cast<const float4*>(color)[i];
  const float4 a = load_half4(albedo, i);
  float4* o = reinterpret_cast<float4*>(out);
  {
    const float4 np = load_half4(normal, i);
    float nx, ny, nz, dx, dy, dz;
    unit_normal(np.x, np.y, np.z, nx, ny, nz);
    primary_dir(cam, x, y, w, h, dx, dy, dz);
    float4* g = reinterpret_cast<float4*>(guides) + 2 * i;
    g[0] = make_float4(nx, ny, nz, a.w < 0.5f ? 0.0f : 1.0f);
    g[1] = make_float4(cam.pos[0] + dx * np.w, cam.pos[1] + dy * np.w, cam.pos[2] + dz * np.w, np.w);
  }
  if (a.w < 0.5f) {
    o[i] = make_float4(c.x, c.y, c.z, 0.0f);
    return;
  }
  const float ir = c.x / fmaxf(a.x, 1e-3f), ig = c.y / fmaxf(a.y, 1e-3f), ib = c.z / fmaxf(a.z, 1e-3f);
  const float samples = fmaxf(history[i], 1.0f);
  const float la = fmaxf(lum3(a.x, a.y, a.z), 1e-3f);
  float variance;
  if (samples >= 4.0f) {
    const float l = lum3(c.x, c.y, c.z);
    variance = fmaxf(0.0f, c.w - l * l) / (la * la) / samples;
  } else {
    const float4 np = load_half4(normal, i);
    float nx, ny, nz;
    unit_normal(np.x, np.y, np.z, nx, ny, nz);
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
}

__device__ inline float fast_exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// atrous.frag:30-82, with the guides of k_variance.
__global__ __launch_bounds__(kTx* kTy) void k_atrous(const float* __restrict__ in, const float* __restrict__ guides,
                                                     float* __restrict__ out, int w, int h, int step, float tan_half_fov,
                                                     float sigma_l, float sigma_n, float sigma_p) {
  const int x = blockIdx.x * kTx + threadIdx.x % kTx;
  const int y = blockIdx.y * kTy + threadIdx.x / kTx;
  if (x >= w || y >= h) return;
  const size_t i = size_t(y) * w + x;
  const float4* I = reinterpret_cast<const float4*>(in);
  const float4* G = reinterpret_cast<const float4*>(guides);
  const float4 cp = I[i];
  float4* O = reinterpret_cast<float4*>(out);
  const float4 gn = G[2 * i];
  if (gn.w < 0.5f) {
    O[i] = cp;
    return;
  }
  const float4 gp = G[2 * i + 1];
  // footprint = 2 tan(fov/2) depth / H, depth = the stored primary hit distance.
  const float footprint = 2.0f * tan_half_fov * gp.w / float(h);

  // 3x3 binomial blur of the variance (atrous.frag:43-54).
  float var = 0.0f, wv = 0.0f;
  for (int oy = -1; oy <= 1; ++oy)
    for (int ox = -1; ox <= 1; ++ox) {
      const int qx = x + ox, qy = y + oy;
      if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
      const float k = (ox == 0 ? 0.5f : 0.25f) * (oy == 0 ? 0.5f : 0.25f);
      var += I[size_t(qy) * w + qx].w * k;
      wv += k;
    }
  const float inv_denom = 1.0f / (sigma_l * sqrtf(fmaxf(var / wv, 0.0f)) + 1e-6f);
  // 1 / (sigma_p * footprint * step * |offset| + 1e-6) for |offset|^2 = 1, 2, 4, 5, 8.
  const float base = sigma_p * footprint * float(step);
  const float inv_reach[9] = {0.0f,
                              1.0f / (base * 1.0f + 1e-6f),
                              1.0f / (base * 1.41421356f + 1e-6f),
                              0.0f,
                              1.0f / (base * 2.0f + 1e-6f),
                              1.0f / (base * 2.23606798f + 1e-6f),
                              0.0f,
                              0.0f,
                              1.0f / (base * 2.82842712f + 1e-6f)};
  const float lp = lum3(cp.x, cp.y, cp.z);
  const float hk[3] = {0.375f, 0.25f, 0.0625f};
  float sr = cp.x * (hk[0] * hk[0]), sg = cp.y * (hk[0] * hk[0]), sb = cp.z * (hk[0] * hk[0]);
  float svar = cp.w * (hk[0] * hk[0]) * (hk[0] * hk[0]);
  float wsum = hk[0] * hk[0];
  for (int oy = -2; oy <= 2; ++oy)
    for (int ox = -2; ox <= 2; ++ox) {
      if (ox == 0 && oy == 0) continue;
      const int qx = x + ox * step, qy = y + oy * step;
      if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
      const size_t q = size_t(qy) * w + qx;
      const float4 qn = G[2 * q];
      if (qn.w < 0.5f) continue;
      const float4 qp = G[2 * q + 1];
      const float4 cq = I[q];
      const float k = hk[ox < 0 ? -ox : ox] * hk[oy < 0 ? -oy : oy];
      const float c = fmaxf(0.0f, gn.x * qn.x + gn.y * qn.y + gn.z * qn.z);
      const float wn = sigma_n == 0.0f ? 1.0f : __builtin_amdgcn_exp2f(sigma_n * __builtin_amdgcn_logf(c));
      const float plane = fabsf(gn.x * (qp.x - gp.x) + gn.y * (qp.y - gp.y) + gn.z * (qp.z - gp.z));
      const float wp = fast_exp(-plane * inv_reach[ox * ox + oy * oy]);
      const float wl = fast_exp(-fabsf(lp - lum3(cq.x, cq.y, cq.z)) * inv_denom);
      const float wgt = k * wn * wp * wl;
      sr += cq.x * wgt;
      sg += cq.y * wgt;
      sb += cq.z * wgt;
      svar += cq.w * wgt * wgt;
      wsum += wgt;
    }
  O[i] = make_float4(sr / wsum, sg / wsum, sb / wsum, svar / (wsum * wsum));
}

// display.frag:32-40 with weight 1: denoised illumination times max(albedo, 1e-3).
__global__ __launch_bounds__(kTx* kTy) void k_remodulate(const float* __restrict__ in, const uint16_t* __restrict__ albedo,
                                                         float* __restrict__ out, int w, int h) {
  const int x = blockIdx.x * kTx + threadIdx.x % kTx;
  const int y = blockIdx.y * kTy + threadIdx.x / kTx;
  if (x >= w || y >= h) return;
  const size_t i = size_t(y) * w + x;
  const float4 d = reinterpret_cast<const float4*>(in)[i];
  const float4 a = load_half4(albedo, i);
  reinterpret_cast<float4*>(out)[i] =
      make_float4(d.x * fmaxf(a.x, 1e-3f), d.y * fmaxf(a.y, 1e-3f), d.z * fmaxf(a.z, 1e-3f), 1.0f);
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
                     a.width, a.height, cam);
  float* read = a.ping;
  float* write = a.pong;
  for (uint32_t it = 0; it < a.iterations; ++it) {
    hipLaunchKernelGGL(k_atrous, grid, block, 0, stream, read, a.guides, write, a.width, a.height, int(1u << it),
                       a.tan_half_fov, a.sigma_l, a.sigma_n, a.sigma_p);
    float* t = read;
    read = write;
    write = t;
  }
  *a.result = read;
  if (a.remodulated)
    hipLaunchKernelGGL(k_remodulate, grid, block, 0, stream, read, a.albedo, a.remodulated, a.width, a.height);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "denoise launch");
  return LUCENT_OK;
}

}  // namespace lucent
