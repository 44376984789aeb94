// Presentation pass (display.frag, reference src/shaders/display.frag:1-68):
// the running mean, optionally blended with the remodulated denoised
// illumination, Catmull-Rom upsampled when the traced film is smaller than
// the view, exposed, tone mapped (ACES fit, Reinhard or none), encoded to
// sRGB, dithered with interleaved gradient noise and stored as RGBA8 the way
// a UNORM8 framebuffer stores a float colour (clamp to [0,1], x255, round to
// nearest even).  One lane per view pixel; rows bottom-up like the film.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "display.hpp"

namespace lucent {
namespace {

__device__ inline float h2f(uint16_t h) { return __half2float(__ushort_as_half(h)); }

// radiance() (display.frag:32-40).
__device__ inline float3 radiance(const DisplayArgs& a, int px, int py) {
  px = min(max(px, 0), a.render_w - 1);
  py = min(max(py, 0), a.render_h - 1);
  const size_t i = size_t(py) * a.render_w + px;
  const float4 raw = reinterpret_cast<const float4*>(a.color)[i];
  if (a.denoise_weight > 0.0f) {
    const ushort4 ah = reinterpret_cast<const ushort4*>(a.albedo)[i];
    if (h2f(ah.w) >= 0.5f) {
      const float4 d = reinterpret_cast<const float4*>(a.denoised)[i];
      const float w = a.denoise_weight;
      // GLSL mix(raw, d * max(albedo, 1e-3), w) = raw (1-w) + (d max(albedo, 1e-3)) w.
      return make_float3(raw.x * (1.0f - w) + d.x * fmaxf(h2f(ah.x), 1e-3f) * w,
                         raw.y * (1.0f - w) + d.y * fmaxf(h2f(ah.y), 1e-3f) * w,
                         raw.z * (1.0f - w) + d.z * fmaxf(h2f(ah.z), 1e-3f) * w);
    }
  }
  return make_float3(raw.x, raw.y, raw.z);
}

// Narkowicz 2015 fit of the ACES filmic curve (display.frag:16-19).
__device__ inline float aces(float x) {
  return fminf(fmaxf((x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f), 0.0f), 1.0f);
}

// linearToSrgb (display.frag:21-25).
__device__ inline float to_srgb(float c) {
  return c >= 0.0031308f ? 1.055f * powf(c, 1.0f / 2.4f) - 0.055f : c * 12.92f;
}

// Interleaved gradient noise (Jimenez 2014; display.frag:27-30) at the
// fragment centre.
__device__ inline float ign(float x, float y) {
  const float a = 0.06711056f * x + 0.00583715f * y;
  const float b = 52.9829189f * (a - floorf(a));
  return b - floorf(b);
}

__device__ inline uint32_t unorm8(float c) {
  return uint32_t(__float2int_rn(fminf(fmaxf(c, 0.0f), 1.0f) * 255.0f));
}

__global__ __launch_bounds__(256) void k_display(DisplayArgs a) {
  const int x = blockIdx.x * 16 + threadIdx.x % 16;
  const int y = blockIdx.y * 16 + threadIdx.x / 16;
  if (x >= a.view_w || y >= a.view_h) return;
  float3 c;
  if (a.render_w == a.view_w && a.render_h == a.view_h) {
    c = radiance(a, x, y);
  } else {
    // Catmull-Rom upsampling (display.frag:46-60).
    const float sx = (float(x) + 0.5f) * (float(a.render_w) / float(a.view_w)) - 0.5f;
    const float sy = (float(y) + 0.5f) * (float(a.render_h) / float(a.view_h)) - 0.5f;
    const float fx0 = floorf(sx), fy0 = floorf(sy);
    const int ix = int(fx0), iy = int(fy0);
    const float fx = sx - fx0, fy = sy - fy0;
    const float wx[4] = {fx * (-0.5f + fx * (1.0f - 0.5f * fx)), 1.0f + fx * fx * (-2.5f + 1.5f * fx),
                         fx * (0.5f + fx * (2.0f - 1.5f * fx)), fx * fx * (-0.5f + 0.5f * fx)};
    const float wy[4] = {fy * (-0.5f + fy * (1.0f - 0.5f * fy)), 1.0f + fy * fy * (-2.5f + 1.5f * fy),
                         fy * (0.5f + fy * (2.0f - 1.5f * fy)), fy * fy * (-0.5f + 0.5f * fy)};
    c = make_float3(0.0f, 0.0f, 0.0f);
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 4; ++i) {
        const float3 r = radiance(a, ix + i - 1, iy + j - 1);
        const float w = wx[i] * wy[j];
        c.x += r.x * w;
        c.y += r.y * w;
        c.z += r.z * w;
      }
    c = make_float3(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
  }
  c = make_float3(c.x * a.exposure, c.y * a.exposure, c.z * a.exposure);
  if (a.tonemap == 0) {
    c = make_float3(aces(c.x), aces(c.y), aces(c.z));
  } else if (a.tonemap == 1) {
    c = make_float3(c.x / (1.0f + c.x), c.y / (1.0f + c.y), c.z / (1.0f + c.z));
  }
  c = make_float3(to_srgb(fminf(fmaxf(c.x, 0.0f), 1.0f)), to_srgb(fminf(fmaxf(c.y, 0.0f), 1.0f)),
                  to_srgb(fminf(fmaxf(c.z, 0.0f), 1.0f)));
  const float dither = (ign(float(x) + 0.5f, float(y) + 0.5f) - 0.5f) / 255.0f;
  const uint32_t r = unorm8(c.x + dither), g = unorm8(c.y + dither), b = unorm8(c.z + dither);
  reinterpret_cast<uint32_t*>(a.out)[size_t(y) * a.view_w + x] = r | (g << 8) | (b << 16) | (255u << 24);
}

}  // namespace

Status launch_display(const DisplayArgs& a, void* stream) {
  const dim3 grid((a.view_w + 15) / 16, (a.view_h + 15) / 16);
  hipLaunchKernelGGL(k_display, grid, dim3(256), 0, static_cast<hipStream_t>(stream), a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_display launch");
  return LUCENT_OK;
}

}  // namespace lucent
