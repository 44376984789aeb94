// Progressive path tracing on gfx950: one launch per sample over the whole
// film (Renderer.traceSamples, reference src/renderer.ts:702-737), one lane
// per pixel, each wave64 covering an 8x8 pixel tile so that neighbouring
// primary rays share BVH nodes in the caches.  The integrator itself is
// pt_device.hpp (trace.frag semantics in float32).
#include <hip/hip_runtime.h>

#include "pt_device.hpp"

namespace lucent {
namespace {

constexpr int kPtBlock = 64;
constexpr int kPtLds = 16;
constexpr int kTile = 8;

__global__ __launch_bounds__(kPtBlock) void k_trace_sample(PtScene sc, PtFrame f, PtFilm film, uint32_t* spill,
                                                          uint32_t spill_depth, uint32_t tiles_x) {
  __shared__ uint32_t lds[kPtLds * kPtBlock];
  const uint32_t tile = blockIdx.x;
  const uint32_t x = (tile % tiles_x) * kTile + (threadIdx.x & (kTile - 1));
  const uint32_t y = (tile / tiles_x) * kTile + (threadIdx.x / kTile);
  if (x >= f.width || y >= f.height) return;
  const uint32_t idx = y * f.width + x;
  pt::Rng rng = pt::seed_rng(idx, f.seed);
  pt::v3 ro, rd;
  pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  pt::Stack<kPtLds, kPtBlock> st{lds + threadIdx.x, spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
                                 0, 0};
  pt::Guide g;
  const pt::v3 c = pt::trace_path(sc, f, ro, rd, rng, st, g);
  pt::accumulate(film, idx, c, g);
}

}  // namespace

Status launch_trace_sample(Context* ctx, const PtScene& sc, const PtFrame& f, const PtFilm& film, void* stream) {
  const uint32_t tiles_x = (f.width + kTile - 1) / kTile;
  const uint32_t tiles_y = (f.height + kTile - 1) / kTile;
  const uint64_t blocks = uint64_t(tiles_x) * tiles_y;
  if (blocks == 0) return LUCENT_OK;
  const uint32_t depth = sc.bvh.stack_depth + 1;
  Status s = ensure_spill(ctx, size_t(blocks) * kPtBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  hipLaunchKernelGGL(k_trace_sample, dim3(uint32_t(blocks)), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), sc, f,
                     film, static_cast<uint32_t*>(ctx->spill), depth, tiles_x);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_sample launch");
  return LUCENT_OK;
}

}  // namespace lucent
