// Wavefront path tracer: device buffers and launch (wavefront.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.hpp"
#include "pt_types.hpp"

namespace lucent {

// Queue control block (device memory).
struct WfCtl {
  uint32_t parity;    // the queue the current iteration consumes
  uint32_t count[2];  // path queue lengths
  uint32_t scount[2]; // shadow queue lengths
  uint32_t cursor[3]; // work cursors of the extend, shadow and shade passes
  uint32_t live;      // paths consumed by the current iteration
  uint32_t slots;
};

// Structure-of-arrays path state, one entry per slot (device pointers).
struct WfBuffers {
  WfCtl* ctl = nullptr;
  float4* ray_o = nullptr;  // origin
  float4* ray_d = nullptr;  // direction, w = BSDF pdf of the sample that made it
  float4* beta = nullptr;   // throughput
  float4* rad = nullptr;    // radiance so far
  float4* tint = nullptr;   // guide tint, w = guide depth
  float4* galb = nullptr;   // guide albedo, w = coverage
  float4* gnrm = nullptr;   // guide normal
  float4* hit = nullptr;    // closest hit: t, u, v, (kind << 30 | primitive) bits
  float4* nee = nullptr;    // 3 per slot: light-sample contribution, w = -1 none, 0 pending, 1 unoccluded
  float4* shadow_o = nullptr;  // 3 per slot: shadow ray origin, w = extent
  float4* shadow_d = nullptr;  // 3 per slot: shadow ray direction
  uint32_t* rng = nullptr;
  uint32_t* sample = nullptr;  // index of the sample the slot is tracing
  uint32_t* flags = nullptr;   // specular, guide done, terminating; bounce << 8
  uint32_t* queue[2] = {nullptr, nullptr};   // live path slots
  uint32_t* squeue[2] = {nullptr, nullptr};  // pending light samples: slot * 3 + which
  uint32_t* spill = nullptr;
  uint32_t spill_depth = 0;
};

struct WavefrontLaunch {
  uint32_t waves_per_cu = 0;  // persistent waves per CU of every pass (0: 20 extend, 24 shadow, 8 shade)
  uint32_t check_every = 0;   // iterations between progress reads (0: 16, 4 in the tail)
};

size_t wavefront_bytes(uint32_t slots);
// it.spp samples of every pixel of it.film, seeds from it.frame.seed; fills it.tiles_x.
Status launch_wavefront(Context* ctx, TileItem& it, const WavefrontLaunch& cfg, void* stream);

}  // namespace lucent
