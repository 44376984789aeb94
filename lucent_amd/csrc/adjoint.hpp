// Path-replay adjoint launch arguments.
#pragma once

#include <cstdint>

#include "internal.hpp"
#include "pt_types.hpp"

namespace lucent {

struct AdjointArgs {
  const PtScene* scene = nullptr;
  PtFrame frame;
  uint32_t spp = 1;
  uint32_t seed0 = 0;
  const float* dl_di = nullptr;     // dLoss/dImage, float4 per pixel (rows bottom-up)
  const float* recorded_radiance = nullptr;  // optional: float4 per pixel per sample from the forward pass
  double* grad_materials = nullptr; // material_count * 5 (albedo rgb, roughness, metallic)
  double* grad_env = nullptr;       // env_w * env_h * 3, or null
  void* scratch = nullptr;          // adjoint_scratch_bytes(...)
  uint32_t occupancy = 0;           // register-allocation variant of the replay kernel: 4 (default; any value but 2 or 3), 2 or 3
  uint32_t tile_w = 0, tile_h = 0;  // job tile of the replay's hand-out (0 = 8 x 8; at most 64 pixels)
};

struct DetachedArgs {
  const PtScene* scene = nullptr;          // sampling parameters θ0
  const float* value_materials = nullptr;  // device, 40 floats per material: BSDF values θ
  PtFrame frame;
  uint32_t spp = 1;
  uint32_t seed0 = 0;
  float* out = nullptr;                    // float4 per pixel (rows bottom-up), mean over spp
};

Status launch_detached(Context* ctx, const DetachedArgs& a, void* stream);
size_t adjoint_scratch_bytes(uint32_t width, uint32_t height, uint32_t material_count, uint32_t env_texels);
Status launch_adjoint(Context* ctx, const AdjointArgs& a, void* stream);

}  // namespace lucent
