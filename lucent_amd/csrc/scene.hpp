// The uploaded scene: device allocations plus the integrator's view of them.
#pragma once

#include <vector>

#include "internal.hpp"
#include "pt_types.hpp"

namespace lucent {

struct Scene {
  Context* ctx = nullptr;
  DevBvh bvh;
  std::vector<void*> allocations;
  lucent_scene_stats stats{};
  PtScene pt;                    // bvh, materials, lights, environment as the kernels read them
  bool lighting_set = false;
  bool blend_alpha = false;      // some material is alpha-blended (draws random numbers inside traversal)
  std::vector<void*> env_allocations;  // replaced by each lucent_scene_set_lighting
  std::vector<void*> tex_allocations;  // replaced by each lucent_scene_set_textures
};

}  // namespace lucent

namespace lucent {
// alphaMode "blend" (code 2 in material texel 3 w, material.ts:117-131).
inline bool any_blend_alpha(const float* materials, uint32_t n) {
  for (uint32_t i = 0; i < n; ++i)
    if (int(materials[size_t(i) * 40 + 15]) == 2) return true;
  return false;
}
}  // namespace lucent

struct lucent_ctx : lucent::Context {};
struct lucent_scene : lucent::Scene {};
