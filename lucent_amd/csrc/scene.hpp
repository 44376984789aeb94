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
  std::vector<void*> env_allocations;  // replaced by each lucent_scene_set_lighting
  std::vector<void*> tex_allocations;  // replaced by each lucent_scene_set_textures
};

}  // namespace lucent

struct lucent_ctx : lucent::Context {};
struct lucent_scene : lucent::Scene {};
