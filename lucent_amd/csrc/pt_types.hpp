// Plain-data structures shared by the host C++ and the path-tracing kernels.
#pragma once

#include <cstdint>

#include "internal.hpp"

namespace lucent {

constexpr int kMaxLights = 8;
constexpr int kMaxSpheres = 64;
constexpr int kMaterialTexels = 10;
constexpr int kSheenLutSize = 16;

// Everything the integrator reads about a scene (device pointers + uniforms).
struct PtScene {
  DevBvh bvh;
  const float* materials = nullptr;  // float4[10 * material_count], material.ts:117-131
  uint32_t material_count = 0;
  const float* spheres = nullptr;    // float4 per sphere (centre, radius)
  const int32_t* sphere_material = nullptr;
  const int32_t* sphere_light = nullptr;
  uint32_t sphere_count = 0;
  int32_t ground_material = -1;
  float ground_height = 0.0f;
  uint32_t light_count = 0;
  float lights[kMaxLights * 16] = {};  // scene.ts:88-95 layout

  // Environment (lights.glsl:16-23).
  const float* env_map = nullptr;    // float4 per texel (rgb, 0), rows top to bottom
  int32_t env_w = 0, env_h = 0;
  const float* env_table = nullptr;  // float4 per cell (keep prob, alias cell, pdfUV, alias pdfUV)
  int32_t grid_w = 0, grid_h = 0;
  float env_cos = 1.0f, env_sin = 0.0f;
  float env_intensity = 1.0f;
  int32_t env_sampling = 0;
  int32_t env_background = 1;
  float backdrop[3] = {0, 0, 0};

  // Analytic sun (lights.glsl:10-14).
  int32_t sun_enabled = 0;
  float sun_dir[3] = {0, 1, 0};
  float sun_radiance[3] = {0, 0, 0};
  float sun_cos = 1.0f;
  float sun_pdf = 0.0f;

  const float* sheen_lut = nullptr;  // 16x16, half-rounded, row = roughness

  // Texture arrays (RGBA8 texels, layer-major, square tex_size).
  const uint32_t* color_maps = nullptr;  // sRGB colour
  const uint32_t* data_maps = nullptr;   // linear data
  int32_t tex_size = 0;
  int32_t color_layers = 0, data_layers = 0;
};

// Per-launch camera and integrator settings (trace.frag:12-40).
struct PtFrame {
  uint32_t width = 0, height = 0;
  uint32_t max_bounces = 8;
  float clamp = 0.0f;
  float cam_pos[3] = {0, 0, 0};
  float cam_right[3] = {1, 0, 0};
  float cam_up[3] = {0, 1, 0};
  float cam_fwd[3] = {0, 0, -1};
  float tan_half_fov = 1.0f;
  float aperture = 0.0f;
  float focus_dist = 1.0f;
  uint32_t seed = 0;
};

// Accumulation film (renderer.ts:490 formats): RGBA32F colour + lum^2,
// RGBA16F albedo + coverage, RGBA16F normal + depth, R32F history.
// Rows bottom to top, as the GL render targets.
struct PtFilm {
  float* color = nullptr;       // float4 per pixel
  uint16_t* albedo = nullptr;   // half4 per pixel
  uint16_t* normal = nullptr;   // half4 per pixel
  float* history = nullptr;     // float per pixel
};

// The previous view of a reprojecting sample (renderer.ts:715-726): the
// running means it accumulated (read only), its resolution and camera.
struct PrevView {
  PtFilm film;
  uint32_t width = 0, height = 0;
  float cam_pos[3] = {0, 0, 0};
  float cam_right[3] = {1, 0, 0};
  float cam_up[3] = {0, 1, 0};
  float cam_fwd[3] = {0, 0, -1};
  float tan_half_fov = 1.0f;
  float history_cap = 8.0f;  // uHistoryCap
};

// One film's share of a persistent tile launch: `spp` samples of every pixel
// with seeds frame.seed .. frame.seed + spp - 1, in that order per pixel.
struct TileItem {
  PtScene scene;
  PtFrame frame;
  PtFilm film;
  uint32_t spp = 0;
  uint32_t tiles_x = 0;
  // Launch shape of the per-sample hand-out (k_trace_units and the batch's
  // record ring): a job is a tile_w x tile_h tile (at most 64 pixels) with
  // all its samples; a wave shades once shade_batch of its lanes wait (or
  // none traces) and runs up to trace_turns traversal steps between its
  // checks for shading and new work.  None of them changes any path, so
  // films are bit-identical for every choice (the autotuner's search space).
  uint32_t tile_w = 8, tile_h = 8;
  uint32_t shade_batch = 32, trace_turns = 8;
  // Host only: the sample count a sized job tile is chosen for (0: spp).
  // A timing trial that traces fewer samples than the real launch sets it
  // to the real count, so that both use the same tiles.
  uint32_t job_spp = 0;
  // Optional: each path's radiance as tracePath returns it, float4 per pixel
  // per sample, sample-major (sample k of the launch at k * width * height).
  float* sample_radiance = nullptr;
  // Optional: a device word holding the first seed, read instead of
  // frame.seed (a replayed graph advances it on the device).
  const uint32_t* seed_src = nullptr;
};

}  // namespace lucent
