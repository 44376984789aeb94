// Denoiser launch arguments (device pointers into a film's targets).
#pragma once

#include <cstdint>

#include "internal.hpp"

namespace lucent {

struct DenoiseArgs {
  const float* color = nullptr;      // film colour (RGBA32F)
  const uint16_t* albedo = nullptr;  // film albedo (RGBA16F)
  const uint16_t* normal = nullptr;  // film normal + depth (RGBA16F)
  const float* history = nullptr;    // film history (R32F)
  float* ping = nullptr;             // filter target A (RGBA32F)
  float* pong = nullptr;             // filter target B (RGBA32F)
  float* remodulated = nullptr;      // RGBA32F output of the last iteration: illumination * albedo
  float* guides = nullptr;           // 4 floats per pixel: unit normal (zero if uncovered), depth / |ray vector|
  float* var_a = nullptr;            // compact variance planes (1 float per pixel)
  float* var_b = nullptr;
  float** result = nullptr;          // receives the target holding the final iteration
  int width = 0, height = 0;
  uint32_t iterations = 5;
  float sigma_l = 4.0f, sigma_n = 64.0f, sigma_p = 1.5f;
  float cam_pos[3], cam_right[3], cam_up[3], cam_fwd[3];
  float tan_half_fov = 1.0f;
  int compute_units = 256;           // sizes the persistent grid of the a-trous passes
};

Status launch_denoise(const DenoiseArgs& a, void* stream);

}  // namespace lucent
