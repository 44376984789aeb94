// Presentation pass launch arguments (display.hip).
#pragma once

#include <cstdint>

#include "internal.hpp"

namespace lucent {

struct DisplayArgs {
  const float* color = nullptr;     // film colour (RGBA32F), render_w x render_h
  const uint16_t* albedo = nullptr; // film albedo (RGBA16F)
  const float* denoised = nullptr;  // denoised illumination (RGBA32F) or null when denoise_weight is 0
  uint8_t* out = nullptr;           // RGBA8, view_w x view_h, rows bottom-up
  int render_w = 0, render_h = 0, view_w = 0, view_h = 0;
  float denoise_weight = 0.0f;
  float exposure = 1.0f;
  int tonemap = 0;                  // 0 ACES, 1 Reinhard, 2 none
};

Status launch_display(const DisplayArgs& a, void* stream);

}  // namespace lucent
