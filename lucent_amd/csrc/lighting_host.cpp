// Host-side producers of the lighting tables the integrator reads:
//  * the Walker alias table built with Vose's worklists
//    (reference src/math/alias.ts:17-64: LIFO small/large lists, float64 pmf);
//  * the environment importance sampler (src/scene/environment.ts:52-85);
//  * the procedural sky bake (src/scene/environment.ts:106-132);
//  * the sheen directional-albedo table (src/scene/sheen.ts:24-61), rounded
//    to half precision as its R16F upload is (src/renderer.ts:241-242);
//  * the camera basis (src/renderer.ts:156-161).
// Numerics are float64 like the JavaScript, without contraction.
#pragma clang fp contract(off)

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "render_host.hpp"

namespace lucent {

Status build_alias_table(const double* w, uint32_t n, float* prob, uint32_t* alias, double* pmf, double* total_out) {
  if (n == 0) return fail(LUCENT_E_INVALID, "alias table needs at least one weight");
  double total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (!(w[i] >= 0) || !std::isfinite(w[i]))
      return fail(LUCENT_E_INVALID, "weight " + std::to_string(i) + " is negative or not finite");
    total += w[i];
  }
  for (uint32_t i = 0; i < n; ++i) pmf[i] = total > 0 ? w[i] / total : 1.0 / n;
  std::vector<double> scaled(n);
  std::vector<uint32_t> small, large;
  small.reserve(n);
  large.reserve(n);
  for (uint32_t i = 0; i < n; ++i) {
    scaled[i] = pmf[i] * n;
    (scaled[i] < 1 ? small : large).push_back(i);
  }
  while (!small.empty() && !large.empty()) {
    const uint32_t s = small.back();
    small.pop_back();
    const uint32_t l = large.back();
    large.pop_back();
    prob[s] = float(scaled[s]);
    alias[s] = l;
    scaled[l] = scaled[l] + scaled[s] - 1;
    (scaled[l] < 1 ? small : large).push_back(l);
  }
  // What remains is 1 up to rounding.
  while (!large.empty()) {
    const uint32_t l = large.back();
    large.pop_back();
    prob[l] = 1.0f;
    alias[l] = l;
  }
  while (!small.empty()) {
    const uint32_t s = small.back();
    small.pop_back();
    prob[s] = 1.0f;
    alias[s] = s;
  }
  if (total_out) *total_out = total;
  return LUCENT_OK;
}

static double lum(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }

void env_grid(uint32_t width, uint32_t height, uint32_t max_width, uint32_t* gw_out, uint32_t* gh_out) {
  const uint32_t gw = std::max(1u, std::min(width, max_width));
  *gw_out = gw;
  *gh_out = std::max(1u, std::min(height, uint32_t(std::round(double(height) * gw / width))));
}

double env_row_sine(uint32_t cy, uint32_t gh) { return std::sin((M_PI * (cy + 0.5)) / gh); }

Status build_env_sampler(const float* rgb, uint32_t width, uint32_t height, uint32_t max_width, EnvSampler* out) {
  if (!rgb || width == 0 || height == 0) return fail(LUCENT_E_INVALID, "empty environment map");
  uint32_t gw, gh;
  env_grid(width, height, max_width, &gw, &gh);
  const size_t n = size_t(gw) * gh;
  std::vector<double> weights(n);
  for (uint32_t cy = 0; cy < gh; ++cy) {
    const uint32_t y0 = uint32_t(std::floor(double(cy) * height / gh));
    const uint32_t y1 = std::max(y0 + 1, uint32_t(std::floor(double(cy + 1) * height / gh)));
    const double sin_theta = env_row_sine(cy, gh);
    for (uint32_t cx = 0; cx < gw; ++cx) {
      const uint32_t x0 = uint32_t(std::floor(double(cx) * width / gw));
      const uint32_t x1 = std::max(x0 + 1, uint32_t(std::floor(double(cx + 1) * width / gw)));
      double sum = 0;
      for (uint32_t y = y0; y < y1; ++y)
        for (uint32_t x = x0; x < x1; ++x) {
          const float* p = rgb + (size_t(y) * width + x) * 3;
          sum += lum(p[0], p[1], p[2]);
        }
      weights[size_t(cy) * gw + cx] = (sum / (double(y1 - y0) * (x1 - x0))) * sin_theta;
    }
  }
  std::vector<float> prob(n);
  std::vector<uint32_t> alias(n);
  std::vector<double> pmf(n);
  double total = 0;
  Status s = build_alias_table(weights.data(), uint32_t(n), prob.data(), alias.data(), pmf.data(), &total);
  if (s != LUCENT_OK) return s;
  out->grid_w = gw;
  out->grid_h = gh;
  out->total = total;
  out->table.resize(n * 4);
  for (size_t i = 0; i < n; ++i) {
    const uint32_t a = alias[i];
    out->table[i * 4] = prob[i];
    out->table[i * 4 + 1] = float(a);
    out->table[i * 4 + 2] = float(pmf[i] * n);
    out->table[i * 4 + 3] = float(pmf[a] * n);
  }
  return LUCENT_OK;
}

static void from_spherical(double elevation, double azimuth, double* o) {
  const double c = std::cos(elevation);
  o[0] = c * std::sin(azimuth);
  o[1] = std::sin(elevation);
  o[2] = c * std::cos(azimuth);
}

static void direction_from_uv(double u, double v, double* d) {
  const double phi = (u - 0.5) * 2 * M_PI;
  const double theta = v * M_PI;
  const double s = std::sin(theta);
  d[0] = s * std::cos(phi);
  d[1] = std::cos(theta);
  d[2] = s * std::sin(phi);
}

void bake_sky(const lucent_sky& sky, const lucent_sun& sun, uint32_t width, std::vector<float>* out) {
  const uint32_t height = width / 2;
  out->assign(size_t(width) * height * 3, 0.0f);
  if (!sky.enabled || sky.intensity <= 0) return;
  double sd[3];
  from_spherical((sun.elevation * M_PI) / 180, (sun.azimuth * M_PI) / 180, sd);
  const bool sun_on = sun.intensity > 0 && sun.elevation > -sun.angular_radius;
  auto mix = [](double a, double b, double t) { return a + (b - a) * t; };
  for (uint32_t y = 0; y < height; ++y)
    for (uint32_t x = 0; x < width; ++x) {
      double d[3];
      direction_from_uv((x + 0.5) / width, (y + 0.5) / height, d);
      const double t = d[1];
      float* o = out->data() + (size_t(y) * width + x) * 3;
      for (int c = 0; c < 3; ++c)
        o[c] = float(t >= 0 ? mix(sky.horizon[c], sky.zenith[c], std::pow(t, 0.55))
                            : mix(sky.horizon[c], sky.ground[c], std::min(1.0, -t * 4)));
      if (sun_on) {
        const double s = std::max(0.0, d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2]);
        const double glow = 0.12 * std::pow(s, 24) + 0.5 * std::pow(s, 400);
        for (int c = 0; c < 3; ++c) o[c] = float(double(o[c]) + sun.color[c] * glow);
      }
      for (int c = 0; c < 3; ++c) o[c] = float(double(o[c]) * sky.intensity);
    }
}

static double charlie_d(double alpha, double cos_h) {
  const double inv = 1 / alpha;
  const double sin2 = std::max(0.0, 1 - cos_h * cos_h);
  return ((2 + inv) * std::pow(sin2, inv * 0.5)) / (2 * M_PI);
}
static double neubelt_v(double cos_l, double cos_v) { return 1 / (4 * (cos_l + cos_v - cos_l * cos_v)); }

double sheen_albedo(double cos_v, double roughness, uint32_t steps) {
  const double r = std::max(0.07, roughness);
  const double alpha = r * r;
  const double v0 = std::sqrt(std::max(0.0, 1 - cos_v * cos_v)), v2 = cos_v;
  const uint32_t phi_steps = steps / 2;
  const double d_theta = M_PI / 2 / steps;
  const double d_phi = M_PI / phi_steps;
  double e = 0;
  for (uint32_t i = 0; i < steps; ++i) {
    const double th = (i + 0.5) * d_theta;
    const double ch = std::cos(th), sh = std::sin(th);
    const double d = charlie_d(alpha, ch);
    for (uint32_t j = 0; j < phi_steps; ++j) {
      const double ph = (j + 0.5) * d_phi;
      const double h0 = sh * std::cos(ph), h2 = ch;
      const double vh = v0 * h0 + v2 * h2;
      if (vh <= 0) continue;
      const double cos_l = 2 * vh * h2 - v2;
      if (cos_l <= 0) continue;
      e += d * neubelt_v(cos_l, cos_v) * cos_l * 4 * vh * sh * d_theta * d_phi;
    }
  }
  return std::min(1.0, 2 * e);
}

// IEEE binary16 round-to-nearest-even of a float.
float round_to_half(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = x & 0x80000000u;
  uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return f;  // inf / nan
  float a;
  std::memcpy(&a, &ax, 4);
  if (a >= 65520.0f) {
    const uint32_t inf = sign | 0x7F800000u;
    float r;
    std::memcpy(&r, &inf, 4);
    return r;
  }
  // Quantum of the half grid at this magnitude (subnormal step 2^-24).
  int e;
  std::frexp(a, &e);  // a = m * 2^e, m in [0.5, 1)
  const int exp_half = std::max(e - 1, -14);
  const double q = std::ldexp(1.0, exp_half - 10);
  const double r = std::nearbyint(double(a) / q) * q;  // ties to even under the default rounding mode
  float out = float(r);
  if (sign) out = -out;
  return out;
}

void sheen_lut(uint32_t size, uint32_t steps, float* out) {
  for (uint32_t y = 0; y < size; ++y)
    for (uint32_t x = 0; x < size; ++x) {
      const double cos_v = std::max(1e-3, double(x) / (size - 1));
      out[y * size + x] = float(sheen_albedo(cos_v, double(y) / (size - 1), steps));
    }
}

void camera_basis(const lucent_camera& cam, double fwd[3], double right[3], double up[3], double* tan_half_fov) {
  double f[3] = {cam.target[0] - cam.position[0], cam.target[1] - cam.position[1], cam.target[2] - cam.position[2]};
  auto normalize = [](double* v) {
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len == 0) {
      v[0] = v[1] = v[2] = 0;
      return;
    }
    const double s = 1 / len;
    v[0] *= s;
    v[1] *= s;
    v[2] *= s;
  };
  auto cross = [](const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
  };
  normalize(f);
  const double world_up[3] = {0, std::fabs(f[1]) > 0.999 ? 0.0 : 1.0, std::fabs(f[1]) > 0.999 ? 1.0 : 0.0};
  double r[3];
  cross(f, world_up, r);
  normalize(r);
  double u[3];
  cross(r, f, u);
  for (int i = 0; i < 3; ++i) {
    fwd[i] = f[i];
    right[i] = r[i];
    up[i] = u[i];
  }
  *tan_half_fov = std::tan((cam.fov * M_PI) / 360);
}

}  // namespace lucent
