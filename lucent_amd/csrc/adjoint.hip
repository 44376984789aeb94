// Path-replay adjoint of the progressive path tracer on gfx950.
//
// For each pixel sample the kernel takes the path radiance L recorded by the
// forward pass (lucent_forward), or re-runs the primal path from its RNG seed
// (trace.frag semantics, pt_device.hpp) to obtain it, then replays the same path (same random numbers, so the same vertices
// and decisions) and differentiates each contribution:
//   * NEE terms  beta * f(θ) * Li * k     -> d/dθ through f and through the
//                                            environment texels behind Li;
//   * escapes    beta * Lenv * w          -> d/d(texels) of the bilinear lookup;
//   * the BSDF continuation weight f(θ) cos / pdf scales every later term,
//     whose sum is the remaining radiance Lr = L - (terms so far), so it
//     contributes Lr * (df/dθ) / f.
// Sampling probabilities, MIS weights, Russian-roulette probabilities and the
// environment sampler's pdf are held constant (detached): the estimator is
// unbiased for any choice of them, so its expectation's gradient is the
// expectation of these terms.  df/dθ comes from forward-mode dual numbers
// over the opaque BSDF (bsdf.glsl:114-166) with θ = (albedo rgb, roughness,
// metallic) of the hit material.
//
// Reductions, FP64 throughout: material gradients are summed per block in
// LDS and written as per-block partials, then reduced by a second kernel;
// environment-texel gradients go through atomics into 8 replicas to spread
// contention and are summed by a final kernel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "adjoint.hpp"
#include "pt_device.hpp"

namespace lucent {
namespace {

using namespace pt;

constexpr int kAdjBlock = 64;
constexpr int kTile = 8;
constexpr int kNumParams = 5;        // albedo r, g, b, roughness, metallic
constexpr int kLdsMaterials = 64;    // materials whose gradients are pre-reduced in LDS
constexpr int kEnvReplicas = 8;
#ifndef LUCENT_ADJ_STACK
#define LUCENT_ADJ_STACK 8
#endif
constexpr int kAdjStack = LUCENT_ADJ_STACK;  // LDS entries of the replay's traversal stack ring

// Forward-mode dual number with the five material-parameter tangents.
struct D5 {
  float v;
  float d[kNumParams];
};
__device__ inline D5 dc(float v) {
  D5 r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = 0.0f;
  return r;
}
__device__ inline D5 dvar(float v, int i) {
  D5 r = dc(v);
  r.d[i] = 1.0f;
  return r;
}
__device__ inline D5 operator+(D5 a, D5 b) {
  D5 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
__device__ inline D5 operator-(D5 a, D5 b) {
  D5 r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
__device__ inline D5 operator*(D5 a, D5 b) {
  D5 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
__device__ inline D5 operator*(D5 a, float s) {
  D5 r;
  r.v = a.v * s;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = a.d[i] * s;
  return r;
}
__device__ inline D5 operator+(D5 a, float s) {
  a.v += s;
  return a;
}
__device__ inline D5 operator/(D5 a, D5 b) {
  D5 r;
  const float inv = 1.0f / b.v;
  r.v = a.v * inv;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
__device__ inline D5 dsqrt(D5 a) {
  D5 r;
  r.v = sqrtf(a.v);
  const float k = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
  for (int i = 0; i < kNumParams; ++i) r.d[i] = a.d[i] * k;
  return r;
}
__device__ inline D5 one_minus(D5 a) { return dc(1.0f) - a; }

// f of evalOpaque (bsdf.glsl:114-166) and its derivatives with respect to
// the hit material's albedo, roughness and metallic.  The sheen energy scale
// and the coat layering do not depend on these parameters.
__device__ __forceinline__ bool eval_opaque_grad(const PtScene& sc, const Surface& s, v3 wo, v3 wi, float f[3],
                                        float df[3][kNumParams]) {
  const v3 n = s.n;
  const float ndl = dot(n, wi), ndv = dot(n, wo);
  if (ndl <= 0.0f || ndv <= 0.0f) return false;
  const D5 alb[3] = {dvar(s.albedo.x, 0), dvar(s.albedo.y, 1), dvar(s.albedo.z, 2)};
  const D5 rough = dvar(s.roughness, 3);
  const D5 metal = dvar(s.metallic, 4);
  const D5 a = (rough.v * rough.v > 1e-3f) ? rough * rough : dc(1e-3f);
  const v3 H = normalize(wo + wi);
  const float ndh = fmaxf(dot(n, H), 0.0f);
  const float vdh = fmaxf(dot(wo, H), 0.0f);
  const float fres = pow5(1.0f - clampf(vdh, 0.0f, 1.0f));
  const float d0 = dielectric_f0(s.ior);
  const D5 a2 = a * a;
  const D5 dd = (a2 + (-1.0f)) * (ndh * ndh) + 1.0f;
  const D5 D = a2 / (dd * dd * PI);
  auto g1 = [&](float x) { return dc(2.0f * x) / (dsqrt(a2 + one_minus(a2) * (x * x)) + x); };
  const D5 G = g1(ndv) * g1(ndl);
  const float denom = 1.0f / fmaxf(4.0f * ndv * ndl, 1e-6f);
  const D5 DG = D * G * denom;
  const D5 one_m = one_minus(metal);
  float scale = 1.0f, sheen[3] = {0.0f, 0.0f, 0.0f};
  const float sheen_max = max3(s.sheen_color);
  if (sheen_max > 0.0f) {
    const float sr = fmaxf(s.sheen_roughness, MIN_SHEEN_ROUGHNESS);
    const float k = d_charlie(sr * sr, ndh) * v_neubelt(ndl, ndv);
    sheen[0] = s.sheen_color.x * k;
    sheen[1] = s.sheen_color.y * k;
    sheen[2] = s.sheen_color.z * k;
    scale = 1.0f - sheen_max * fmaxf(sheen_albedo(sc, ndv, sr), sheen_albedo(sc, ndl, sr));
  }
  float coat_mul = 1.0f, coat_add = 0.0f;
  if (coat_probability(s, wo) > 0.0f) {
    const float cv = dot(s.nc, wo), cl = dot(s.nc, wi);
    float coat = 0.0f;
    if (cv > 0.0f && cl > 0.0f) {
      const float ac = alpha_from_roughness(s.clearcoat_roughness);
      const float ch = fmaxf(dot(s.nc, H), 0.0f);
      coat = d_ggx(ch, ac) * g1_ggx(cv, ac) * g1_ggx(cl, ac) * fresnel_schlick1(vdh, 0.04f) / (4.0f * cv * cl);
    }
    coat_mul = 1.0f - s.clearcoat * fresnel_schlick1(fmaxf(cv, 0.0f), 0.04f);
    coat_add = s.clearcoat * coat;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const D5 f0 = dc(d0) * one_m + alb[c] * metal;
    const D5 F = f0 + one_minus(f0) * fres;
    const D5 spec = DG * F;
    const D5 diff = one_minus(F) * alb[c] * one_m * INV_PI;
    D5 fc = (diff + spec) * scale + sheen[c];
    fc = fc * coat_mul + coat_add;
    f[c] = fc.v;
  #pragma unroll
  for (int i = 0; i < kNumParams; ++i) df[c][i] = fc.d[i];
  }
  return true;
}

// Scales of the surface parameters with respect to the material's own
// (surfaceAt, trace.frag:72-91: checker colour, base/metal-rough maps).
struct ParamScale {
  float albedo[3];
  float rough, metal;
};

__device__ __forceinline__ ParamScale param_scale(const PtScene& sc, const Hit& hit) {
  const Material m = fetch_material(sc, hit.material);
  ParamScale p{{1.0f, 1.0f, 1.0f}, 1.0f, 1.0f};
  float tu, tv;
  transform_uv(hit.uv_u, hit.uv_v, m.uv0, m.uv1, tu, tv);
  if (m.checker_scale > 0.0f) {
    const float sum = floorf(hit.position.x * m.checker_scale + 0.5f) + floorf(hit.position.y * m.checker_scale + 0.5f) +
                      floorf(hit.position.z * m.checker_scale + 0.5f);
    if (sum - 2.0f * floorf(sum / 2.0f) >= 0.5f) p.albedo[0] = p.albedo[1] = p.albedo[2] = 0.0f;
  }
  const bool tex = sc.tex_size > 0;
  if (m.map_base >= 0 && m.map_base < sc.color_layers && tex) {
    const float4 t = sample_layer(sc.color_maps, sc.tex_size, m.map_base, tu, tv, true);
    p.albedo[0] *= t.x;
    p.albedo[1] *= t.y;
    p.albedo[2] *= t.z;
  }
  if (m.map_mr >= 0 && m.map_mr < sc.data_layers && tex) {
    const float4 mr = sample_layer(sc.data_maps, sc.tex_size, m.map_mr, tu, tv, false);
    p.rough = mr.y;
    p.metal = mr.z;
  }
  return p;
}

// Gradient sums are FP64 throughout (LDS and global double atomics): a
// persistent launch sums every sample of the frame into them.
//
// Environment texels go through a per-block cache in LDS first.  The
// environment sampler draws most light samples from a few bright texels,
// and every addition to global memory is a memory-side atomic (the MI355X
// executes device atomics beyond the XCD's L2), so many lanes queued on the
// same few addresses.  The block sums additions to up to kEnvSlots texels
// in LDS (direct-mapped by a hash of the texel index; a slot is claimed by
// the first texel that reaches it) and adds each claimed slot to its global
// replica once, after the block's last path; a texel whose slot another
// texel holds goes to global memory directly.
#ifndef LUCENT_ADJ_ENV_SLOTS_LOG2
#define LUCENT_ADJ_ENV_SLOTS_LOG2 7
#endif
constexpr int kEnvSlots = 1 << LUCENT_ADJ_ENV_SLOTS_LOG2;
constexpr uint32_t kFreeSlot = 0xFFFFFFFFu;
__device__ inline uint32_t env_slot(uint32_t texel) { return (texel * 2654435761u) >> (32 - LUCENT_ADJ_ENV_SLOTS_LOG2); }

// FP64 atomic additions with the memory named: the sink's pointers are read
// from memory, so left generic they would compile to flat atomics (counted
// on the LDS counter as well).  Both are the hardware's FP64 add.
__device__ __forceinline__ void add_global_f64(double* p, double v) {
  [[clang::atomic(ignore_denormal_mode)]] {
    __hip_atomic_fetch_add((__attribute__((address_space(1))) double*)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
  }
}
__device__ __forceinline__ void add_lds_f64(double* p, double v) {
  [[clang::atomic(ignore_denormal_mode)]] {
    __hip_atomic_fetch_add((__attribute__((address_space(3))) double*)p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

struct GradSink {
  double* lds;  // kLdsMaterials * 5 (block-shared)
  double* env;  // kEnvReplicas replicas of env_w * env_h * 3
  size_t env_stride;
  uint32_t replica;
  int material_count;
  double* global_mat;  // materials beyond the LDS table
  uint32_t* env_tag;   // kEnvSlots texel indices (block-shared LDS)
  double* env_val;     // kEnvSlots * 3 sums (block-shared LDS)
};

__device__ __forceinline__ void add_material(GradSink& g, int mat, const float v[kNumParams]) {
  if (mat < kLdsMaterials) {
  #pragma unroll
  for (int i = 0; i < kNumParams; ++i)
      if (v[i] != 0.0f) add_lds_f64(&g.lds[mat * kNumParams + i], double(v[i]));
  } else {
  #pragma unroll
  for (int i = 0; i < kNumParams; ++i)
      if (v[i] != 0.0f) add_global_f64(&g.global_mat[mat * kNumParams + i], double(v[i]));
  }
}

__device__ __forceinline__ void add_env(const PtScene& sc, GradSink& g, float u, float v, v3 weight) {
  int idx[4];
  float w[4];
  env_taps(sc, u, v, idx, w);
  double* base = g.env + g.replica * g.env_stride;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (w[k] == 0.0f) continue;
    const float s = w[k] * sc.env_intensity;
    const double v[3] = {double(weight.x * s), double(weight.y * s), double(weight.z * s)};
    const uint32_t texel = uint32_t(idx[k]), slot = env_slot(texel);
    uint32_t owner = g.env_tag[slot];
    if (owner == kFreeSlot) {
      owner = atomicCAS(&g.env_tag[slot], kFreeSlot, texel);
      if (owner == kFreeSlot) owner = texel;
    }
    if (owner == texel) {
      for (int c = 0; c < 3; ++c) add_lds_f64(&g.env_val[slot * 3 + c], v[c]);
    } else {
      for (int c = 0; c < 3; ++c) add_global_f64(&base[size_t(texel) * 3 + c], v[c]);
    }
  }
}

// Adds the block's cached texel sums to its replica (after the block's
// last path, the block synchronised).
__device__ inline void flush_env(GradSink& g, int lane, int lanes) {
  double* base = g.env + g.replica * g.env_stride;
  for (int slot = lane; slot < kEnvSlots; slot += lanes) {
    const uint32_t texel = g.env_tag[slot];
    if (texel == kFreeSlot) continue;
    for (int c = 0; c < 3; ++c) add_global_f64(&base[size_t(texel) * 3 + c], g.env_val[slot * 3 + c]);
  }
}

// Gradient of δL · (beta * f * Li * k) through f.
__device__ __forceinline__ void add_f_term(GradSink& g, int mat, const ParamScale& ps, v3 coeff, const float df[3][kNumParams]) {
  const float c[3] = {coeff.x, coeff.y, coeff.z};
  float out[kNumParams] = {0, 0, 0, 0, 0};
  for (int ch = 0; ch < 3; ++ch)
  #pragma unroll
  for (int i = 0; i < kNumParams; ++i) out[i] += c[ch] * df[ch][i];
  out[0] *= ps.albedo[0];
  out[1] *= ps.albedo[1];
  out[2] *= ps.albedo[2];
  out[3] *= ps.rough;
  out[4] *= ps.metal;
  add_material(g, mat, out);
}

// The replay of tracePath (trace.frag:136-272) accumulating d(dL · L)/dθ.
// A path's replay state between bounces: the ray, the radiance still ahead
// of this vertex (lr), the throughput.
struct ReplayState {
  v3 ro, rd, lr, beta;
  float bsdf_pdf;
  int bounce;
  bool specular;
};

__device__ __forceinline__ void replay_begin(ReplayState& r, v3 ro, v3 rd, v3 L) {
  r.ro = ro;
  r.rd = rd;
  r.lr = L;
  r.beta = mk(1.f, 1.f, 1.f);
  r.bsdf_pdf = 0.0f;
  r.bounce = 0;
  r.specular = true;
}

// One bounce of the replay; false once the path has ended.
template <class StackT>
__device__ __forceinline__ bool replay_step(const PtScene& sc, const PtFrame& f, ReplayState& r, Rng& rng, StackT& st,
                                            v3 dl, GradSink& g) {
  v3 ro = r.ro, rd = r.rd, lr = r.lr, beta = r.beta;
  bool specular = r.specular;
  float bsdf_pdf = r.bsdf_pdf;
  const int bounce = r.bounce;
  {
    Hit hit;
    if (!trace_scene(sc, ro, rd, hit, rng, st)) {
      if (!(bounce == 0 && sc.env_background == 0)) {
        float w = 1.0f;
        if (!specular && sc.env_sampling != 0) w = mis_weight(bsdf_pdf, env_pdf(sc, rd));
        float u, v;
        equirect_uv(env_to_local(sc, rd), u, v);
        add_env(sc, g, u, v, dl * beta * w);
      }
      return false;
    }
    const v3 wo = -rd;
    v3 emission;
    const Surface s = surface_at(sc, hit, wo, emission);
    const v3 ng = hit.geom_normal;
    if (!hit.front_face && !s.thin_walled && s.transmission > 0.0f) beta = beta * vexp(-s.absorption * hit.t);
    if (emission.x > 0.0f || emission.y > 0.0f || emission.z > 0.0f) {
      float w = 1.0f;
      if (!specular && hit.light >= 0) w = mis_weight(bsdf_pdf, light_pdf(sc, hit.light, ro, rd, hit.t) / float(sc.light_count));
      lr = lr - beta * emission * w;
      if (hit.light >= 0) return false;
    }
    const bool glass = s.transmission > 0.0f && rng.rand() < s.transmission;
    const ParamScale ps = param_scale(sc, hit);
    if (glass) {
      v3 wi, weight;
      bool reflected;
      if (!sample_dielectric(s, hit.front_face, ng, wo, rng, wi, weight, reflected)) return false;
      if (!reflected) {
        // Refraction weight = albedo: later terms scale with it.
        float out[kNumParams] = {0, 0, 0, 0, 0};
        if (s.albedo.x > 0.0f) out[0] = dl.x * lr.x / s.albedo.x * ps.albedo[0];
        if (s.albedo.y > 0.0f) out[1] = dl.y * lr.y / s.albedo.y * ps.albedo[1];
        if (s.albedo.z > 0.0f) out[2] = dl.z * lr.z / s.albedo.z * ps.albedo[2];
        add_material(g, hit.material, out);
      }
      beta = beta * weight;
      ro = offset_ray(hit.position, reflected ? ng : -ng);
      rd = wi;
      specular = true;
    } else {
      const v3 shadow_origin = offset_ray(hit.position, ng);
      const v3 n = s.n;
      float fv[3], dfv[3][kNumParams];
      if (sc.sun_enabled != 0) {
        const float u0 = rng.rand(), u1 = rng.rand();
        const v3 wi = sample_sun_direction(sc, u0, u1);
        const float ndl = dot(n, wi);
        if (ndl > 0.0f && dot(ng, wi) > 0.0f && !occluded(sc, shadow_origin, wi, INF, rng, st)) {
          float pb;
          eval_opaque(sc, s, wo, wi, pb);
          const float k = ndl * mis_weight(sc.sun_pdf, pb) / sc.sun_pdf;
          if (eval_opaque_grad(sc, s, wo, wi, fv, dfv)) {
            const v3 li = mk(sc.sun_radiance) * k;
            lr = lr - beta * mk(fv[0], fv[1], fv[2]) * li;
            add_f_term(g, hit.material, ps, dl * beta * li, dfv);
          }
        }
      }
      if (sc.env_sampling != 0) {
        const LightSample ls = sample_environment(sc, rng);
        const float ndl = dot(n, ls.dir);
        if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f && !occluded(sc, shadow_origin, ls.dir, INF, rng, st)) {
          float pb;
          eval_opaque(sc, s, wo, ls.dir, pb);
          const float k = ndl * mis_weight(ls.pdf, pb) / ls.pdf;
          if (eval_opaque_grad(sc, s, wo, ls.dir, fv, dfv)) {
            const v3 fvec = mk(fv[0], fv[1], fv[2]);
            lr = lr - beta * fvec * ls.radiance * k;
            add_f_term(g, hit.material, ps, dl * beta * ls.radiance * k, dfv);
            add_env(sc, g, ls.u, ls.v, dl * beta * fvec * k);
          }
        }
      }
      if (sc.light_count > 0) {
        const int li = min(int(rng.rand() * float(sc.light_count)), int(sc.light_count) - 1);
        const float u0 = rng.rand(), u1 = rng.rand();
        const LightSample ls = sample_light(sc, li, hit.position, u0, u1);
        const float ndl = dot(n, ls.dir);
        if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f &&
            !occluded(sc, shadow_origin, ls.dir, ls.dist * (1.0f - 1e-4f), rng, st)) {
          const float pl = ls.pdf / float(sc.light_count);
          float pb;
          eval_opaque(sc, s, wo, ls.dir, pb);
          const float k = ndl * mis_weight(pl, pb) / pl;
          if (eval_opaque_grad(sc, s, wo, ls.dir, fv, dfv)) {
            lr = lr - beta * mk(fv[0], fv[1], fv[2]) * ls.radiance * k;
            add_f_term(g, hit.material, ps, dl * beta * ls.radiance * k, dfv);
          }
        }
      }
      const BsdfSample bs = sample_opaque(sc, s, wo, rng);
      if (bs.pdf <= 0.0f || dot(bs.dir, ng) <= 0.0f) return false;
      // Every later term carries this vertex's f: δL · Lr * (df/dθ) / f.
      if (eval_opaque_grad(sc, s, wo, bs.dir, fv, dfv)) {
        v3 coeff = mk(0.f, 0.f, 0.f);
        if (fv[0] > 1e-30f) coeff.x = dl.x * lr.x / fv[0];
        if (fv[1] > 1e-30f) coeff.y = dl.y * lr.y / fv[1];
        if (fv[2] > 1e-30f) coeff.z = dl.z * lr.z / fv[2];
        add_f_term(g, hit.material, ps, coeff, dfv);
      }
      beta = beta * bs.weight;
      bsdf_pdf = bs.pdf;
      specular = false;
      ro = shadow_origin;
      rd = bs.dir;
    }
    if (bounce >= 2) {
      const float p = fminf(0.95f, max3(beta));
      if (rng.rand() >= p) return false;
      beta = beta / p;
    }
  }
  r.ro = ro;
  r.rd = rd;
  r.lr = lr;
  r.beta = beta;
  r.specular = specular;
  r.bsdf_pdf = bsdf_pdf;
  r.bounce = bounce + 1;
  return r.bounce <= int(f.max_bounces);
}

// The detached estimator as a function of the material parameters: every
// random decision, direction, pdf, MIS weight and roulette probability comes
// from the scene's materials `sc` (θ0), the BSDF values, refraction tints,
// absorption and emission from `sv` (θ, same geometry and lighting).  At θ = θ0
// this is tracePath (trace.frag:136-272, clamp 0) up to rounding; its
// derivative in θ at θ0 is the sum replay_step accumulates.  Two throughputs run
// side by side: beta0 drives roulette, beta carries the θ values.
template <class StackT>
__device__ inline v3 detached_path(const PtScene& sc, const PtScene& sv, const PtFrame& f, v3 ro, v3 rd, Rng& rng,
                                   StackT& st) {
  v3 radiance = mk(0.f, 0.f, 0.f);
  v3 beta0 = mk(1.f, 1.f, 1.f), beta = mk(1.f, 1.f, 1.f);
  bool specular = true;
  float bsdf_pdf = 0.0f;
  for (int bounce = 0; bounce <= int(f.max_bounces); ++bounce) {
    Hit hit;
    if (!trace_scene(sc, ro, rd, hit, rng, st)) {
      const v3 env = (bounce == 0 && sc.env_background == 0) ? mk(sc.backdrop) : environment(sc, rd, specular, bsdf_pdf);
      radiance = radiance + beta * env;
      break;
    }
    const v3 wo = -rd;
    v3 emission, emission_v;
    const Surface s = surface_at(sc, hit, wo, emission);
    const Surface s_v = surface_at(sv, hit, wo, emission_v);
    const v3 ng = hit.geom_normal;
    if (!hit.front_face && !s.thin_walled && s.transmission > 0.0f) {
      beta0 = beta0 * vexp(-s.absorption * hit.t);
      beta = beta * vexp(-s_v.absorption * hit.t);
    }
    if (emission.x > 0.0f || emission.y > 0.0f || emission.z > 0.0f) {
      float w = 1.0f;
      if (!specular && hit.light >= 0) w = mis_weight(bsdf_pdf, light_pdf(sc, hit.light, ro, rd, hit.t) / float(sc.light_count));
      radiance = radiance + beta * emission_v * w;
      if (hit.light >= 0) break;
    }
    const bool glass = s.transmission > 0.0f && rng.rand() < s.transmission;
    if (glass) {
      v3 wi, weight;
      bool reflected;
      if (!sample_dielectric(s, hit.front_face, ng, wo, rng, wi, weight, reflected)) break;
      beta0 = beta0 * weight;
      beta = beta * (reflected ? weight : s_v.albedo);
      ro = offset_ray(hit.position, reflected ? ng : -ng);
      rd = wi;
      specular = true;
    } else {
      const v3 shadow_origin = offset_ray(hit.position, ng);
      const v3 n = s.n;
      if (sc.sun_enabled != 0) {
        const float u0 = rng.rand(), u1 = rng.rand();
        const v3 wi = sample_sun_direction(sc, u0, u1);
        const float ndl = dot(n, wi);
        if (ndl > 0.0f && dot(ng, wi) > 0.0f && !occluded(sc, shadow_origin, wi, INF, rng, st)) {
          float pb, unused;
          eval_opaque(sc, s, wo, wi, pb);
          const v3 fv = eval_opaque(sv, s_v, wo, wi, unused);
          radiance = radiance + beta * fv * ndl * mk(sc.sun_radiance) * (mis_weight(sc.sun_pdf, pb) / sc.sun_pdf);
        }
      }
      if (sc.env_sampling != 0) {
        const LightSample ls = sample_environment(sc, rng);
        const float ndl = dot(n, ls.dir);
        if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f && !occluded(sc, shadow_origin, ls.dir, INF, rng, st)) {
          float pb, unused;
          eval_opaque(sc, s, wo, ls.dir, pb);
          const v3 fv = eval_opaque(sv, s_v, wo, ls.dir, unused);
          radiance = radiance + beta * fv * ndl * ls.radiance * (mis_weight(ls.pdf, pb) / ls.pdf);
        }
      }
      if (sc.light_count > 0) {
        const int li = min(int(rng.rand() * float(sc.light_count)), int(sc.light_count) - 1);
        const float u0 = rng.rand(), u1 = rng.rand();
        const LightSample ls = sample_light(sc, li, hit.position, u0, u1);
        const float ndl = dot(n, ls.dir);
        if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f &&
            !occluded(sc, shadow_origin, ls.dir, ls.dist * (1.0f - 1e-4f), rng, st)) {
          const float pl = ls.pdf / float(sc.light_count);
          float pb, unused;
          eval_opaque(sc, s, wo, ls.dir, pb);
          const v3 fv = eval_opaque(sv, s_v, wo, ls.dir, unused);
          radiance = radiance + beta * fv * ndl * ls.radiance * (mis_weight(pl, pb) / pl);
        }
      }
      const BsdfSample bs = sample_opaque(sc, s, wo, rng);
      if (bs.pdf <= 0.0f || dot(bs.dir, ng) <= 0.0f) break;
      float unused;
      const v3 fv = eval_opaque(sv, s_v, wo, bs.dir, unused);
      beta0 = beta0 * bs.weight;
      beta = beta * fv * (dot(n, bs.dir) / bs.pdf);
      bsdf_pdf = bs.pdf;
      specular = false;
      ro = shadow_origin;
      rd = bs.dir;
    }
    if (bounce >= 2) {
      const float p = fminf(0.95f, max3(beta0));
      if (rng.rand() >= p) break;
      beta0 = beta0 / p;
      beta = beta / p;
    }
  }
  return radiance;
}

// One sample per pixel of detached_path: out += L / spp (the first launch
// overwrites), so after spp launches out holds the mean.
__global__ __launch_bounds__(kAdjBlock) void k_detached(PtScene sc, PtScene sv, PtFrame f, float inv_spp, int first,
                                                        float* __restrict__ out, uint32_t* spill, uint32_t spill_depth,
                                                        uint32_t tiles_x) {
  __shared__ uint32_t lds[kPtStackWords * kAdjBlock];
  const uint32_t tile = blockIdx.x;
  const uint32_t x = (tile % tiles_x) * kTile + (threadIdx.x & (kTile - 1));
  const uint32_t y = (tile / tiles_x) * kTile + (threadIdx.x / kTile);
  if (x >= f.width || y >= f.height) return;
  const uint32_t idx = y * f.width + x;
  Rng rng = seed_rng(idx, f.seed);
  v3 ro, rd;
  camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  PtStack<kAdjBlock> st{as_lds(lds + threadIdx.x), spill + size_t(blockIdx.x * kAdjBlock + threadIdx.x) * spill_depth, 0,
                               0};
  v3 L = detached_path(sc, sv, f, ro, rd, rng, st);
  if (isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z)) L = mk(0.f, 0.f, 0.f);
  float4* o = reinterpret_cast<float4*>(out) + idx;
  const float4 prev = first ? make_float4(0.f, 0.f, 0.f, 0.f) : *o;
  *o = make_float4(prev.x + L.x * inv_spp, prev.y + L.y * inv_spp, prev.z + L.z * inv_spp, 0.0f);
}

// Persistent replay over all samples: waves take 8x8 tile jobs from an
// atomic queue and hand out the job's pixel samples to their lanes, one path
// per lane at a time, one bounce per loop turn (the job's pixels at sample
// 0, then at sample 1, ...; sample k: seed seed0 + k, radiance
// recorded_L[k]).  A lane whose path ends takes the next sample of the job
// or of the next job, so lanes neither idle behind the longest path of the
// wave nor behind the longest pixel at the end of the queue (at 512x512 a
// lane sees only about one pixel).  Material gradients gather in the block's
// LDS table over everything the block replays and are written once as the
// block's partials.
// kWaves: the occupancy the register allocation targets (waves per SIMD).
// The scene and frame of one replay launch, in device memory: k_adjoint
// reads them through a pointer, so the compiler loads their fields with
// scalar loads where they are used instead of keeping more of them live
// (by value, 52 B per lane spilled at 3 waves per SIMD).
struct AdjointView {
  PtScene sc;
  PtFrame f;
};

__global__ void k_store_view(AdjointView* dst, PtScene sc, PtFrame f) {
  dst->sc = sc;
  dst->f = f;
}

template <int kWaves>
__global__ __launch_bounds__(kAdjBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_adjoint(
    const AdjointView* __restrict__ view, uint32_t spp, uint32_t n_jobs, uint32_t* __restrict__ counter, const float* __restrict__ dl_di,
    const float4* __restrict__ recorded_L, float inv_spp, double* __restrict__ mat_partials, double* __restrict__ env_acc,
    size_t env_stride, double* __restrict__ mat_overflow, uint32_t* spill, uint32_t spill_depth, uint32_t tiles_x,
    uint32_t tile_w, uint32_t tile_h) {
  const PtScene& sc = view->sc;
  const PtFrame& f = view->f;
  // LDS per block: the stack ring (kAdjStack entries, 2 words each), the
  // material table and the texel cache: 10,240 B, so that 16 blocks of one
  // wave fit a CU's 160 KiB at 4 waves per SIMD.  (Measured: 128 slots and
  // an 8-entry ring beat 64 slots and 10 entries, and 32 and 10.)
  __shared__ uint32_t lds[2 * kAdjStack * kAdjBlock];
  __shared__ double lds_grad[kLdsMaterials * kNumParams];
  __shared__ uint32_t env_tag[kEnvSlots];
  __shared__ double env_val[kEnvSlots * 3];
  for (int i = threadIdx.x; i < kLdsMaterials * kNumParams; i += kAdjBlock) lds_grad[i] = 0.0;
  for (int i = threadIdx.x; i < kEnvSlots; i += kAdjBlock) {
    env_tag[i] = kFreeSlot;
    env_val[i * 3] = env_val[i * 3 + 1] = env_val[i * 3 + 2] = 0.0;
  }
  __syncthreads();
  Stack4<kAdjStack, kAdjBlock> st{as_lds(lds + threadIdx.x), spill + size_t(blockIdx.x * kAdjBlock + threadIdx.x) * spill_depth,
                                  0, 0};
  GradSink g{lds_grad, env_acc, env_stride, blockIdx.x % kEnvReplicas, int(sc.material_count), mat_overflow,
             env_tag, env_val};
  const uint32_t lane = threadIdx.x;
  const size_t frame_px = size_t(f.width) * f.height;
  auto take = [&]() {
    uint32_t v = 0;
    if (lane == 0) v = atomicAdd(counter, 1u);
    return __builtin_amdgcn_readlane(v, 0);
  };
  uint32_t job = take();
  // The job being handed out (wave-uniform): tile corner, width, pixels,
  // and the next of its px * spp samples.
  uint32_t x0 = 0, y0 = 0, tw = 1, px = 0, units = 0, handed = 0;
  auto open_job = [&](uint32_t j) {
    x0 = (j % tiles_x) * tile_w;
    y0 = (j / tiles_x) * tile_h;
    tw = min(tile_w, f.width - x0);
    px = tw * min(tile_h, f.height - y0);
    units = px * spp;
    handed = 0;
  };
  bool taking = job < n_jobs;
  if (taking) open_job(job);
  bool active = false;
  v3 dl = mk(0.f, 0.f, 0.f);
  Rng rng{};
  ReplayState r;
  for (;;) {
    uint64_t idle = __ballot(!active);
    while (taking && idle != 0) {
      if (handed == units) {
        const uint32_t j = take();
        if (j >= n_jobs) {
          taking = false;
          break;
        }
        open_job(j);
        continue;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
      const uint32_t avail = units - handed;
      if (((idle >> lane) & 1u) && rank < avail) {
        const uint32_t u = handed + rank;
        const uint32_t slot = u % px, k = u / px;
        const uint32_t x = x0 + slot % tw, y = y0 + slot / tw;
        const uint32_t idx = y * f.width + x;
        const float4 d = reinterpret_cast<const float4*>(dl_di)[idx];
        dl = mk(d.x * inv_spp, d.y * inv_spp, d.z * inv_spp);
        // A pixel without loss gradient, or a sample whose radiance is not
        // finite (the film's accumulation drops it), needs no replay: the
        // lane asks again.
        if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) {
          rng = seed_rng(idx, f.seed + k);
          v3 ro, rd;
          camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
          const Rng replay_rng = rng;
          v3 L;
          if (recorded_L) {  // the forward pass kept this path's radiance
            const float4 rr = recorded_L[size_t(k) * frame_px + idx];
            L = mk(rr.x, rr.y, rr.z);
          } else {
            Guide gd;
            L = trace_path(sc, f, ro, rd, rng, st, gd);
          }
          if (!(isnan(L.x) || isnan(L.y) || isnan(L.z) || isinf(L.x) || isinf(L.y) || isinf(L.z))) {
            rng = replay_rng;
            replay_begin(r, ro, rd, L);
            active = true;
          }
        }
      }
      handed += min(uint32_t(__popcll(idle)), avail);
      idle = __ballot(!active);
    }
    if (active && !replay_step(sc, f, r, rng, st, dl, g)) active = false;
    if (!taking && __ballot(active) == 0) break;
  }
  __syncthreads();
  const int nm = min(int(sc.material_count), kLdsMaterials);
  for (int i = threadIdx.x; i < nm * kNumParams; i += kAdjBlock)
    mat_partials[size_t(blockIdx.x) * kLdsMaterials * kNumParams + i] = lds_grad[i];
  flush_env(g, int(threadIdx.x), kAdjBlock);
}

// FP64 reductions: per-block material partials, then the env replicas.
__global__ void k_reduce_materials(const double* __restrict__ partials, uint32_t blocks, uint32_t entries,
                                   const double* __restrict__ overflow, uint32_t overflow_entries,
                                   double* __restrict__ out) {
  const uint32_t e = blockIdx.x;
  __shared__ double part[256];
  double acc = 0.0;
  if (e < entries)
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) acc += partials[size_t(b) * kLdsMaterials * kNumParams + e];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = e < entries ? part[0] : 0.0;
    if (e < overflow_entries) v += overflow[e];
    out[e] += v;
  }
}

__global__ void k_reduce_env(const double* __restrict__ acc, size_t stride, size_t n, double* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int r = 0; r < kEnvReplicas; ++r) s += acc[r * stride + i];
  out[i] += s;
}

__global__ void k_zero(double* p, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

size_t adjoint_scratch_bytes(uint32_t width, uint32_t height, uint32_t material_count, uint32_t env_texels) {
  const size_t blocks = size_t((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile);
  const size_t partials = blocks * kLdsMaterials * kNumParams;
  const size_t overflow = size_t(material_count) * kNumParams;
  const size_t env = size_t(kEnvReplicas) * env_texels * 3;
  // + the replay's job counter, then its AdjointView at a 256-byte boundary.
  return (partials + overflow + env + 1) * sizeof(double) + 256 + sizeof(AdjointView);
}

Status launch_detached(Context* ctx, const DetachedArgs& a, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const uint32_t tiles_x = (a.frame.width + kTile - 1) / kTile;
  const uint32_t tiles_y = (a.frame.height + kTile - 1) / kTile;
  const uint32_t blocks = tiles_x * tiles_y;
  const uint32_t depth = pt_spill_words(a.scene->bvh);
  Status s = ensure_spill(ctx, size_t(blocks) * kAdjBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  PtScene sv = *a.scene;
  sv.materials = a.value_materials;
  for (uint32_t k = 0; k < a.spp; ++k) {
    PtFrame f = a.frame;
    f.seed = a.seed0 + k;
    hipLaunchKernelGGL(k_detached, dim3(blocks), dim3(kAdjBlock), 0, stream, *a.scene, sv, f, 1.0f / float(a.spp),
                       k == 0 ? 1 : 0, a.out, static_cast<uint32_t*>(ctx->spill), depth, tiles_x);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "detached forward launch");
  return LUCENT_OK;
}

Status launch_adjoint(Context* ctx, const AdjointArgs& a, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const PtScene& sc = *a.scene;
  // Jobs: tiles of the launch shape (0 = 8 x 8, at most 64 pixels) with all
  // their samples.  The per-block partials hold one entry per 8 x 8 tile
  // (adjoint_scratch_bytes), which bounds the grid.
  const uint32_t tw = a.tile_w ? a.tile_w : kTile, th = a.tile_h ? a.tile_h : kTile;
  if (tw * th > 64) return fail(LUCENT_E_INVALID, "launch tile holds more than 64 pixels");
  const uint32_t tiles_x = (a.frame.width + tw - 1) / tw;
  const uint32_t tiles_y = (a.frame.height + th - 1) / th;
  const uint32_t n_jobs = tiles_x * tiles_y;
  const uint32_t blocks8 = ((a.frame.width + kTile - 1) / kTile) * ((a.frame.height + kTile - 1) / kTile);
  const uint32_t waves = (a.occupancy == 2 || a.occupancy == 3) ? a.occupancy : 4;
  const uint32_t grid =
      std::min<uint32_t>(std::min(n_jobs, blocks8), uint32_t(std::max(ctx->compute_units, 1)) * 4 * waves);
  const size_t env_texels = size_t(sc.env_w) * sc.env_h;
  const size_t partials = size_t(blocks8) * kLdsMaterials * kNumParams;
  const size_t overflow = size_t(sc.material_count) * kNumParams;
  const size_t env_stride = env_texels * 3;
  // Scratch: [block partials | overflow | env replicas | job counter | view] (adjoint_scratch_bytes).
  double* base = static_cast<double*>(a.scratch);
  double* mat_partials = base;
  double* mat_overflow = base + partials;
  double* env_acc = mat_overflow + overflow;
  uint32_t* counter = reinterpret_cast<uint32_t*>(env_acc + kEnvReplicas * env_stride);
  const size_t view_off = ((reinterpret_cast<char*>(counter + 2) - static_cast<char*>(a.scratch)) + 255) / 256 * 256;
  AdjointView* view = reinterpret_cast<AdjointView*>(static_cast<char*>(a.scratch) + view_off);
  const uint32_t depth = pt_spill_words(sc.bvh);
  Status s = ensure_spill(ctx, size_t(grid) * kAdjBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  // Overflow, env replicas and the counter (0.0 has the bits of 0u).
  const size_t zero_n = overflow + kEnvReplicas * env_stride + 1;
  hipLaunchKernelGGL(k_zero, dim3(uint32_t((zero_n + 255) / 256)), dim3(256), 0, stream, mat_overflow, zero_n);
  PtFrame f = a.frame;
  f.seed = a.seed0;
  const float4* rec = reinterpret_cast<const float4*>(a.recorded_radiance);
  hipLaunchKernelGGL(k_store_view, dim3(1), dim3(1), 0, stream, view, sc, f);
  auto kern = waves == 3 ? k_adjoint<3> : waves == 4 ? k_adjoint<4> : k_adjoint<2>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kAdjBlock), 0, stream, view, a.spp, n_jobs, counter, a.dl_di, rec,
                     1.0f / float(a.spp), mat_partials, env_acc, env_stride, mat_overflow,
                     static_cast<uint32_t*>(ctx->spill), depth, tiles_x, tw, th);
  const uint32_t entries = uint32_t(min(int(sc.material_count), kLdsMaterials)) * kNumParams;
  const uint32_t all = uint32_t(sc.material_count) * kNumParams;
  if (all > 0)
    hipLaunchKernelGGL(k_reduce_materials, dim3(all), dim3(256), 0, stream, mat_partials, grid, entries, mat_overflow, all,
                       a.grad_materials);
  if (env_stride > 0 && a.grad_env)
    hipLaunchKernelGGL(k_reduce_env, dim3(uint32_t((env_stride + 255) / 256)), dim3(256), 0, stream, env_acc, env_stride,
                       env_stride, a.grad_env);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "adjoint launch");
  return LUCENT_OK;
}

}  // namespace lucent
