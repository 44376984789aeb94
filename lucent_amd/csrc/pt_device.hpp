// Device-side integrator library: the float32 semantics of the reference
// shaders (src/shaders/common.glsl, scene.glsl, lights.glsl, bsdf.glsl,
// trace.frag) written for gfx950 wave64 lanes.  Included by the path-tracing
// translation units only.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "pt_types.hpp"

namespace lucent {
namespace pt {

constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.31830988618379067154f;
constexpr float TWO_PI = 6.28318530717958647692f;
constexpr float INF = 1e30f;
constexpr float EPS = 1e-6f;
constexpr float FAR = 1e4f;
constexpr float MIN_SHEEN_ROUGHNESS = 0.07f;

// Scene data read as global loads: the scene's arrays reach the kernels
// through pointers read from memory, which the compiler would otherwise
// access with flat instructions (counted on the LDS counter too, so that
// LDS waits also wait for them).  Global<T> indexes such an array.
typedef float gf32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const gf32x4 global_f4;
typedef __attribute__((address_space(1))) const float global_f32;
typedef __attribute__((address_space(1))) const uint32_t global_u32;
__device__ __forceinline__ float4 ld_global(const float4* p) {
  const gf32x4 v = *(global_f4*)p;
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float ld_global(const float* p) { return *(global_f32*)p; }
__device__ __forceinline__ uint32_t ld_global(const uint32_t* p) { return *(global_u32*)p; }
__device__ __forceinline__ int32_t ld_global(const int32_t* p) {
  return *(__attribute__((address_space(1))) const int32_t*)p;
}
__device__ __forceinline__ void st_global(float4* p, float4 v) {
  gf32x4 w;
  w.x = v.x;
  w.y = v.y;
  w.z = v.z;
  w.w = v.w;
  *(__attribute__((address_space(1))) gf32x4*)p = w;
}
template <class T>
struct Global {
  const T* p;
  __device__ __forceinline__ T operator[](size_t i) const { return ld_global(p + i); }
};

// ---------------------------------------------------------------------------
// Small vector algebra with GLSL meaning (normalize = v * inversesqrt(v.v)).
// ---------------------------------------------------------------------------
struct v3 {
  float x, y, z;
};
__device__ inline v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ inline v3 mk(const float* p) { return v3{p[0], p[1], p[2]}; }
__device__ inline v3 mk(float4 p) { return v3{p.x, p.y, p.z}; }
__device__ inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
__device__ inline v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ inline v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline v3 operator*(float s, v3 a) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline v3 operator/(v3 a, v3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline float length(v3 a) { return sqrtf(dot(a, a)); }
__device__ inline v3 normalize(v3 a) { return a * rsqrtf(dot(a, a)); }
__device__ inline v3 vmax(v3 a, float s) { return {fmaxf(a.x, s), fmaxf(a.y, s), fmaxf(a.z, s)}; }
__device__ inline float max3(v3 v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }
__device__ inline v3 mixv(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ inline float clampf(float x, float a, float b) { return fminf(fmaxf(x, a), b); }
// GLSL smoothstep.
__device__ inline float smoothstep(float e0, float e1, float x) {
  const float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}
__device__ inline v3 reflect(v3 i, v3 n) { return i - n * (2.0f * dot(n, i)); }
__device__ inline v3 refract(v3 i, v3 n, float eta) {
  const float d = dot(n, i);
  const float k = 1.0f - eta * eta * (1.0f - d * d);
  if (k < 0.0f) return {0.0f, 0.0f, 0.0f};
  return i * eta - n * (eta * d + sqrtf(k));
}
__device__ inline v3 vexp(v3 a) { return {expf(a.x), expf(a.y), expf(a.z)}; }
// GLSL pow(x, y) for x >= 0 as the specification defines it, exp2(y * log2(x)),
// on the hardware exp2/log2 (pow(0, y > 0) = exp2(-inf) = 0).
__device__ inline float gpow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ inline float luminance(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// ---------------------------------------------------------------------------
// common.glsl
// ---------------------------------------------------------------------------
// PCG hash stream (common.glsl:13-22); rand() may round to exactly 1.0.
struct Rng {
  uint32_t s;
  __device__ inline uint32_t next() {
    const uint32_t st = s;
    s = st * 747796405u + 2891336453u;
    const uint32_t word = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
    return (word >> 22u) ^ word;
  }
  __device__ inline float rand() { return float(next()) * (1.0f / 4294967296.0f); }
};

// Seeding of trace.frag:326-328.
__device__ inline Rng seed_rng(uint32_t pixel_index, uint32_t seed) {
  Rng r{pixel_index * 1973u + seed * 9277u + 26699u};
  r.next();
  return r;
}

__device__ inline void basis(v3 n, v3& t, v3& b) {
  const float s = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (s + n.z);
  const float c = n.x * n.y * a;
  t = mk(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
  b = mk(c, s + n.y * n.y * a, -n.y);
}
__device__ inline v3 to_world(v3 v, v3 t, v3 b, v3 n) { return t * v.x + b * v.y + n * v.z; }
__device__ inline v3 to_local(v3 v, v3 t, v3 b, v3 n) { return mk(dot(v, t), dot(v, b), dot(v, n)); }

__device__ inline v3 cosine_hemisphere(float u0, float u1) {
  const float r = sqrtf(u0);
  const float phi = TWO_PI * u1;
  return mk(r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - u0)));
}
__device__ inline v3 uniform_cone(float u0, float u1, float cos_max) {
  const float ct = 1.0f - u0 * (1.0f - cos_max);
  const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
  const float phi = TWO_PI * u1;
  return mk(st * cosf(phi), st * sinf(phi), ct);
}
__device__ inline void concentric_disk(float u0, float u1, float& dx, float& dy) {
  const float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
  if (ox == 0.0f && oy == 0.0f) {
    dx = dy = 0.0f;
    return;
  }
  float r, theta;
  if (fabsf(ox) > fabsf(oy)) {
    r = ox;
    theta = (PI / 4.0f) * (oy / ox);
  } else {
    r = oy;
    theta = (PI / 2.0f) - (PI / 4.0f) * (ox / oy);
  }
  dx = r * cosf(theta);
  dy = r * sinf(theta);
}
__device__ inline float mis_weight(float a, float b) {
  const float a2 = a * a, b2 = b * b;
  return a2 / fmaxf(a2 + b2, 1e-10f);
}
// Wächter & Binder offset (common.glsl:95-108).
__device__ inline float offset_axis(float p, float n) {
  const int o = int(256.0f * n);
  const float pi = __int_as_float(__float_as_int(p) + (p < 0.0f ? -o : o));
  return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : pi;
}
__device__ inline v3 offset_ray(v3 p, v3 n) { return mk(offset_axis(p.x, n.x), offset_axis(p.y, n.y), offset_axis(p.z, n.z)); }

// ---------------------------------------------------------------------------
// Textures: bilinear RGBA8 arrays, REPEAT, level 0 (context.ts:238-242).
// ---------------------------------------------------------------------------
__device__ inline float srgb_to_linear(float c) {
  return c <= 0.04045f ? c * (1.0f / 12.92f) : gpow((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}
__device__ inline float4 texel_rgba8(const uint32_t* maps, int size, int layer, int x, int y, bool srgb) {
  x = ((x % size) + size) % size;
  y = ((y % size) + size) % size;
  const uint32_t p = ld_global(maps + (size_t(layer) * size + y) * size + x);
  float4 c = make_float4(float(p & 255u) / 255.0f, float((p >> 8) & 255u) / 255.0f, float((p >> 16) & 255u) / 255.0f,
                         float(p >> 24) / 255.0f);
  if (srgb) {
    c.x = srgb_to_linear(c.x);
    c.y = srgb_to_linear(c.y);
    c.z = srgb_to_linear(c.z);
  }
  return c;
}
__device__ inline float4 sample_layer(const uint32_t* maps, int size, int layer, float u, float v, bool srgb) {
  const float x = u * float(size) - 0.5f, y = v * float(size) - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const int x0 = int(fx), y0 = int(fy);
  const float tx = x - fx, ty = y - fy;
  const float4 a = texel_rgba8(maps, size, layer, x0, y0, srgb);
  const float4 b = texel_rgba8(maps, size, layer, x0 + 1, y0, srgb);
  const float4 c = texel_rgba8(maps, size, layer, x0, y0 + 1, srgb);
  const float4 d = texel_rgba8(maps, size, layer, x0 + 1, y0 + 1, srgb);
  auto lerp4 = [](float4 p, float4 q, float t) {
    return make_float4(p.x + (q.x - p.x) * t, p.y + (q.y - p.y) * t, p.z + (q.z - p.z) * t, p.w + (q.w - p.w) * t);
  };
  return lerp4(lerp4(a, b, tx), lerp4(c, d, tx), ty);
}

// ---------------------------------------------------------------------------
// scene.glsl: materials, hits, traversal
// ---------------------------------------------------------------------------
struct Material {
  v3 albedo;
  float metallic, roughness, ior, transmission, checker_scale;
  v3 emission;
  float alpha;
  v3 albedo2;
  int alpha_mode;
  int map_base, map_mr, map_normal, map_emissive;
  float normal_scale, clearcoat, clearcoat_roughness, alpha_cutoff;
  v3 sheen_color;
  float sheen_roughness;
  v3 atten_color;
  float atten_distance;
  v3 uv0, uv1;
  bool thin_walled;
};

__device__ inline Material fetch_material(const PtScene& s, int index) {
  const Global<float4> m{reinterpret_cast<const float4*>(s.materials) + size_t(index) * kMaterialTexels};
  const float4 t0 = m[0], t1 = m[1], t2 = m[2], t3 = m[3], t4 = m[4], t5 = m[5], t6 = m[6], t7 = m[7], t8 = m[8],
               t9 = m[9];
  Material r;
  r.albedo = mk(t0);
  r.metallic = t0.w;
  r.roughness = t1.x;
  r.ior = t1.y;
  r.transmission = t1.z;
  r.checker_scale = t1.w;
  r.emission = mk(t2);
  r.alpha = t2.w;
  r.albedo2 = mk(t3);
  r.alpha_mode = int(t3.w);
  r.map_base = int(t4.x);
  r.map_mr = int(t4.y);
  r.map_normal = int(t4.z);
  r.map_emissive = int(t4.w);
  r.normal_scale = t5.x;
  r.clearcoat = t5.y;
  r.clearcoat_roughness = t5.z;
  r.alpha_cutoff = t5.w;
  r.sheen_color = mk(t6);
  r.sheen_roughness = t6.w;
  r.atten_color = mk(t7);
  r.atten_distance = t7.w;
  r.uv0 = mk(t8);
  r.thin_walled = t8.w > 0.5f;
  r.uv1 = mk(t9);
  return r;
}

__device__ inline void transform_uv(float u, float v, v3 row0, v3 row1, float& tu, float& tv) {
  tu = row0.x * u + row0.y * v + row0.z;
  tv = row1.x * u + row1.y * v + row1.z;
}

struct Hit {
  float t;
  v3 position, normal, geom_normal;
  bool front_face;
  int material, light;
  float uv_u, uv_v;
  v3 dpdu, dpdv;
};

// hitTriangle (scene.glsl:221-238).
__device__ inline bool hit_triangle(v3 ro, v3 rd, v3 v0, v3 v1, v3 v2, float tmax, float& t, float& bu, float& bv) {
  const v3 e1 = v1 - v0, e2 = v2 - v0;
  const v3 p = cross(rd, e2);
  const float det = dot(e1, p);
  if (fabsf(det) < 1e-12f) return false;
  // GLSL's division is accurate to 2.5 ulp; the hardware reciprocal (1 ulp)
  // stays within that and replaces the IEEE division sequence.
  const float inv_det = __builtin_amdgcn_rcpf(det);
  const v3 s = ro - v0;
  const float u = dot(s, p) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  const v3 q = cross(s, e1);
  const float v = dot(rd, q) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = dot(e2, q) * inv_det;
  if (t <= EPS || t >= tmax) return false;
  bu = u;
  bv = v;
  return true;
}

__device__ inline float hit_sphere(v3 ro, v3 rd, float4 s, float tmax) {
  const v3 oc = ro - mk(s);
  const float b = dot(oc, rd);
  const float c = dot(oc, oc) - s.w * s.w;
  const float disc = b * b - c;
  if (disc < 0.0f) return INF;
  const float sq = sqrtf(disc);
  float t = -b - sq;
  if (t > EPS && t < tmax) return t;
  t = -b + sq;
  if (t > EPS && t < tmax) return t;
  return INF;
}

__device__ inline float hit_ground(const PtScene& s, v3 ro, v3 rd, float tmax) {
  if (s.ground_material < 0 || fabsf(rd.y) < 1e-9f) return INF;
  const float t = (s.ground_height - ro.y) / rd.y;
  return t > EPS && t < tmax ? t : INF;
}

// Interpolated uv of a leaf slot (triangleUV, scene.glsl:271-281).
__device__ inline void triangle_uvs(const PtScene& s, uint32_t slot, float& u0, float& v0, float& u1, float& v1,
                                    float& u2, float& v2) {
  const Global<float4> a{reinterpret_cast<const float4*>(s.bvh.attrs) + size_t(slot) * 4};
  const float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
  u0 = a0.w;
  v0 = a3.x;
  u1 = a1.w;
  v1 = a3.y;
  u2 = a2.w;
  v2 = a3.z;
}

// alphaOpaque (scene.glsl:284-297): consumes a random number in blend mode.
__device__ inline bool alpha_opaque(const PtScene& s, uint32_t slot, float bu, float bv, int mat, Rng& rng) {
  float u0 = 0, v0 = 0, u1 = 0, v1 = 0, u2 = 0, v2 = 0;
  if (s.bvh.attrs) triangle_uvs(s, slot, u0, v0, u1, v1, u2, v2);
  const float w = 1.0f - bu - bv;
  const float uu = u0 * w + u1 * bu + u2 * bv, vv = v0 * w + v1 * bu + v2 * bv;
  const Global<float4> m{reinterpret_cast<const float4*>(s.materials) + size_t(mat) * kMaterialTexels};
  float alpha = m[2].w;
  const int mode = int(m[3].w);
  const int map = int(m[4].x);
  if (map >= 0 && s.color_maps && map < s.color_layers) {
    float tu, tv;
    transform_uv(uu, vv, mk(m[8]), mk(m[9]), tu, tv);
    alpha *= sample_layer(s.color_maps, s.tex_size, map, tu, tv, true).w;
  }
  if (mode == 1) return alpha >= m[5].w;
  return rng.rand() < alpha;
}

// ---------------------------------------------------------------------------
// Traversal of the 4-wide BVH (DevBvh::nodes4).  A stack of (link, entry
// distance) pairs; one step pops an entry or tests one triangle of the leaf
// being worked through.  A popped node has its four child boxes slab-tested
// (hitAABB, scene.glsl:261-269: tn = max(min(t0,t1).xyz, 0), tf = min(max
// (t0,t1).xyz, tmax), hit iff tf >= tn) and its hit children pushed, farthest
// first, so the nearest is popped next; entries at or beyond the closest hit
// so far are dropped when popped (a triangle must be strictly nearer to be
// accepted).  Shadow rays push hit children unsorted and stop at the first
// accepted triangle.  The closest hit equals that of any complete traversal
// (ties between equal distances aside).
// ---------------------------------------------------------------------------
constexpr uint32_t kLinkLeafBit = 0x80000000u, kLinkNone = 0xFFFFFFFFu;
__device__ inline bool link_is_leaf(uint32_t l) { return (l & kLinkLeafBit) != 0u; }
__device__ inline uint32_t link_first(uint32_t l) { return l & 0x03FFFFFFu; }
__device__ inline uint32_t link_count(uint32_t l) { return ((l >> 26) & 31u) + 1u; }

// (link, distance) stack: kLds entries per lane in LDS (two lane-major
// planes, a ring), deeper ones spilled to 2 words per entry of a per-lane
// area.  Entries [base, sp) are in the ring at slots top-1, ..., bot
// (mod kLds); [0, base) are spilled.  The slot indices are kept
// incrementally (no division by kLds).
// A pointer into LDS typed as such: every stack access compiles to a ds_
// instruction.  (Through a generic pointer the compiler merges the LDS read
// and the rare spill read of a pop into one flat load, which waits on both
// the LDS and the vector-memory counters.)
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ inline lds_u32* as_lds(uint32_t* p) { return (lds_u32*)p; }

template <int kLds, int kBlock>
struct Stack4 {
  static_assert(kLds >= 4, "a node visit reserves ring room for its four children at once");
  lds_u32* lds;     // this lane's word of plane 0; plane 1 starts kLds * kBlock words later
  uint32_t* spill;  // 2 words per entry
  uint32_t sp, base;
  uint32_t top, bot;  // sp % kLds, base % kLds
  static constexpr bool kPow2 = (kLds & (kLds - 1)) == 0;
  __device__ static inline uint32_t next(uint32_t i) {
    if constexpr (kPow2)
      return (i + 1u) & uint32_t(kLds - 1);
    else
      return i + 1u == uint32_t(kLds) ? 0u : i + 1u;
  }
  __device__ static inline uint32_t prev(uint32_t i) {
    if constexpr (kPow2)
      return (i - 1u) & uint32_t(kLds - 1);
    else
      return i == 0u ? uint32_t(kLds - 1) : i - 1u;
  }
  __device__ inline void reset() { sp = base = top = bot = 0; }
  __device__ inline bool empty() const { return sp == 0; }
  // Room in the ring for n more entries: the oldest entries move to the
  // spill area (the same entries, in the same order, that n pushes that
  // each spill when the ring is full would move; n <= kLds).
  __device__ inline void reserve(uint32_t n) {
    while (sp - base + n > uint32_t(kLds)) {
      spill[2 * base] = lds[bot * kBlock];
      spill[2 * base + 1] = lds[(kLds + bot) * kBlock];
      ++base;
      bot = next(bot);
    }
  }
  // A push into a ring with room for it.
  __device__ inline void push_reserved(uint32_t link, float dist) {
    lds[top * kBlock] = link;
    lds[(kLds + top) * kBlock] = __float_as_uint(dist);
    ++sp;
    top = next(top);
  }
  __device__ inline void push(uint32_t link, float dist) {
    reserve(1);
    push_reserved(link, dist);
  }
  __device__ inline void pop(uint32_t& link, float& dist) {
    --sp;
    top = prev(top);
    if (sp >= base) {
      link = lds[top * kBlock];
      dist = __uint_as_float(lds[(kLds + top) * kBlock]);
      return;
    }
    --base;
    bot = top;
    link = spill[2 * sp];
    dist = __uint_as_float(spill[2 * sp + 1]);
  }
};

constexpr int kStack4Lds = 12;
constexpr int kPtStackWords = 2 * kStack4Lds;  // LDS words per lane
template <int kBlock>
using PtStack = Stack4<kStack4Lds, kBlock>;
// Spill words per lane for a scene's 4-wide BVH.
__host__ __device__ inline uint32_t pt_spill_words(const DevBvh& b) { return 2u * b.stack4; }

// Entry distances of the four children of 4-wide node n (INF: no child or
// missed).  A child's slab distance (origin + code * scale - o) * inv is one
// multiply-add per plane, code * (scale * inv) + (origin - o) * inv, and each
// child box is widened by 1/16 of a quantisation step on both sides, which
// covers the rounding of that split form (a few ulps of the terms) with room
// to spare: no box that the ray touches is missed.
__device__ inline void node4_hits(const float* nodes, uint32_t n, v3 o, v3 inv, float tmax, float tn[4], uint32_t link[4]) {
  const float4* r = reinterpret_cast<const float4*>(nodes) + size_t(n) * 4;
  const float4 A = r[0], Bq = r[1], Cq = r[2], Lf = r[3];
  const float sx = A.w, sy = Bq.x, sz = Bq.y;
  const uint32_t lx = __float_as_uint(Bq.z), hx = __float_as_uint(Bq.w);
  const uint32_t ly = __float_as_uint(Cq.x), hy = __float_as_uint(Cq.y);
  const uint32_t lz = __float_as_uint(Cq.z), hz = __float_as_uint(Cq.w);
  link[0] = __float_as_uint(Lf.x);
  link[1] = __float_as_uint(Lf.y);
  link[2] = __float_as_uint(Lf.z);
  link[3] = __float_as_uint(Lf.w);
  const float kx = sx * inv.x, ky = sy * inv.y, kz = sz * inv.z;
  const float dx = A.x - o.x, dy = A.y - o.y, dz = A.z - o.z;
  const float lox = fmaf(sx, -0.0625f, dx) * inv.x, hix = fmaf(sx, 0.0625f, dx) * inv.x;
  const float loy = fmaf(sy, -0.0625f, dy) * inv.y, hiy = fmaf(sy, 0.0625f, dy) * inv.y;
  const float loz = fmaf(sz, -0.0625f, dz) * inv.z, hiz = fmaf(sz, 0.0625f, dz) * inv.z;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int sh = 8 * i;
    const float x0 = fmaf(float((lx >> sh) & 255u), kx, lox), x1 = fmaf(float((hx >> sh) & 255u), kx, hix);
    const float y0 = fmaf(float((ly >> sh) & 255u), ky, loy), y1 = fmaf(float((hy >> sh) & 255u), ky, hiy);
    const float z0 = fmaf(float((lz >> sh) & 255u), kz, loz), z1 = fmaf(float((hz >> sh) & 255u), kz, hiz);
    const float nn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float ff = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    tn[i] = (link[i] != kLinkNone && ff >= nn) ? nn : INF;
  }
}

// The same test on the float32 records (DevBvh::nodes4f): child bounds are
// offsets from the node origin, so a plane is one multiply-add onto the
// per-node (origin - o) * inv, with no decoding.
__device__ inline void node4_hits_f32(const float* nodes, uint32_t n, v3 o, v3 inv, float tmax, float tn[4],
                                      uint32_t link[4]) {
  const float4* r = reinterpret_cast<const float4*>(nodes) + size_t(n) * 8;
  const float4 A = ld_global(r), lx = ld_global(r + 1), hx = ld_global(r + 2), ly = ld_global(r + 3),
               hy = ld_global(r + 4), lz = ld_global(r + 5), hz = ld_global(r + 6), L = ld_global(r + 7);
  link[0] = __float_as_uint(L.x);
  link[1] = __float_as_uint(L.y);
  link[2] = __float_as_uint(L.z);
  link[3] = __float_as_uint(L.w);
  const float bx = (A.x - o.x) * inv.x, by = (A.y - o.y) * inv.y, bz = (A.z - o.z) * inv.z;
  const float LX[4] = {lx.x, lx.y, lx.z, lx.w}, HX[4] = {hx.x, hx.y, hx.z, hx.w};
  const float LY[4] = {ly.x, ly.y, ly.z, ly.w}, HY[4] = {hy.x, hy.y, hy.z, hy.w};
  const float LZ[4] = {lz.x, lz.y, lz.z, lz.w}, HZ[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float x0 = fmaf(LX[i], inv.x, bx), x1 = fmaf(HX[i], inv.x, bx);
    const float y0 = fmaf(LY[i], inv.y, by), y1 = fmaf(HY[i], inv.y, by);
    const float z0 = fmaf(LZ[i], inv.z, bz), z1 = fmaf(HZ[i], inv.z, bz);
    const float nn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
    const float ff = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
    tn[i] = (link[i] != kLinkNone && ff >= nn) ? nn : INF;
  }
}

struct Trav4 {
  v3 ro, rd, inv;
  float tmax;  // closest hit so far (closest queries) or the ray's extent (shadow rays)
  float bu, bv;
  int best;
  uint32_t la, na;  // leaf being tested: next slot, triangles left
};

template <class StackT>
__device__ inline void trav4_begin(const PtScene& s, Trav4& t, v3 ro, v3 rd, float tmax, StackT& st) {
  t.ro = ro;
  t.rd = rd;
  t.inv = mk(1.0f / rd.x, 1.0f / rd.y, 1.0f / rd.z);
  t.tmax = tmax;
  t.bu = t.bv = 0.0f;
  t.best = -1;
  t.la = t.na = 0;
  st.reset();
  if (s.bvh.tri_count > 0) st.push(s.bvh.root4, 0.0f);
}

template <class StackT>
__device__ inline bool trav4_done(const Trav4& t, const StackT& st) { return t.na == 0 && st.empty(); }

// Sort key of a missed child: INF's bits with the two index bits cleared.
constexpr uint32_t kMissKey = 0x7149F2CAu & ~3u;

// Tests the pending leaf's next triangle (t.na > 0); for shadow rays (any)
// true when it is accepted, for closest rays it updates the hit.
__device__ __forceinline__ bool test_next_triangle(const PtScene& s, Trav4& t, Rng& rng, bool any) {
  const uint32_t k = t.la++;
  --t.na;
  const float4* tris = reinterpret_cast<const float4*>(s.bvh.tris);
  const float4* tp = tris + size_t(k) * 3;
  const float4 p0 = ld_global(tp), p1 = ld_global(tp + 1), p2 = ld_global(tp + 2);
  float th, u, v;
  if (hit_triangle(t.ro, t.rd, mk(p0), mk(p1), mk(p2), t.tmax, th, u, v) &&
      (__float_as_int(p2.w) == 0 || alpha_opaque(s, k, u, v, __float_as_int(p0.w), rng))) {
    if (any) return true;
    t.tmax = th;
    t.best = int(k);
    t.bu = u;
    t.bv = v;
  }
  return false;
}

// Pushes the children whose keys are not kMissKey in descending key order
// (the key's two low bits are the child index, so keys are distinct): a
// sorting network of compare-exchanges that carry each key's link, then one
// check for ring room for all of them.
template <class StackT>
__device__ __forceinline__ void push_sorted(StackT& st, uint32_t k[4], uint32_t l[4]) {
  auto cx = [&](int a, int b) {
    const bool swap = k[a] < k[b];
    const uint32_t ka = k[a], kb = k[b], la = l[a], lb = l[b];
    k[a] = swap ? kb : ka;
    k[b] = swap ? ka : kb;
    l[a] = swap ? lb : la;
    l[b] = swap ? la : lb;
  };
  cx(0, 1);
  cx(2, 3);
  cx(0, 2);
  cx(1, 3);
  cx(1, 2);
  bool hit[4];
  uint32_t n = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    hit[i] = (k[i] & ~3u) != kMissKey;
    n += uint32_t(hit[i]);
  }
  st.reserve(n);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (hit[i]) st.push_reserved(l[i], __uint_as_float(k[i] & ~3u));
}

// One step: with no leaf pending, pop an entry and visit it (a node's
// children are pushed; a leaf's first triangle is tested in the same step),
// else test the pending leaf's next triangle.  For shadow rays (kAny) true
// on the first accepted triangle.
template <bool kAny, class StackT>
__device__ inline bool trav4_step(const PtScene& s, Trav4& t, StackT& st, Rng& rng) {
  if (t.na == 0) {
    uint32_t link;
    float dist;
    st.pop(link, dist);
    if (!kAny && dist >= t.tmax) return false;
    if (!link_is_leaf(link)) {
      float tn[4];
      uint32_t l[4];
#ifdef LUCENT_PT_QUANTIZED_NODES
      node4_hits(s.bvh.nodes4, link, t.ro, t.inv, t.tmax, tn, l);
#else
      node4_hits_f32(s.bvh.nodes4f, link, t.ro, t.inv, t.tmax, tn, l);
#endif
      if (kAny) {
        st.reserve(uint32_t(tn[0] != INF) + uint32_t(tn[1] != INF) + uint32_t(tn[2] != INF) + uint32_t(tn[3] != INF));
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (tn[i] != INF) st.push_reserved(l[i], tn[i]);
        return false;
      }
      // Farthest first, so that the nearest child ends on top.  Entry
      // distances are >= 0, so their bit patterns order like the floats: the
      // keys carry the child index in the two low bits (the distance
      // truncated by up to 3 ulps, which only weakens culling; distinct
      // keys, so the order is fixed) and sort descending with a branch-free
      // network that moves each child's link with its key; missed children
      // (kMissKey) sort first and are not pushed.
      uint32_t k[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) k[i] = (__float_as_uint(tn[i]) & ~3u) | uint32_t(i);
      push_sorted(st, k, l);
      return false;
    }
    t.la = link_first(link);
    t.na = link_count(link);
  }
  return test_next_triangle(s, t, rng, kAny);
}

// trav4_step for a ray whose kind is a run-time value: one body for closest
// rays (any = false, children nearest first, as trav4_step<false>) and
// shadow rays (any = true, children pushed in index order, true on the first
// accepted triangle, as trav4_step<true>), so that a wave whose lanes trace
// both kinds runs the node and leaf code once.  Shadow rays ride the same
// sorting network with keys that keep their index order.
template <class StackT>
__device__ inline bool trav4_step_mixed(const PtScene& s, Trav4& t, StackT& st, Rng& rng, bool any) {
  if (t.na == 0) {
    uint32_t link;
    float dist;
    st.pop(link, dist);
    if (!any && dist >= t.tmax) return false;
    if (!link_is_leaf(link)) {
      float tn[4];
      uint32_t l[4];
#ifdef LUCENT_PT_QUANTIZED_NODES
      node4_hits(s.bvh.nodes4, link, t.ro, t.inv, t.tmax, tn, l);
#else
      node4_hits_f32(s.bvh.nodes4f, link, t.ro, t.inv, t.tmax, tn, l);
#endif
      // Closest rays: the entry distance's bits with the child index in the
      // two low bits (as trav4_step).  Shadow rays: (3 - i) above the index,
      // so the descending sort leaves them in index order; missed children
      // sort first with kMissKey either way and are not pushed.
      // (A missed child's closest-ray key is kMissKey | i already: its
      // distance is INF.)
      uint32_t k[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t closest = (__float_as_uint(tn[i]) & ~3u) | uint32_t(i);
        const uint32_t shadow = ((3u - uint32_t(i)) << 2) | uint32_t(i);
        k[i] = any && tn[i] != INF ? shadow : closest;
      }
      push_sorted(st, k, l);
      return false;
    }
    t.la = link_first(link);
    t.na = link_count(link);
  }
  return test_next_triangle(s, t, rng, any);
}

// traceScene's triangle part (scene.glsl:321-376): the closest accepted hit
// with t < closest.  Returns the best slot or -1.
template <class StackT>
__device__ inline int trace_bvh(const PtScene& s, v3 ro, v3 rd, float& closest, float& bu, float& bv, Rng& rng,
                                StackT& st) {
  Trav4 t;
  trav4_begin(s, t, ro, rd, closest, st);
  while (!trav4_done(t, st)) trav4_step<false>(s, t, st, rng);
  if (t.best >= 0) {
    closest = t.tmax;
    bu = t.bu;
    bv = t.bv;
  }
  return t.best;
}

// occluded (scene.glsl:336-377): any accepted hit below tmax.
template <class StackT>
__device__ inline bool occluded(const PtScene& s, v3 ro, v3 rd, float tmax, Rng& rng, StackT& st) {
  for (uint32_t i = 0; i < s.sphere_count; ++i) {
    const float4 sp = ld_global(reinterpret_cast<const float4*>(s.spheres) + i);
    if (hit_sphere(ro, rd, sp, tmax) < tmax) return true;
  }
  if (hit_ground(s, ro, rd, tmax) < tmax) return true;
  Trav4 t;
  trav4_begin(s, t, ro, rd, tmax, st);
  while (!trav4_done(t, st))
    if (trav4_step<true>(s, t, st, rng)) return true;
  return false;
}

// The closest primitive found by traceScene's search (scene.glsl:192-333):
// kind 0 = miss, 1 = triangle (prim = leaf slot), 2 = sphere, 3 = ground.
enum : uint32_t { kHitMiss = 0, kHitTri = 1, kHitSphere = 2, kHitGround = 3 };
struct HitRecord {
  float t, u, v;
  uint32_t kind, prim;
};

// Spheres, then the ground, with t < closest (scene.glsl:307-320).
__device__ inline void search_analytic(const PtScene& s, v3 ro, v3 rd, HitRecord& r) {
  r.t = INF;
  r.kind = kHitMiss;
  r.prim = 0;
  r.u = r.v = 0.0f;
  for (uint32_t i = 0; i < s.sphere_count; ++i) {
    const float t = hit_sphere(ro, rd, ld_global(reinterpret_cast<const float4*>(s.spheres) + i), r.t);
    if (t < r.t) {
      r.t = t;
      r.kind = kHitSphere;
      r.prim = i;
    }
  }
  const float tg = hit_ground(s, ro, rd, r.t);
  if (tg < r.t) {
    r.t = tg;
    r.kind = kHitGround;
  }
}

// The hit record of traceScene from its search result.
__device__ inline bool build_hit(const PtScene& s, v3 ro, v3 rd, const HitRecord& r, Hit& hit) {
  if (r.kind == kHitMiss) return false;
  const float closest = r.t;
  hit.t = closest;
  hit.position = ro + rd * closest;
  hit.uv_u = hit.uv_v = 0.0f;
  hit.dpdu = hit.dpdv = mk(0.0f, 0.0f, 0.0f);
  hit.light = -1;
  if (r.kind == kHitGround) {
    hit.front_face = rd.y < 0.0f;
    hit.normal = mk(0.0f, hit.front_face ? 1.0f : -1.0f, 0.0f);
    hit.geom_normal = hit.normal;
    hit.material = s.ground_material;
    hit.uv_u = hit.position.x;
    hit.uv_v = hit.position.z;
    return true;
  }
  if (r.kind == kHitSphere) {
    const float4 sp = ld_global(reinterpret_cast<const float4*>(s.spheres) + r.prim);
    const v3 n = (hit.position - mk(sp)) / sp.w;
    hit.front_face = dot(rd, n) < 0.0f;
    hit.normal = hit.front_face ? n : -n;
    hit.geom_normal = hit.normal;
    hit.material = ld_global(s.sphere_material + r.prim);
    hit.light = ld_global(s.sphere_light + r.prim);
    return true;
  }
  const float bu = r.u, bv = r.v;
  const Global<float4> tris{reinterpret_cast<const float4*>(s.bvh.tris) + size_t(r.prim) * 3};
  const float4 p0 = tris[0], p1 = tris[1], p2 = tris[2];
  v3 n0 = mk(0.f, 0.f, 0.f), n1 = n0, n2 = n0;
  float u0 = 0, v0 = 0, u1 = 0, v1 = 0, u2 = 0, v2 = 0;
  const v3 e1 = mk(p1) - mk(p0), e2 = mk(p2) - mk(p0);
  v3 ng = normalize(cross(e1, e2));
  v3 ns;
  const float w = 1.0f - bu - bv;
  if (s.bvh.attrs) {
    const Global<float4> a{reinterpret_cast<const float4*>(s.bvh.attrs) + size_t(r.prim) * 4};
    const float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    n0 = mk(a0);
    n1 = mk(a1);
    n2 = mk(a2);
    u0 = a0.w;
    v0 = a3.x;
    u1 = a1.w;
    v1 = a3.y;
    u2 = a2.w;
    v2 = a3.z;
    hit.uv_u = u0 * w + u1 * bu + u2 * bv;
    hit.uv_v = v0 * w + v1 * bu + v2 * bv;
    const float d1x = u1 - u0, d1y = v1 - v0, d2x = u2 - u0, d2y = v2 - v0;
    const float det = d1x * d2y - d1y * d2x;
    if (fabsf(det) > 1e-12f) {
      const float rr = 1.0f / det;
      hit.dpdu = (e1 * d2y - e2 * d1y) * rr;
      hit.dpdv = (e2 * d1x - e1 * d2x) * rr;
    }
    ns = normalize(n0 * w + n1 * bu + n2 * bv);
  } else {
    ns = ng;
  }
  hit.front_face = dot(rd, ng) < 0.0f;
  if (!hit.front_face) {
    ng = -ng;
    ns = -ns;
  }
  if (dot(ns, ng) < 0.05f) ns = ng;
  hit.normal = ns;
  hit.geom_normal = ng;
  hit.material = __float_as_int(p0.w);
  hit.light = __float_as_int(p1.w);
  return true;
}

// traceScene (scene.glsl:192-333): spheres, ground, then triangles.
template <class StackT>
__device__ inline bool trace_scene(const PtScene& s, v3 ro, v3 rd, Hit& hit, Rng& rng, StackT& st) {
  HitRecord r;
  search_analytic(s, ro, rd, r);
  float bu = 0.0f, bv = 0.0f;
  float closest = r.t;
  const int best_tri = trace_bvh(s, ro, rd, closest, bu, bv, rng, st);
  if (best_tri >= 0) {
    r.t = closest;
    r.kind = kHitTri;
    r.prim = uint32_t(best_tri);
    r.u = bu;
    r.v = bv;
  }
  return build_hit(s, ro, rd, r, hit);
}

// ---------------------------------------------------------------------------
// lights.glsl
// ---------------------------------------------------------------------------
struct LightSample {
  v3 dir;
  float dist, pdf;
  v3 radiance;
  float u, v;  // environment samples: the map coordinates looked up
};

__device__ inline v3 env_to_local(const PtScene& s, v3 d) {
  return mk(s.env_cos * d.x - s.env_sin * d.z, d.y, s.env_sin * d.x + s.env_cos * d.z);
}
__device__ inline v3 env_to_world(const PtScene& s, v3 l) {
  return mk(s.env_cos * l.x + s.env_sin * l.z, l.y, -s.env_sin * l.x + s.env_cos * l.z);
}
__device__ inline void equirect_uv(v3 l, float& u, float& v) {
  u = atan2f(l.z, l.x) * (0.5f * INV_PI) + 0.5f;
  v = acosf(clampf(l.y, -1.0f, 1.0f)) * INV_PI;
}
// Texels and weights of the bilinear lookup at (u, v): LINEAR filter,
// REPEAT in u, CLAMP_TO_EDGE in v (renderer.ts:397-403).  Order: (x0,y0),
// (x1,y0), (x0,y1), (x1,y1).
__device__ inline void env_taps(const PtScene& s, float u, float v, int idx[4], float w[4]) {
  const float x = u * float(s.env_w) - 0.5f, y = v * float(s.env_h) - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float tx = x - fx, ty = y - fy;
  int x0 = int(fx) % s.env_w;
  if (x0 < 0) x0 += s.env_w;
  const int x1 = x0 + 1 == s.env_w ? 0 : x0 + 1;
  const int y0 = min(max(int(fy), 0), s.env_h - 1), y1 = min(max(int(fy) + 1, 0), s.env_h - 1);
  idx[0] = y0 * s.env_w + x0;
  idx[1] = y0 * s.env_w + x1;
  idx[2] = y1 * s.env_w + x0;
  idx[3] = y1 * s.env_w + x1;
  w[0] = (1.0f - tx) * (1.0f - ty);
  w[1] = tx * (1.0f - ty);
  w[2] = (1.0f - tx) * ty;
  w[3] = tx * ty;
}
__device__ inline v3 env_lookup(const PtScene& s, float u, float v) {
  const float x = u * float(s.env_w) - 0.5f, y = v * float(s.env_h) - 0.5f;
  const float fx = floorf(x), fy = floorf(y);
  const float tx = x - fx, ty = y - fy;
  int x0 = int(fx) % s.env_w;
  if (x0 < 0) x0 += s.env_w;
  const int x1 = x0 + 1 == s.env_w ? 0 : x0 + 1;
  const int y0 = min(max(int(fy), 0), s.env_h - 1), y1 = min(max(int(fy) + 1, 0), s.env_h - 1);
  const Global<float4> m{reinterpret_cast<const float4*>(s.env_map)};
  const float4 a = m[size_t(y0) * s.env_w + x0], b = m[size_t(y0) * s.env_w + x1];
  const float4 c = m[size_t(y1) * s.env_w + x0], d = m[size_t(y1) * s.env_w + x1];
  const v3 top = mk(a) * (1.0f - tx) + mk(b) * tx;
  const v3 bot = mk(c) * (1.0f - tx) + mk(d) * tx;
  return (top * (1.0f - ty) + bot * ty) * s.env_intensity;
}
__device__ inline v3 env_radiance(const PtScene& s, v3 d) {
  float u, v;
  equirect_uv(env_to_local(s, d), u, v);
  return env_lookup(s, u, v);
}
__device__ inline float env_pdf(const PtScene& s, v3 d) {
  const v3 l = env_to_local(s, d);
  const float st = sqrtf(fmaxf(0.0f, 1.0f - l.y * l.y));
  if (st < 1e-6f) return 0.0f;
  float u, v;
  equirect_uv(l, u, v);
  const int cx = min(int(u * float(s.grid_w)), s.grid_w - 1);
  const int cy = min(int(v * float(s.grid_h)), s.grid_h - 1);
  return ld_global(reinterpret_cast<const float4*>(s.env_table) + size_t(cy) * s.grid_w + cx).z /
         (2.0f * PI * PI * st);
}
__device__ inline LightSample sample_environment(const PtScene& s, Rng& rng) {
  LightSample ls;
  const int n = s.grid_w * s.grid_h;
  const int i = min(int(rng.rand() * float(n)), n - 1);
  const float4 e = ld_global(reinterpret_cast<const float4*>(s.env_table) + i);
  int cell = i;
  float pdf_uv = e.z;
  if (rng.rand() >= e.x) {
    cell = int(e.y);
    pdf_uv = e.w;
  }
  const float r0 = rng.rand();
  const float r1 = rng.rand();
  const float u = (float(cell % s.grid_w) + r0) / float(s.grid_w);
  const float v = (float(cell / s.grid_w) + r1) / float(s.grid_h);
  const float phi = (u - 0.5f) * TWO_PI;
  const float theta = v * PI;
  const float st = sinf(theta);
  ls.dir = env_to_world(s, mk(st * cosf(phi), cosf(theta), st * sinf(phi)));
  ls.dist = INF;
  ls.pdf = st > 1e-6f ? pdf_uv / (2.0f * PI * PI * st) : 0.0f;
  ls.radiance = env_lookup(s, u, v);
  ls.u = u;
  ls.v = v;
  return ls;
}
__device__ inline v3 environment(const PtScene& s, v3 d, bool specular, float bsdf_pdf) {
  v3 L = env_radiance(s, d);
  if (!specular && s.env_sampling != 0) L = L * mis_weight(bsdf_pdf, env_pdf(s, d));
  if (s.sun_enabled != 0 && dot(d, mk(s.sun_dir)) >= s.sun_cos) {
    const float w = specular ? 1.0f : mis_weight(bsdf_pdf, s.sun_pdf);
    L = L + mk(s.sun_radiance) * w;
  }
  return L;
}
__device__ inline v3 sample_sun_direction(const PtScene& s, float u0, float u1) {
  v3 t, b;
  const v3 sd = mk(s.sun_dir);
  basis(sd, t, b);
  return to_world(uniform_cone(u0, u1, s.sun_cos), t, b, sd);
}
__device__ inline LightSample sample_light(const PtScene& s, int index, v3 p, float u0, float u1) {
  LightSample ls;
  ls.pdf = 0.0f;
  ls.dir = mk(0.f, 0.f, 0.f);
  ls.dist = 0.0f;
  const float* L = s.lights + index * 16;
  ls.radiance = mk(L + 12);
  if (L[3] < 0.5f) {
    const v3 to_c = mk(L) - p;
    const float d2 = dot(to_c, to_c);
    const float r = L[7];
    if (d2 <= r * r) return ls;
    const float cos_max = sqrtf(fmaxf(0.0f, 1.0f - r * r / d2));
    const v3 axis = to_c * rsqrtf(d2);
    v3 t, b;
    basis(axis, t, b);
    ls.dir = to_world(uniform_cone(u0, u1, cos_max), t, b, axis);
    ls.pdf = 1.0f / (TWO_PI * (1.0f - cos_max));
    const float t0 = hit_sphere(p, ls.dir, make_float4(L[0], L[1], L[2], r), INF);
    ls.dist = t0 < INF ? t0 : sqrtf(d2);
  } else {
    const v3 e1 = mk(L + 4), e2 = mk(L + 8);
    const v3 q = mk(L) + e1 * u0 + e2 * u1;
    const v3 to_q = q - p;
    const float d2 = dot(to_q, to_q);
    const float d = sqrtf(d2);
    ls.dir = to_q / d;
    ls.dist = d;
    const v3 nc = cross(e1, e2);
    const float area = length(nc);
    const float cos_l = fabsf(dot(nc / area, ls.dir));
    if (cos_l < 1e-4f) return ls;
    ls.pdf = d2 / (cos_l * area);
  }
  return ls;
}
__device__ inline float light_pdf(const PtScene& s, int index, v3 p, v3 dir, float dist) {
  const float* L = s.lights + index * 16;
  if (L[3] < 0.5f) {
    const v3 to_c = mk(L) - p;
    const float d2 = dot(to_c, to_c);
    const float r = L[7];
    if (d2 <= r * r) return 0.0f;
    const float cos_max = sqrtf(fmaxf(0.0f, 1.0f - r * r / d2));
    return 1.0f / (TWO_PI * (1.0f - cos_max));
  }
  const v3 nc = cross(mk(L + 4), mk(L + 8));
  const float area = length(nc);
  const float cos_l = fabsf(dot(nc / area, dir));
  if (cos_l < 1e-4f) return 0.0f;
  return dist * dist / (cos_l * area);
}

// ---------------------------------------------------------------------------
// bsdf.glsl
// ---------------------------------------------------------------------------
struct Surface {
  v3 albedo;
  float metallic, roughness, ior, transmission;
  v3 emission;
  float clearcoat, clearcoat_roughness;
  v3 sheen_color;
  float sheen_roughness;
  bool thin_walled;
  v3 absorption;
  v3 n, nc;
};

__device__ inline float d_ggx(float ndh, float a) {
  const float a2 = a * a;
  const float d = ndh * ndh * (a2 - 1.0f) + 1.0f;
  return a2 / (PI * d * d);
}
__device__ inline float g1_ggx(float ndx, float a) {
  const float a2 = a * a;
  return 2.0f * ndx / (ndx + sqrtf(a2 + (1.0f - a2) * ndx * ndx));
}
// GLSL pow(x, 5.0): products (GLSL leaves pow's precision to the implementation).
__device__ inline float pow5(float x) {
  const float x2 = x * x;
  return x2 * x2 * x;
}
__device__ inline v3 fresnel_schlick3(float c, v3 f0) {
  const float f = pow5(1.0f - clampf(c, 0.0f, 1.0f));
  return f0 + (mk(1.f, 1.f, 1.f) - f0) * f;
}
__device__ inline float fresnel_schlick1(float c, float f0) { return f0 + (1.0f - f0) * pow5(1.0f - clampf(c, 0.0f, 1.0f)); }
__device__ inline float dielectric_f0(float ior) {
  const float r = (ior - 1.0f) / (ior + 1.0f);
  return r * r;
}
__device__ inline v3 sample_ggx_vndf(v3 ve, float a, float u0, float u1) {
  const v3 vh = normalize(mk(a * ve.x, a * ve.y, ve.z));
  const float lensq = vh.x * vh.x + vh.y * vh.y;
  const v3 t1v = lensq > 0.0f ? mk(-vh.y, vh.x, 0.0f) * rsqrtf(lensq) : mk(1.0f, 0.0f, 0.0f);
  const v3 t2v = cross(vh, t1v);
  const float r = sqrtf(u0);
  const float phi = TWO_PI * u1;
  const float t1 = r * cosf(phi);
  float t2 = r * sinf(phi);
  const float sv = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - sv) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + sv * t2;
  const v3 nh = t1v * t1 + t2v * t2 + vh * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
  return normalize(mk(a * nh.x, a * nh.y, fmaxf(0.0f, nh.z)));
}
__device__ inline float alpha_from_roughness(float r) { return fmaxf(r * r, 1e-3f); }
__device__ inline float d_charlie(float alpha, float ndh) {
  const float inv = 1.0f / alpha;
  const float sin2 = fmaxf(1.0f - ndh * ndh, 0.0f);
  return (2.0f + inv) * gpow(sin2, inv * 0.5f) / TWO_PI;
}
__device__ inline float v_neubelt(float ndl, float ndv) { return 1.0f / (4.0f * (ndl + ndv - ndl * ndv)); }
// Bilinear R16F LUT, CLAMP_TO_EDGE (renderer.ts:241-242, bsdf.glsl:89-92).
__device__ inline float sheen_albedo(const PtScene& s, float ndv, float r) {
  const float fx = ndv * float(kSheenLutSize - 1), fy = r * float(kSheenLutSize - 1);
  const float x = ((fx + 0.5f) / float(kSheenLutSize)) * float(kSheenLutSize) - 0.5f;
  const float y = ((fy + 0.5f) / float(kSheenLutSize)) * float(kSheenLutSize) - 0.5f;
  const float flx = floorf(x), fly = floorf(y);
  const float tx = x - flx, ty = y - fly;
  const int x0 = min(max(int(flx), 0), kSheenLutSize - 1), x1 = min(max(int(flx) + 1, 0), kSheenLutSize - 1);
  const int y0 = min(max(int(fly), 0), kSheenLutSize - 1), y1 = min(max(int(fly) + 1, 0), kSheenLutSize - 1);
  const Global<float> L{s.sheen_lut};
  const float top = L[y0 * kSheenLutSize + x0] * (1.0f - tx) + L[y0 * kSheenLutSize + x1] * tx;
  const float bot = L[y1 * kSheenLutSize + x0] * (1.0f - tx) + L[y1 * kSheenLutSize + x1] * tx;
  return top * (1.0f - ty) + bot * ty;
}
__device__ inline float specular_probability(v3 f0, float metallic, float ndv) {
  const float f = luminance(fresnel_schlick3(ndv, f0));
  return mixf(clampf(f, 0.1f, 0.9f), 1.0f, metallic);
}
__device__ inline float coat_probability(const Surface& s, v3 wo) {
  if (s.clearcoat <= 0.0f) return 0.0f;
  const float fc = s.clearcoat * fresnel_schlick1(fmaxf(dot(s.nc, wo), 0.0f), 0.04f);
  const float base = (1.0f - fc) * fmaxf(luminance(s.albedo), 0.04f);
  return clampf(fc / (fc + base), 0.15f, 0.85f);
}
__device__ inline v3 base_f0(const Surface& s) {
  const float d = dielectric_f0(s.ior);
  return mixv(mk(d, d, d), s.albedo, s.metallic);
}

__device__ inline v3 eval_opaque(const PtScene& sc, const Surface& s, v3 wo, v3 wi, float& pdf) {
  const v3 n = s.n;
  const float ndl = dot(n, wi), ndv = dot(n, wo);
  pdf = 0.0f;
  if (ndl <= 0.0f || ndv <= 0.0f) return mk(0.f, 0.f, 0.f);
  const float a = alpha_from_roughness(s.roughness);
  const v3 H = normalize(wo + wi);
  const float ndh = fmaxf(dot(n, H), 0.0f);
  const float vdh = fmaxf(dot(wo, H), 0.0f);
  const v3 f0 = base_f0(s);
  const v3 F = fresnel_schlick3(vdh, f0);
  const float D = d_ggx(ndh, a);
  const float g1v = g1_ggx(ndv, a);
  const float G = g1v * g1_ggx(ndl, a);
  const v3 spec = F * (D * G) / fmaxf(4.0f * ndv * ndl, 1e-6f);
  const v3 diff = (mk(1.f, 1.f, 1.f) - F) * s.albedo * ((1.0f - s.metallic) * INV_PI);
  v3 f = diff + spec;
  const float p_spec = specular_probability(f0, s.metallic, ndv);
  pdf = mixf(ndl * INV_PI, g1v * D / fmaxf(4.0f * ndv, 1e-6f), p_spec);
  const float sheen_max = max3(s.sheen_color);
  if (sheen_max > 0.0f) {
    const float sr = fmaxf(s.sheen_roughness, MIN_SHEEN_ROUGHNESS);
    const v3 sheen = s.sheen_color * (d_charlie(sr * sr, ndh) * v_neubelt(ndl, ndv));
    const float scale = 1.0f - sheen_max * fmaxf(sheen_albedo(sc, ndv, sr), sheen_albedo(sc, ndl, sr));
    f = sheen + f * scale;
  }
  const float p_coat = coat_probability(s, wo);
  if (p_coat > 0.0f) {
    const float cv = dot(s.nc, wo), cl = dot(s.nc, wi);
    float coat = 0.0f, pdf_coat = 0.0f;
    if (cv > 0.0f && cl > 0.0f) {
      const float ac = alpha_from_roughness(s.clearcoat_roughness);
      const float ch = fmaxf(dot(s.nc, H), 0.0f);
      const float dc = d_ggx(ch, ac);
      const float g1c = g1_ggx(cv, ac);
      coat = dc * g1c * g1_ggx(cl, ac) * fresnel_schlick1(vdh, 0.04f) / (4.0f * cv * cl);
      pdf_coat = g1c * dc / (4.0f * cv);
    }
    const float fcv = fresnel_schlick1(fmaxf(cv, 0.0f), 0.04f);
    f = f * (1.0f - s.clearcoat * fcv) + mk(coat, coat, coat) * s.clearcoat;
    pdf = mixf(pdf, pdf_coat, p_coat);
  }
  return f;
}

struct BsdfSample {
  v3 dir, weight;
  float pdf;
};

__device__ inline BsdfSample sample_opaque(const PtScene& sc, const Surface& s, v3 wo, Rng& rng) {
  BsdfSample bs;
  bs.pdf = 0.0f;
  bs.weight = mk(0.f, 0.f, 0.f);
  bs.dir = mk(0.f, 0.f, 0.f);
  v3 wi;
  if (rng.rand() < coat_probability(s, wo)) {
    v3 t, b;
    basis(s.nc, t, b);
    const v3 wol = to_local(wo, t, b, s.nc);
    if (wol.z <= 0.0f) return bs;
    const float u0 = rng.rand(), u1 = rng.rand();
    const v3 H = sample_ggx_vndf(wol, alpha_from_roughness(s.clearcoat_roughness), u0, u1);
    wi = to_world(reflect(-wol, H), t, b, s.nc);
  } else {
    v3 t, b;
    basis(s.n, t, b);
    const v3 wol = to_local(wo, t, b, s.n);
    if (wol.z <= 0.0f) return bs;
    v3 wil;
    if (rng.rand() < specular_probability(base_f0(s), s.metallic, wol.z)) {
      const float u0 = rng.rand(), u1 = rng.rand();
      const v3 H = sample_ggx_vndf(wol, alpha_from_roughness(s.roughness), u0, u1);
      wil = reflect(-wol, H);
    } else {
      const float u0 = rng.rand(), u1 = rng.rand();
      wil = cosine_hemisphere(u0, u1);
    }
    wi = to_world(wil, t, b, s.n);
  }
  const float ndl = dot(s.n, wi);
  if (ndl <= 0.0f) return bs;
  float pdf;
  const v3 f = eval_opaque(sc, s, wo, wi, pdf);
  if (pdf <= 0.0f) return bs;
  bs.dir = wi;
  bs.pdf = pdf;
  bs.weight = f * (ndl / pdf);
  return bs;
}

__device__ inline bool sample_dielectric(const Surface& s, bool front_face, v3 ng, v3 wo, Rng& rng, v3& wi, v3& weight,
                                         bool& reflected) {
  const v3 n = s.n;
  v3 t, b;
  basis(n, t, b);
  const v3 wol = to_local(wo, t, b, n);
  if (wol.z <= 0.0f) return false;
  v3 H = n;
  if (s.roughness > 0.0f) {
    const float u0 = rng.rand(), u1 = rng.rand();
    const v3 hl = sample_ggx_vndf(wol, alpha_from_roughness(s.roughness), u0, u1);
    H = to_world(hl, t, b, n);
  }
  float cos_i = dot(wo, H);
  if (cos_i <= 0.0f) {
    H = n;
    cos_i = wol.z;
  }
  const float eta = front_face ? 1.0f / s.ior : s.ior;
  const float sin2t = eta * eta * (1.0f - cos_i * cos_i);
  float fr = fresnel_schlick1(cos_i, dielectric_f0(s.ior));
  if (s.thin_walled)
    fr = 2.0f * fr / (1.0f + fr);
  else if (sin2t > 1.0f)
    fr = 1.0f;
  if (rng.rand() < fr) {
    wi = reflect(-wo, H);
    weight = mk(1.f, 1.f, 1.f);
    reflected = true;
    return dot(wi, ng) > 0.0f;
  }
  wi = s.thin_walled ? -wo : refract(-wo, H, eta);
  weight = s.albedo;
  reflected = false;
  return dot(wi, ng) < 0.0f;
}

// ---------------------------------------------------------------------------
// trace.frag
// ---------------------------------------------------------------------------
__device__ inline v3 clamp_contribution(v3 c, int bounce, float clamp) {
  if (bounce == 0 || clamp <= 0.0f) return c;
  const float l = luminance(c);
  return l > clamp ? c * (clamp / l) : c;
}

// surfaceAt (trace.frag:72-124).
__device__ inline Surface surface_at(const PtScene& sc, const Hit& hit, v3 wo, v3& emission) {
  const Material m = fetch_material(sc, hit.material);
  Surface s;
  float tu, tv;
  transform_uv(hit.uv_u, hit.uv_v, m.uv0, m.uv1, tu, tv);
  s.albedo = m.albedo;
  if (m.checker_scale > 0.0f) {
    const v3 q = mk(floorf(hit.position.x * m.checker_scale + 0.5f), floorf(hit.position.y * m.checker_scale + 0.5f),
                    floorf(hit.position.z * m.checker_scale + 0.5f));
    const float sum = q.x + q.y + q.z;
    if (sum - 2.0f * floorf(sum / 2.0f) >= 0.5f) s.albedo = m.albedo2;
  }
  const bool tex = sc.tex_size > 0;
  if (m.map_base >= 0 && m.map_base < sc.color_layers && tex) s.albedo = s.albedo * mk(sample_layer(sc.color_maps, sc.tex_size, m.map_base, tu, tv, true));
  s.metallic = m.metallic;
  s.roughness = m.roughness;
  if (m.map_mr >= 0 && m.map_mr < sc.data_layers && tex) {
    const float4 mr = sample_layer(sc.data_maps, sc.tex_size, m.map_mr, tu, tv, false);
    s.roughness *= mr.y;
    s.metallic *= mr.z;
  }
  emission = m.emission;
  if (m.map_emissive >= 0 && m.map_emissive < sc.color_layers && tex) emission = emission * mk(sample_layer(sc.color_maps, sc.tex_size, m.map_emissive, tu, tv, true));
  s.nc = hit.normal;
  s.n = hit.normal;
  if (m.map_normal >= 0 && m.map_normal < sc.data_layers && tex && dot(hit.dpdu, hit.dpdu) > 0.0f) {
    const v3 N = hit.front_face ? hit.normal : -hit.normal;
    const v3 T = normalize(hit.dpdu - N * dot(N, hit.dpdu));
    v3 B = cross(N, T);
    if (dot(B, -hit.dpdv) < 0.0f) B = -B;
    const float4 tm = sample_layer(sc.data_maps, sc.tex_size, m.map_normal, tu, tv, false);
    v3 tn = mk(tm.x * 2.0f - 1.0f, tm.y * 2.0f - 1.0f, tm.z * 2.0f - 1.0f);
    tn.x *= m.normal_scale;
    tn.y *= m.normal_scale;
    v3 mapped = normalize(T * tn.x + B * tn.y + N * fmaxf(tn.z, 1e-3f));
    mapped = hit.front_face ? mapped : -mapped;
    if (dot(mapped, wo) > 1e-3f && dot(mapped, hit.geom_normal) > 0.0f) s.n = mapped;
  }
  s.ior = m.ior;
  s.transmission = m.transmission * (1.0f - s.metallic);
  s.emission = emission;
  s.clearcoat = m.clearcoat;
  s.clearcoat_roughness = m.clearcoat_roughness;
  s.sheen_color = m.sheen_color;
  s.sheen_roughness = m.sheen_roughness;
  s.thin_walled = m.thin_walled;
  if (m.atten_distance > 0.0f) {
    const v3 c = vmax(m.atten_color, 1e-4f);
    s.absorption = mk(-logf(c.x), -logf(c.y), -logf(c.z)) / m.atten_distance;
  } else {
    s.absorption = mk(0.f, 0.f, 0.f);
  }
  return s;
}

struct Guide {
  v3 albedo, normal;
  float depth, coverage;
  bool done;
  // Set only by tracers with kRecordPrimary (the reprojecting sample):
  v3 position;  // primary hit
  float reuse;  // 0-1: how view-independent the primary surface is (trace.frag:187-193)
};

// cameraRay (trace.frag:53-62); frag = gl_FragCoord (pixel centre, rows bottom-up).
__device__ inline void camera_ray(const PtFrame& f, float fx, float fy, Rng& rng, v3& ro, v3& rd) {
  const float r0 = rng.rand(), r1 = rng.rand();
  const float ux = (fx + r0) / float(f.width) * 2.0f - 1.0f;
  const float uy = (fy + r1) / float(f.height) * 2.0f - 1.0f;
  const float aspect = float(f.width) / float(f.height);
  const v3 F = mk(f.cam_fwd), R = mk(f.cam_right), U = mk(f.cam_up), P = mk(f.cam_pos);
  const v3 dir = normalize(F + R * (ux * aspect * f.tan_half_fov) + U * (uy * f.tan_half_fov));
  const v3 focus = P + dir * (f.focus_dist / dot(dir, F));
  const float l0 = rng.rand(), l1 = rng.rand();
  float lx, ly;
  concentric_disk(l0, l1, lx, ly);
  lx *= f.aperture;
  ly *= f.aperture;
  ro = P + R * lx + U * ly;
  rd = normalize(focus - ro);
}

// The state of one camera path between bounces of tracePath
// (trace.frag:136-272).  trace_path() runs a path to the end; the persistent
// tile kernel advances each lane's path one bounce at a time and starts the
// lane's next sample as soon as a path ends.
struct PathState {
  v3 ro, rd;
  v3 radiance, throughput, tint;
  float bsdf_pdf;
  int bounce;
  bool specular;
  Guide guide;
};

__device__ __forceinline__ void path_begin(PathState& p, v3 ro, v3 rd) {
  p.ro = ro;
  p.rd = rd;
  p.radiance = mk(0.f, 0.f, 0.f);
  p.throughput = mk(1.f, 1.f, 1.f);
  p.tint = mk(1.f, 1.f, 1.f);
  p.bsdf_pdf = 0.0f;
  p.bounce = 0;
  p.specular = true;
  p.guide.albedo = mk(1.f, 1.f, 1.f);
  p.guide.normal = -rd;
  p.guide.depth = FAR;
  p.guide.coverage = 0.0f;
  p.guide.done = false;
  p.guide.position = mk(0.f, 0.f, 0.f);
  p.guide.reuse = 1.0f;
}

// One iteration of tracePath's bounce loop; false once the path has ended
// (p.radiance and p.guide are then final).
// Rays of a path step are cast through a tracer policy:
//   bool closest(v3 ro, v3 rd, Hit& hit, Rng& rng)  - traceScene;
//   void nee(int which, v3 o, v3 d, float tmax, Rng& rng, PathState& p, F contribution)
//     - a shadow ray for light sample `which` (0 sun, 1 environment, 2 light)
//       whose unoccluded contribution(), added to p.radiance, is the term.
// InlineTracer traces both on the spot (the megakernels); the wavefront's
// tracer reads the closest hit from its extend pass and defers the shadow
// rays, adding their contributions in the same order afterwards.
template <class StackT>
struct InlineTracer {
  static constexpr bool kRecordPrimary = false;  // fill Guide::position and Guide::reuse
  const PtScene& sc;
  StackT& st;
  __device__ __forceinline__ bool closest(v3 ro, v3 rd, Hit& hit, Rng& rng) { return trace_scene(sc, ro, rd, hit, rng, st); }
  template <class F>
  __device__ __forceinline__ void nee(int, v3 o, v3 d, float tmax, Rng& rng, PathState& p, F contribution) {
    if (!occluded(sc, o, d, tmax, rng, st)) p.radiance = p.radiance + contribution();
  }
};

// InlineTracer that also records the primary hit position and the surface's
// reuse factor into the guide (for the reprojecting sample).
template <class StackT>
struct PrimaryTracer : InlineTracer<StackT> {
  static constexpr bool kRecordPrimary = true;
};

template <class Tracer>
__device__ __forceinline__ bool path_step_with(const PtScene& sc, const PtFrame& f, PathState& p, Rng& rng, Tracer& tr) {
  const int bounce = p.bounce;
  Hit hit;
  if (!tr.closest(p.ro, p.rd, hit, rng)) {
    const v3 env = (bounce == 0 && sc.env_background == 0) ? mk(sc.backdrop) : environment(sc, p.rd, p.specular, p.bsdf_pdf);
    p.radiance = p.radiance + clamp_contribution(p.throughput * env, bounce, f.clamp);
    if (!p.guide.done) {
      p.guide.albedo = p.tint;
      p.guide.normal = -p.rd;
      p.guide.done = true;
    }
    return false;
  }
  if (bounce == 0) {
    p.guide.depth = hit.t;
    p.guide.coverage = 1.0f;
    if constexpr (Tracer::kRecordPrimary) p.guide.position = hit.position;
  }
  const v3 wo = -p.rd;
  v3 emission;
  const Surface s = surface_at(sc, hit, wo, emission);
  const v3 ng = hit.geom_normal;
  if (!hit.front_face && !s.thin_walled && s.transmission > 0.0f) p.throughput = p.throughput * vexp(-s.absorption * hit.t);
  if (emission.x > 0.0f || emission.y > 0.0f || emission.z > 0.0f) {
    float w = 1.0f;
    if (!p.specular && hit.light >= 0) {
      const float pl = light_pdf(sc, hit.light, p.ro, p.rd, hit.t) / float(sc.light_count);
      w = mis_weight(p.bsdf_pdf, pl);
    }
    p.radiance = p.radiance + clamp_contribution(p.throughput * emission * w, bounce, f.clamp);
    if (!p.guide.done) {
      p.guide.albedo = p.tint;
      p.guide.normal = s.n;
      p.guide.done = true;
    }
    if (hit.light >= 0) return false;
  }
  const bool glass = s.transmission > 0.0f && rng.rand() < s.transmission;
  if constexpr (Tracer::kRecordPrimary) {
    if (bounce == 0) {
      // Reflections and refractions do not move with the surface (trace.frag:187-193).
      const float sharp = 1.0f - smoothstep(0.1f, 0.7f, s.roughness);
      float glossy = fmaxf(sharp * (0.5f * (1.0f - s.metallic) + 1.0f * s.metallic), s.transmission);
      glossy = fmaxf(glossy, s.clearcoat * 0.5f * (1.0f - smoothstep(0.08f, 0.5f, s.clearcoat_roughness)));
      p.guide.reuse = 1.0f - glossy;
    }
  }
  const bool mirror = !glass && s.metallic > 0.9f && s.roughness < 0.1f;
  if (!p.guide.done && !((glass || mirror) && bounce < 3)) {
    p.guide.albedo = p.tint * s.albedo;
    p.guide.normal = s.n;
    p.guide.done = true;
  }
  if (glass) {
    v3 wi, weight;
    bool reflected;
    if (!sample_dielectric(s, hit.front_face, ng, wo, rng, wi, weight, reflected)) return false;
    p.throughput = p.throughput * weight;
    if (!p.guide.done) p.tint = p.tint * weight;
    p.ro = offset_ray(hit.position, reflected ? ng : -ng);
    p.rd = wi;
    p.specular = true;
  } else {
    if (!p.guide.done) p.tint = p.tint * s.albedo;
    const v3 shadow_origin = offset_ray(hit.position, ng);
    const v3 n = s.n;
    if (sc.sun_enabled != 0) {
      const float u0 = rng.rand(), u1 = rng.rand();
      const v3 wi = sample_sun_direction(sc, u0, u1);
      const float ndl = dot(n, wi);
      if (ndl > 0.0f && dot(ng, wi) > 0.0f)
        tr.nee(0, shadow_origin, wi, INF, rng, p, [&]() {
          float pb;
          const v3 fv = eval_opaque(sc, s, wo, wi, pb);
          const float w = mis_weight(sc.sun_pdf, pb);
          return clamp_contribution(p.throughput * fv * ndl * mk(sc.sun_radiance) * (w / sc.sun_pdf), bounce, f.clamp);
        });
    }
    if (sc.env_sampling != 0) {
      const LightSample ls = sample_environment(sc, rng);
      const float ndl = dot(n, ls.dir);
      if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f)
        tr.nee(1, shadow_origin, ls.dir, INF, rng, p, [&]() {
          float pb;
          const v3 fv = eval_opaque(sc, s, wo, ls.dir, pb);
          const float w = mis_weight(ls.pdf, pb);
          return clamp_contribution(p.throughput * fv * ndl * ls.radiance * (w / ls.pdf), bounce, f.clamp);
        });
    }
    if (sc.light_count > 0) {
      const int li = min(int(rng.rand() * float(sc.light_count)), int(sc.light_count) - 1);
      const float u0 = rng.rand(), u1 = rng.rand();
      const LightSample ls = sample_light(sc, li, hit.position, u0, u1);
      const float ndl = dot(n, ls.dir);
      if (ls.pdf > 0.0f && ndl > 0.0f && dot(ng, ls.dir) > 0.0f)
        tr.nee(2, shadow_origin, ls.dir, ls.dist * (1.0f - 1e-4f), rng, p, [&]() {
          const float pl = ls.pdf / float(sc.light_count);
          float pb;
          const v3 fv = eval_opaque(sc, s, wo, ls.dir, pb);
          const float w = mis_weight(pl, pb);
          return clamp_contribution(p.throughput * fv * ndl * ls.radiance * (w / pl), bounce, f.clamp);
        });
    }
    const BsdfSample bs = sample_opaque(sc, s, wo, rng);
    if (bs.pdf <= 0.0f || dot(bs.dir, ng) <= 0.0f) return false;
    p.throughput = p.throughput * bs.weight;
    p.bsdf_pdf = bs.pdf;
    p.specular = false;
    p.ro = shadow_origin;
    p.rd = bs.dir;
  }
  if (bounce >= 2) {
    const float q = fminf(0.95f, max3(p.throughput));
    if (rng.rand() >= q) return false;
    p.throughput = p.throughput / q;
  }
  p.bounce = bounce + 1;
  return p.bounce <= int(f.max_bounces);
}

template <class StackT>
__device__ __forceinline__ bool path_step(const PtScene& sc, const PtFrame& f, PathState& p, Rng& rng, StackT& st) {
  InlineTracer<StackT> tr{sc, st};
  return path_step_with(sc, f, p, rng, tr);
}

template <class StackT>
__device__ inline v3 trace_path(const PtScene& sc, const PtFrame& f, v3 ro, v3 rd, Rng& rng, StackT& st, Guide& guide) {
  PathState p;
  path_begin(p, ro, rd);
  if (int(f.max_bounces) >= 0)
    while (path_step(sc, f, p, rng, st)) {
    }
  guide = p.guide;
  return p.radiance;
}

// A pixel's accumulation targets (renderer.ts:490 formats) as float values;
// the FP16 guides hold exactly representable half values.
struct PixelAccum {
  float4 color, albedo, normal;
  float n;
};

// The film's targets, read and written as global memory (as the scene's
// arrays): four halves travel as two 32-bit words, low half first.
typedef uint32_t gu32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) gu32x2 global_u32x2;
typedef __attribute__((address_space(1))) float global_f32_rw;

__device__ inline PixelAccum load_accum(const PtFilm& film, uint32_t idx) {
  auto h2f = [](uint32_t h) { return __half2float(__ushort_as_half((unsigned short)h)); };
  PixelAccum a;
  a.color = ld_global(reinterpret_cast<const float4*>(film.color) + idx);
  const gu32x2 pa = *(const global_u32x2*)(reinterpret_cast<const uint2*>(film.albedo) + idx);
  const gu32x2 pn = *(const global_u32x2*)(reinterpret_cast<const uint2*>(film.normal) + idx);
  a.albedo = make_float4(h2f(pa.x & 0xFFFFu), h2f(pa.x >> 16), h2f(pa.y & 0xFFFFu), h2f(pa.y >> 16));
  a.normal = make_float4(h2f(pn.x & 0xFFFFu), h2f(pn.x >> 16), h2f(pn.y & 0xFFFFu), h2f(pn.y >> 16));
  a.n = ld_global(film.history + idx);
  return a;
}

__device__ inline float half_round(float x) { return __half2float(__float2half_rn(x)); }

// One sample into the running means (trace.frag:330-354): NaN/Inf -> 0,
// k = 1/(n+1), GLSL mix = x(1-k) + yk, guides rounded to FP16 as the RGBA16F
// targets store them.  Contraction is off so that every kernel that folds
// samples (one per launch, or many in registers) rounds identically.
__device__ inline float mix_nc(float x, float y, float j, float k) {
#pragma clang fp contract(off)
  return x * j + y * k;
}

__device__ inline void fold_sample(PixelAccum& a, v3 color, const Guide& g) {
#pragma clang fp contract(off)
  if (isnan(color.x) || isnan(color.y) || isnan(color.z) || isinf(color.x) || isinf(color.y) || isinf(color.z))
    color = mk(0.f, 0.f, 0.f);
  const float lum = 0.2126f * color.x + 0.7152f * color.y + 0.0722f * color.z;
  const float k = 1.0f / (a.n + 1.0f);
  const float j = 1.0f - k;
  a.color = make_float4(mix_nc(a.color.x, color.x, j, k), mix_nc(a.color.y, color.y, j, k),
                        mix_nc(a.color.z, color.z, j, k), mix_nc(a.color.w, lum * lum, j, k));
  a.albedo = make_float4(half_round(mix_nc(a.albedo.x, g.albedo.x, j, k)), half_round(mix_nc(a.albedo.y, g.albedo.y, j, k)),
                         half_round(mix_nc(a.albedo.z, g.albedo.z, j, k)), half_round(mix_nc(a.albedo.w, g.coverage, j, k)));
  a.normal = make_float4(half_round(mix_nc(a.normal.x, g.normal.x, j, k)), half_round(mix_nc(a.normal.y, g.normal.y, j, k)),
                         half_round(mix_nc(a.normal.z, g.normal.z, j, k)), half_round(mix_nc(a.normal.w, g.depth, j, k)));
  a.n = a.n + 1.0f;
}

__device__ inline void store_accum(const PtFilm& film, uint32_t idx, const PixelAccum& a) {
  auto f2h = [](float x) { return uint32_t(__half_as_ushort(__float2half_rn(x))); };
  st_global(reinterpret_cast<float4*>(film.color) + idx, a.color);
  gu32x2 pa, pn;
  pa.x = f2h(a.albedo.x) | (f2h(a.albedo.y) << 16);
  pa.y = f2h(a.albedo.z) | (f2h(a.albedo.w) << 16);
  pn.x = f2h(a.normal.x) | (f2h(a.normal.y) << 16);
  pn.y = f2h(a.normal.z) | (f2h(a.normal.w) << 16);
  *(global_u32x2*)(reinterpret_cast<uint2*>(film.albedo) + idx) = pa;
  *(global_u32x2*)(reinterpret_cast<uint2*>(film.normal) + idx) = pn;
  *(global_f32_rw*)(film.history + idx) = a.n;
}

__device__ inline void accumulate(const PtFilm& film, uint32_t idx, v3 color, const Guide& g) {
  PixelAccum a = load_accum(film, idx);
  fold_sample(a, color, g);
  store_accum(film, idx, a);
}

}  // namespace pt
}  // namespace lucent
