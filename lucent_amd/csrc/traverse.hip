// Batched BVH ray queries on gfx950.
//
// One lane per ray, grid-stride over the batch.  The traversal walks the
// 64-byte child-pair records (internal.hpp): one record fetch tests both
// children, leaves are tested as soon as their box is hit, and the nearer of
// two interior children is descended first while the farther goes on a
// per-lane stack.  The stack's top kLdsStack entries live in LDS (lane-major
// columns, so the 64 lanes of a wave hit 64 distinct banks); deeper entries
// spill to a per-lane global area sized from the tree depth, so no entry is
// ever dropped.
//
// Exact mode reproduces intersectBVH (reference src/math/bvh.ts:265-320):
//  * box tests run in float32 but are made conservative — the far distance
//    and the current closest hit are inflated by 2^-20 relative, far more
//    than the <= 6 ulp error of the float32 slab, so every box the FP64 CPU
//    test accepts is accepted here (extra boxes only cost time);
//  * triangle tests run in FP64 with the exact operation order of
//    src/math/ray.ts:22-45 (no contraction), so t, u and v are bit-identical;
//  * the CPU's left-first walk with strict `t < closest` keeps the lowest
//    leaf slot among equal t; here the lexicographic minimum of (t, slot) is
//    kept, which is the same result in any visiting order.
// GLSL mode follows the shader's float32 hitTriangle/hitAABB
// (src/shaders/scene.glsl:221-269) with EPS = 1e-6 as the lower bound.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "internal.hpp"

namespace lucent {
namespace {

constexpr int kBlock = 128;
constexpr int kLdsStack = 16;
constexpr float kMargin32 = 1.0f + 1.0f / 1048576.0f;       // 1 + 2^-20
constexpr double kMargin64 = 1.0 + 1.0 / 1099511627776.0;  // 1 + 2^-40
constexpr float kGlslEps = 1e-6f;

enum class Prec { kExact, kGlsl };

// kLds entries of the stack stay in LDS as a ring; older entries spill.
template <int kLds>
struct LaneStack {
  uint32_t* lds;    // this lane's column: entry k at lds[k * kBlock]
  uint32_t* spill;  // this lane's global area
  uint32_t sp = 0, base = 0;

  __device__ void push(uint32_t x) {
    if (sp - base == kLds) {
      spill[base] = lds[(base % kLds) * kBlock];
      ++base;
    }
    lds[(sp % kLds) * kBlock] = x;
    ++sp;
  }
  __device__ uint32_t pop() {
    --sp;
    if (sp >= base) return lds[(sp % kLds) * kBlock];
    --base;
    return spill[sp];
  }
  __device__ bool empty() const { return sp == 0; }
};

struct PairRecord {
  float4 x, y, z;  // (lo0, lo1, hi0, hi1) per axis
  uint4 link;      // (link0, count0, link1, count1)
};

__device__ inline PairRecord load_pair(const float4* pairs, uint32_t p) {
  const float4* r = pairs + size_t(p) * 4;
  PairRecord rec;
  rec.x = r[0];
  rec.y = r[1];
  rec.z = r[2];
  const float4 l = r[3];
  rec.link = make_uint4(__float_as_uint(l.x), __float_as_uint(l.y), __float_as_uint(l.z), __float_as_uint(l.w));
  return rec;
}

// Slab test of both children in float32.  min/max follow IEEE minNum/maxNum,
// so a NaN slab (zero direction component with the origin on the plane)
// leaves the interval unconstrained, as the reference's comparisons do.
template <bool kConservative>
__device__ inline void slab_pair(const PairRecord& r, const float o[3], const float inv[3], float limit, float& tn0,
                                 float& tn1, bool& h0, bool& h1) {
  const float x0 = (r.x.x - o[0]) * inv[0], x1 = (r.x.z - o[0]) * inv[0];
  const float y0 = (r.y.x - o[1]) * inv[1], y1 = (r.y.z - o[1]) * inv[1];
  const float z0 = (r.z.x - o[2]) * inv[2], z1 = (r.z.z - o[2]) * inv[2];
  const float X0 = (r.x.y - o[0]) * inv[0], X1 = (r.x.w - o[0]) * inv[0];
  const float Y0 = (r.y.y - o[1]) * inv[1], Y1 = (r.y.w - o[1]) * inv[1];
  const float Z0 = (r.z.y - o[2]) * inv[2], Z1 = (r.z.w - o[2]) * inv[2];
  tn0 = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), 0.0f));
  tn1 = fmaxf(fmaxf(fminf(X0, X1), fminf(Y0, Y1)), fmaxf(fminf(Z0, Z1), 0.0f));
  float tf0 = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
  float tf1 = fminf(fminf(fmaxf(X0, X1), fmaxf(Y0, Y1)), fmaxf(Z0, Z1));
  if (kConservative) {
    tf0 = fminf(tf0, limit) * kMargin32;
    tf1 = fminf(tf1, limit) * kMargin32;
  } else {
    tf0 = fminf(tf0, limit);
    tf1 = fminf(tf1, limit);
  }
  h0 = tf0 >= tn0;
  h1 = tf1 >= tn1;
}

// Double-precision slab of both children (for double-precision rays, whose
// origin cannot be rounded to float32 without an absolute error).
__device__ inline void slab_pair64(const PairRecord& r, const double o[3], const double inv[3], double limit,
                                   double& tn0, double& tn1, bool& h0, bool& h1) {
#pragma clang fp contract(off)
  const float lo[2][3] = {{r.x.x, r.y.x, r.z.x}, {r.x.y, r.y.y, r.z.y}};
  const float hi[2][3] = {{r.x.z, r.y.z, r.z.z}, {r.x.w, r.y.w, r.z.w}};
  double tn[2], tf[2];
  for (int c = 0; c < 2; ++c) {
    double a = 0.0, b = limit;
    for (int k = 0; k < 3; ++k) {
      double t0 = (double(lo[c][k]) - o[k]) * inv[k];
      double t1 = (double(hi[c][k]) - o[k]) * inv[k];
      if (t0 > t1) {
        const double tt = t0;
        t0 = t1;
        t1 = tt;
      }
      if (t0 > a) a = t0;
      if (t1 < b) b = t1;
    }
    tn[c] = a;
    tf[c] = b * kMargin64;
  }
  tn0 = tn[0];
  tn1 = tn[1];
  h0 = tf[0] >= tn[0];
  h1 = tf[1] >= tn[1];
}

// Möller–Trumbore in FP64, operation for operation as src/math/ray.ts:22-45
// with vec3.ts:15-25 (sub, dot left to right, cross).  Returns false where
// the reference returns null before its t bound check; t is returned raw.
__device__ inline bool mt_exact(const double ro[3], const double rd[3], const float4 a, const float4 b,
                                const float4 c, double& t, double& u, double& v) {
#pragma clang fp contract(off)
  const double v0[3] = {a.x, a.y, a.z};
  const double e1[3] = {double(b.x) - v0[0], double(b.y) - v0[1], double(b.z) - v0[2]};
  const double e2[3] = {double(c.x) - v0[0], double(c.y) - v0[1], double(c.z) - v0[2]};
  const double p[3] = {rd[1] * e2[2] - rd[2] * e2[1], rd[2] * e2[0] - rd[0] * e2[2], rd[0] * e2[1] - rd[1] * e2[0]};
  const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (fabs(det) < 1e-12) return false;
  const double s[3] = {ro[0] - v0[0], ro[1] - v0[1], ro[2] - v0[2]};
  const double un = s[0] * p[0] + s[1] * p[1] + s[2] * p[2];
  // Sign pre-check: a nonzero un whose sign differs from det's gives a
  // negative u = un * (1/det) (no underflow at scene magnitudes), which the
  // reference rejects with `u < 0`; skipping the division here changes no
  // result.  NaN fails both comparisons and takes the reference's path.
  if ((un < 0.0 && det > 0.0) || (un > 0.0 && det < 0.0)) return false;
  const double inv_det = 1.0 / det;
  u = un * inv_det;
  if (u < 0.0 || u > 1.0) return false;
  const double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
  v = (rd[0] * q[0] + rd[1] * q[1] + rd[2] * q[2]) * inv_det;
  if (v < 0.0 || u + v > 1.0) return false;
  t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv_det;
  return true;
}

// hitTriangle (scene.glsl:221-238) in float32; t returned raw.
__device__ inline bool mt_glsl(const float ro[3], const float rd[3], const float4 a, const float4 b, const float4 c,
                               float& t, float& u, float& v) {
  const float e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
  const float e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
  const float p[3] = {rd[1] * e2[2] - rd[2] * e2[1], rd[2] * e2[0] - rd[0] * e2[2], rd[0] * e2[1] - rd[1] * e2[0]};
  const float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (fabsf(det) < 1e-12f) return false;
  const float inv_det = 1.0f / det;
  const float s[3] = {ro[0] - a.x, ro[1] - a.y, ro[2] - a.z};
  u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  const float q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
  v = (rd[0] * q[0] + rd[1] * q[1] + rd[2] * q[2]) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv_det;
  return true;
}

struct Best {
  double t, u, v;
  uint32_t slot;
};

template <Prec kPrec, bool kAny>
__device__ inline bool test_leaf(const float4* tris, uint32_t first, uint32_t count, const double rod[3],
                                 const double rdd[3], const float rof[3], const float rdf[3], double tmin,
                                 double tmax, Best& best) {
  for (uint32_t s = first; s < first + count; ++s) {
    const float4 a = tris[size_t(s) * 3], b = tris[size_t(s) * 3 + 1], c = tris[size_t(s) * 3 + 2];
    if (kPrec == Prec::kExact) {
      double t, u, v;
      if (!mt_exact(rod, rdd, a, b, c, t, u, v)) continue;
      if (!(t > tmin && t < tmax)) continue;
      if (t < best.t || (t == best.t && s < best.slot)) {
        best = Best{t, u, v, s};
        if (kAny) return true;
      }
    } else {
      float t, u, v;
      if (!mt_glsl(rof, rdf, a, b, c, t, u, v)) continue;
      if (!(t > kGlslEps && double(t) < tmax)) continue;
      if (double(t) < best.t) {
        best = Best{double(t), double(u), double(v), s};
        if (kAny) return true;
      }
    }
  }
  return false;
}

template <Prec kPrec, bool kAny, bool kF64Rays, int kLds>
__global__ __launch_bounds__(kBlock) void k_intersect(DevBvh bvh, const void* __restrict__ org_,
                                                      const void* __restrict__ dir_, uint32_t n, double tmin,
                                                      double tmax, double* __restrict__ t_out,
                                                      uint32_t* __restrict__ id_out, double* __restrict__ u_out,
                                                      double* __restrict__ v_out, uint32_t* __restrict__ spill,
                                                      uint32_t spill_depth) {
  __shared__ uint32_t lds_stack[kLds * kBlock];
  const float4* pairs = reinterpret_cast<const float4*>(bvh.pairs);
  const float4* tris = reinterpret_cast<const float4*>(bvh.tris);
  const uint32_t lane_global = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t stride = gridDim.x * kBlock;

  for (uint32_t r = lane_global; r < n; r += stride) {
    double rod[3], rdd[3];
    float rof[3], rdf[3];
    if (kF64Rays) {
      const double* o = static_cast<const double*>(org_) + size_t(r) * 3;
      const double* d = static_cast<const double*>(dir_) + size_t(r) * 3;
      for (int k = 0; k < 3; ++k) {
        rod[k] = o[k];
        rdd[k] = d[k];
        rof[k] = float(o[k]);
        rdf[k] = float(d[k]);
      }
    } else {
      const float* o = static_cast<const float*>(org_) + size_t(r) * 3;
      const float* d = static_cast<const float*>(dir_) + size_t(r) * 3;
      for (int k = 0; k < 3; ++k) {
        rof[k] = o[k];
        rdf[k] = d[k];
        rod[k] = o[k];
        rdd[k] = d[k];
      }
    }
    const float invf[3] = {1.0f / rdf[0], 1.0f / rdf[1], 1.0f / rdf[2]};
    double invd[3];
    if (kF64Rays) {
#pragma clang fp contract(off)
      invd[0] = 1.0 / rdd[0];
      invd[1] = 1.0 / rdd[1];
      invd[2] = 1.0 / rdd[2];
    }
    const double lower = kPrec == Prec::kExact ? tmin : double(kGlslEps);
    Best best{tmax, 0.0, 0.0, 0xFFFFFFFFu};
    bool done = false;

    LaneStack<kLds> st{lds_stack + threadIdx.x, spill + size_t(lane_global) * spill_depth};
    bool have = false;
    uint32_t cur = 0;
    if (bvh.tri_count == 0) {
      done = true;
    } else if (bvh.root_count > 0) {
      test_leaf<kPrec, kAny>(tris, bvh.root_link, bvh.root_count, rod, rdd, rof, rdf, lower, tmax, best);
      done = true;
    } else {
      cur = bvh.root_link;
      have = true;
    }
    while (!done) {
      if (!have) {
        if (st.empty()) break;
        cur = st.pop();
      }
      have = false;
      const PairRecord rec = load_pair(pairs, cur);
      bool h0, h1;
      float tn0, tn1;
      if (kF64Rays && kPrec == Prec::kExact) {
        double a0, a1;
        slab_pair64(rec, rod, invd, best.t, a0, a1, h0, h1);
        tn0 = float(a0);
        tn1 = float(a1);
      } else if (kPrec == Prec::kExact) {
        slab_pair<true>(rec, rof, invf, __double2float_ru(best.t), tn0, tn1, h0, h1);
      } else {
        slab_pair<false>(rec, rof, invf, float(best.t), tn0, tn1, h0, h1);
      }
      const uint32_t link0 = rec.link.x, cnt0 = rec.link.y, link1 = rec.link.z, cnt1 = rec.link.w;
      if (h0 && cnt0 > 0) {
        if (test_leaf<kPrec, kAny>(tris, link0, cnt0, rod, rdd, rof, rdf, lower, tmax, best)) break;
      }
      if (h1 && cnt1 > 0) {
        if (test_leaf<kPrec, kAny>(tris, link1, cnt1, rod, rdd, rof, rdf, lower, tmax, best)) break;
      }
      const bool g0 = h0 && cnt0 == 0, g1 = h1 && cnt1 == 0;
      if (g0 && g1) {
        const bool first0 = tn0 <= tn1;
        cur = first0 ? link0 : link1;
        st.push(first0 ? link1 : link0);
        have = true;
      } else if (g0) {
        cur = link0;
        have = true;
      } else if (g1) {
        cur = link1;
        have = true;
      }
    }
    if (best.slot != 0xFFFFFFFFu) {
      t_out[r] = best.t;
      id_out[r] = bvh.tri_ids ? bvh.tri_ids[best.slot] : best.slot;
      u_out[r] = best.u;
      v_out[r] = best.v;
    } else {
      t_out[r] = INFINITY;
      id_out[r] = 0xFFFFFFFFu;
      u_out[r] = 0.0;
      v_out[r] = 0.0;
    }
  }
}

template <Prec kPrec, bool kAny, bool kF64Rays, int kLds = kLdsStack>
Status launch_variant(Context* ctx, const DevBvh& bvh, const RayQuery& q, hipStream_t stream) {
  auto kern = k_intersect<kPrec, kAny, kF64Rays, kLds>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess || per_cu < 1) per_cu = 4;
  uint64_t blocks_needed = (uint64_t(q.n) + kBlock - 1) / kBlock;
  uint64_t resident = uint64_t(per_cu) * uint64_t(ctx->compute_units > 0 ? ctx->compute_units : 256);
  uint32_t grid = uint32_t(blocks_needed < resident ? blocks_needed : resident);
  if (grid == 0) return LUCENT_OK;
  const uint32_t depth = bvh.stack_depth + 1;
  Status s = ensure_spill(ctx, size_t(grid) * kBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, bvh, q.org, q.dir, q.n, q.tmin, q.tmax, q.t, q.id,
                     q.u, q.v, static_cast<uint32_t*>(ctx->spill), depth);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_intersect launch");
  return LUCENT_OK;
}

}  // namespace

Status launch_intersect(Context* ctx, const DevBvh& bvh, const RayQuery& q, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  const bool any = (q.flags & LUCENT_QUERY_ANY) != 0;
  const bool glsl = (q.flags & LUCENT_PREC_GLSL_FP32) != 0;
  if (q.flags & LUCENT_FLAG_TINY_LDS_STACK) {
    // Test aid: a 2-entry LDS ring forces the global spill path on any tree.
    if (q.f64_rays || any || glsl) return fail(LUCENT_E_INVALID, "LUCENT_FLAG_TINY_LDS_STACK: exact closest-hit float32 rays only");
    return launch_variant<Prec::kExact, false, false, 2>(ctx, bvh, q, stream);
  }
  if (q.f64_rays) {
    if (glsl) return any ? launch_variant<Prec::kGlsl, true, true>(ctx, bvh, q, stream)
                         : launch_variant<Prec::kGlsl, false, true>(ctx, bvh, q, stream);
    return any ? launch_variant<Prec::kExact, true, true>(ctx, bvh, q, stream)
               : launch_variant<Prec::kExact, false, true>(ctx, bvh, q, stream);
  }
  if (glsl) return any ? launch_variant<Prec::kGlsl, true, false>(ctx, bvh, q, stream)
                       : launch_variant<Prec::kGlsl, false, false>(ctx, bvh, q, stream);
  return any ? launch_variant<Prec::kExact, true, false>(ctx, bvh, q, stream)
             : launch_variant<Prec::kExact, false, false>(ctx, bvh, q, stream);
}

}  // namespace lucent
