// Batched BVH ray queries on gfx950.
//
// One lane per ray, grid-stride over the batch.  The traversal walks the
// 4-wide nodes (internal.hpp: a 64-byte record per node, child boxes as 8-bit
// codes rounded outward in a per-node frame): one record fetch tests four
// children, the hit children go on a per-lane stack of (link, entry
// distance) pairs, nearest on top, and a popped leaf has all its triangles
// tested.  The stack's top kLdsStack entries live in LDS (lane-major
// columns, so the lanes of a wave hit distinct banks); deeper entries spill
// to a per-lane global area sized from the tree depth, so no entry is ever
// dropped.
//
// Exact mode reproduces intersectBVH (reference src/math/bvh.ts:265-320):
//  * box tests run in float32 but are made conservative — the decoded boxes
//    contain the reference's float32 boxes, the far distance and the current
//    closest hit are inflated by 2^-20 relative, far more than the <= 6 ulp
//    error of the float32 slab, and a popped entry is skipped only when its
//    entry distance, shrunk by the same margin, lies strictly beyond the
//    closest hit; so every box the FP64 CPU test accepts is visited here
//    (extra boxes only cost time);
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
#include <mutex>
#include <string>

#include "internal.hpp"

namespace lucent {
namespace {

constexpr int kBlock = 128;
constexpr int kLdsStack = 12;
constexpr float kMargin32 = 1.0f + 1.0f / 1048576.0f;       // 1 + 2^-20
constexpr double kMargin64 = 1.0 + 1.0 / 1099511627776.0;  // 1 + 2^-40
constexpr float kGlslEps = 1e-6f;

enum class Prec { kExact, kGlsl };

// (link, entry distance) pairs; kLds entries stay in LDS as a ring (two
// lane-major planes), older ones spill to 2 words per entry.
// LDS pointer typed as such, so that stack accesses are ds_ instructions
// (not flat loads merging the LDS and spill reads of a pop).
typedef __attribute__((address_space(3))) uint32_t lds_word;

template <int kLds>
struct LaneStack {
  lds_word* lds;    // this lane's column of plane 0: entry k at lds[k * kBlock]; plane 1 kLds * kBlock later
  uint32_t* spill;  // this lane's global area
  uint32_t sp = 0, base = 0;
  uint32_t top = 0, bot = 0;  // sp % kLds and base % kLds, kept incrementally (no division)

  __device__ static uint32_t next(uint32_t i) { return i + 1u == uint32_t(kLds) ? 0u : i + 1u; }
  __device__ static uint32_t prev(uint32_t i) { return i == 0u ? uint32_t(kLds - 1) : i - 1u; }
  // Room in the ring for n more entries: the oldest move to the spill area
  // (the entries, in the order, that n pushes spilling one at a time when
  // the ring is full would move; n <= kLds).
  __device__ void reserve(uint32_t n) {
    while (sp - base + n > uint32_t(kLds)) {
      spill[2 * base] = lds[bot * kBlock];
      spill[2 * base + 1] = lds[(kLds + bot) * kBlock];
      ++base;
      bot = next(bot);
    }
  }
  __device__ void push_reserved(uint32_t link, float dist) {
    lds[top * kBlock] = link;
    lds[(kLds + top) * kBlock] = __float_as_uint(dist);
    ++sp;
    top = next(top);
  }
  __device__ void push(uint32_t link, float dist) {
    reserve(1);
    push_reserved(link, dist);
  }
  __device__ void pop(uint32_t& link, float& dist) {
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
  __device__ bool empty() const { return sp == 0; }
};

constexpr uint32_t kLinkLeaf = 0x80000000u, kLinkNone = 0xFFFFFFFFu;
constexpr float kShrink32 = 1.0f - 1.0f / 1048576.0f;  // 1 - 2^-20

// A 4-wide node's children: decoded boxes (fmaf(code, scale, origin), the
// expression the host rounded the codes outward for) and links.
struct Node4 {
  float lo[4][3], hi[4][3];
  uint32_t link[4];
};

__device__ inline void load_node4(const float4* nodes, uint32_t n, Node4& nd) {
  const float4* r = nodes + size_t(n) * 4;
  const float4 A = r[0], B = r[1], C = r[2], L = r[3];
  const float org[3] = {A.x, A.y, A.z}, sc[3] = {A.w, B.x, B.y};
  const uint32_t ql[3] = {__float_as_uint(B.z), __float_as_uint(C.x), __float_as_uint(C.z)};
  const uint32_t qh[3] = {__float_as_uint(B.w), __float_as_uint(C.y), __float_as_uint(C.w)};
  nd.link[0] = __float_as_uint(L.x);
  nd.link[1] = __float_as_uint(L.y);
  nd.link[2] = __float_as_uint(L.z);
  nd.link[3] = __float_as_uint(L.w);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      nd.lo[i][k] = fmaf(float((ql[k] >> (8 * i)) & 255u), sc[k], org[k]);
      nd.hi[i][k] = fmaf(float((qh[k] >> (8 * i)) & 255u), sc[k], org[k]);
    }
}

// Slab test of child i in float32 (conservative: the far distance inflated).
template <bool kConservative>
__device__ inline bool slab4(const Node4& nd, int i, const float o[3], const float inv[3], float limit, float& tn) {
  float a = 0.0f, b = limit;
  float lo = 0.0f, hi = 0.0f;
  const float x0 = (nd.lo[i][0] - o[0]) * inv[0], x1 = (nd.hi[i][0] - o[0]) * inv[0];
  const float y0 = (nd.lo[i][1] - o[1]) * inv[1], y1 = (nd.hi[i][1] - o[1]) * inv[1];
  const float z0 = (nd.lo[i][2] - o[2]) * inv[2], z1 = (nd.hi[i][2] - o[2]) * inv[2];
  lo = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), a));
  hi = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
  hi = kConservative ? fminf(hi, b) * kMargin32 : fminf(hi, b);
  tn = lo;
  return hi >= lo;
}

// Double-precision slab of child i (double-precision rays).
__device__ inline bool slab4_64(const Node4& nd, int i, const double o[3], const double inv[3], double limit,
                                double& tn) {
#pragma clang fp contract(off)
  double a = 0.0, b = limit;
  for (int k = 0; k < 3; ++k) {
    double t0 = (double(nd.lo[i][k]) - o[k]) * inv[k];
    double t1 = (double(nd.hi[i][k]) - o[k]) * inv[k];
    if (t0 > t1) {
      const double tt = t0;
      t0 = t1;
      t1 = tt;
    }
    if (t0 > a) a = t0;
    if (t1 < b) b = t1;
  }
  tn = a;
  return b * kMargin64 >= a;
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
  uint32_t slot;  // device slot
  uint32_t key;   // the caller's slot of it: ties go to the lowest
};

template <Prec kPrec, bool kAny>
__device__ inline bool test_leaf(const float4* tris, const uint32_t* ref_slots, uint32_t first, uint32_t count,
                                 const double rod[3], const double rdd[3], const float rof[3], const float rdf[3],
                                 double tmin, double tmax, Best& best) {
  for (uint32_t s = first; s < first + count; ++s) {
    const float4 a = tris[size_t(s) * 3], b = tris[size_t(s) * 3 + 1], c = tris[size_t(s) * 3 + 2];
    if (kPrec == Prec::kExact) {
      double t, u, v;
      if (!mt_exact(rod, rdd, a, b, c, t, u, v)) continue;
      if (!(t > tmin && t < tmax)) continue;
      if (t <= best.t) {
        const uint32_t key = ref_slots ? ref_slots[s] : s;
        if (t < best.t || key < best.key) {
          best = Best{t, u, v, s, key};
          if (kAny) return true;
        }
      }
    } else {
      float t, u, v;
      if (!mt_glsl(rof, rdf, a, b, c, t, u, v)) continue;
      if (!(t > kGlslEps && double(t) < tmax)) continue;
      if (double(t) < best.t) {
        best = Best{double(t), double(u), double(v), s, s};
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
  __shared__ uint32_t lds_stack[2 * kLds * kBlock];
  const float4* nodes = reinterpret_cast<const float4*>(bvh.nodes4);
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
    Best best{tmax, 0.0, 0.0, 0xFFFFFFFFu, 0xFFFFFFFFu};

    LaneStack<kLds> st{(lds_word*)(lds_stack + threadIdx.x), spill + size_t(lane_global) * spill_depth};
    if (bvh.tri_count > 0) st.push(bvh.root4, 0.0f);
    while (!st.empty()) {
      uint32_t link;
      float dist;
      st.pop(link, dist);
      // dist is a lower bound of the entry distance (shrunk by 2^-20 in exact
      // mode); entries strictly beyond the closest hit cannot hold a hit that
      // wins, ties included.
      if (!kAny && double(dist) > best.t) continue;
      if (link & kLinkLeaf) {
        const uint32_t first = link & 0x03FFFFFFu, count = ((link >> 26) & 31u) + 1u;
        if (test_leaf<kPrec, kAny>(tris, bvh.ref_slots, first, count, rod, rdd, rof, rdf, lower, tmax, best)) break;
        continue;
      }
      Node4 nd;
      load_node4(nodes, link, nd);
      float tn[4];
      uint32_t l[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        l[i] = nd.link[i];
        bool hit;
        float t;
        if (kF64Rays && kPrec == Prec::kExact) {
          double td;
          hit = slab4_64(nd, i, rod, invd, best.t, td);
          t = __double2float_rd(td) * kShrink32;
        } else if (kPrec == Prec::kExact) {
          hit = slab4<true>(nd, i, rof, invf, __double2float_ru(best.t), t);
          t *= kShrink32;
        } else {
          hit = slab4<false>(nd, i, rof, invf, float(best.t), t);
        }
        tn[i] = (l[i] != kLinkNone && hit) ? t : INFINITY;
      }
      if (!kAny) {
        auto cs = [&](int a, int b) {
          if (tn[a] < tn[b]) {
            const float td = tn[a];
            tn[a] = tn[b];
            tn[b] = td;
            const uint32_t ld = l[a];
            l[a] = l[b];
            l[b] = ld;
          }
        };
        cs(0, 1);
        cs(2, 3);
        cs(0, 2);
        cs(1, 3);
        cs(1, 2);
      }
      if constexpr (kLds >= 4) {
        st.reserve(uint32_t(tn[0] != INFINITY) + uint32_t(tn[1] != INFINITY) + uint32_t(tn[2] != INFINITY) +
                   uint32_t(tn[3] != INFINITY));
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (tn[i] != INFINITY) st.push_reserved(l[i], tn[i]);
      } else {  // a ring smaller than a node's children: room checked per push
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (tn[i] != INFINITY) st.push(l[i], tn[i]);
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

// Where a launch's traversal stacks spill: the context's growable area, or
// a fixed area (the no-allocation entry point) that bounds the grid.
struct SpillArea {
  uint32_t* fixed = nullptr;  // null: the context's area, grown to fit
  size_t fixed_bytes = 0;
};

template <Prec kPrec, bool kAny, bool kF64Rays, int kLds = kLdsStack>
Status launch_variant(Context* ctx, const DevBvh& bvh, const RayQuery& q, hipStream_t stream,
                      const SpillArea& area) {
  auto kern = k_intersect<kPrec, kAny, kF64Rays, kLds>;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0) != hipSuccess || per_cu < 1) per_cu = 4;
  uint64_t blocks_needed = (uint64_t(q.n) + kBlock - 1) / kBlock;
  uint64_t resident = uint64_t(per_cu) * uint64_t(ctx->compute_units > 0 ? ctx->compute_units : 256);
  uint32_t grid = uint32_t(blocks_needed < resident ? blocks_needed : resident);
  if (grid == 0) return LUCENT_OK;
  uint32_t* spill = nullptr;
  uint32_t depth = 0;  // words per lane
  if (!area.fixed) {
    depth = 2 * bvh.stack4;  // (link, distance) entries
    Status s = ensure_spill(ctx, size_t(grid) * kBlock * depth * sizeof(uint32_t));
    if (s != LUCENT_OK) return s;
    spill = static_cast<uint32_t*>(ctx->spill);
  } else {
    // Only entries beyond the kLds held in LDS reach the spill area: a push
    // spills entry `base` when kLds entries are above it, so spill indexes
    // stay below 2 * (stack4 - kLds).  The grid-stride loop lets a smaller
    // grid cover every ray, so the grid shrinks to what the area holds.
    depth = bvh.stack4 > uint32_t(kLds) ? 2 * (bvh.stack4 - kLds) : 2;
    const size_t per_block = size_t(kBlock) * depth * sizeof(uint32_t);
    const size_t fit = area.fixed_bytes / per_block;
    if (fit == 0)
      return fail(LUCENT_E_UNSUPPORTED, "lucent_intersect_device_noalloc: the tree needs " +
                                            std::to_string(per_block) + " B of stack spill per block, above the " +
                                            std::to_string(area.fixed_bytes) + " B scratch area");
    if (grid > fit) grid = uint32_t(fit);
    spill = area.fixed;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, bvh, q.org, q.dir, q.n, q.tmin, q.tmax, q.t, q.id,
                     q.u, q.v, spill, depth);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_intersect launch");
  return LUCENT_OK;
}

Status launch_intersect_in(Context* ctx, const DevBvh& bvh, const RayQuery& q, hipStream_t stream,
                          const SpillArea& a) {
  const bool any = (q.flags & LUCENT_QUERY_ANY) != 0;
  const bool glsl = (q.flags & LUCENT_PREC_GLSL_FP32) != 0;
  if (q.flags & LUCENT_FLAG_TINY_LDS_STACK) {
    // Test aid: a 2-entry LDS ring forces the global spill path on any tree.
    if (q.f64_rays || any || glsl) return fail(LUCENT_E_INVALID, "LUCENT_FLAG_TINY_LDS_STACK: exact closest-hit float32 rays only");
    return launch_variant<Prec::kExact, false, false, 2>(ctx, bvh, q, stream, a);
  }
  if (q.f64_rays) {
    if (glsl) return any ? launch_variant<Prec::kGlsl, true, true>(ctx, bvh, q, stream, a)
                         : launch_variant<Prec::kGlsl, false, true>(ctx, bvh, q, stream, a);
    return any ? launch_variant<Prec::kExact, true, true>(ctx, bvh, q, stream, a)
               : launch_variant<Prec::kExact, false, true>(ctx, bvh, q, stream, a);
  }
  if (glsl) return any ? launch_variant<Prec::kGlsl, true, false>(ctx, bvh, q, stream, a)
                       : launch_variant<Prec::kGlsl, false, false>(ctx, bvh, q, stream, a);
  return any ? launch_variant<Prec::kExact, true, false>(ctx, bvh, q, stream, a)
             : launch_variant<Prec::kExact, false, false>(ctx, bvh, q, stream, a);
}

}  // namespace

Status launch_intersect(Context* ctx, const DevBvh& bvh, const RayQuery& q, void* stream_) {
  return launch_intersect_in(ctx, bvh, q, static_cast<hipStream_t>(stream_), SpillArea{});
}

}  // namespace lucent

// ---------------------------------------------------------------------------
// The no-allocation query (lucent_intersect_device_noalloc): its traversal
// stacks spill into this fixed device array instead of the context's
// growable area, so a call allocates nothing.  Every device the library runs
// on holds its own copy for the life of the process; calls on one device
// share it, so calls that overlap on one device race.  Host code uses the
// address hipGetSymbolAddress returns for the current device (kept per
// device by noalloc_scratch_init), never the array's name.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint32_t scratch[lucent::kNoAllocScratchBytes / sizeof(uint32_t)];

namespace lucent {
namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxNoAllocDevices] = {};
}  // namespace

Status noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxNoAllocDevices) return fail(LUCENT_E_UNSUPPORTED, "device index above the no-allocation table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return LUCENT_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress(scratch)");
  if (!p) return fail(LUCENT_E_HIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch_ptr[device] = p;
  return LUCENT_OK;
}

Status launch_intersect_noalloc(Context* ctx, const DevBvh& bvh, const RayQuery& q, void* stream_) {
  void* p = nullptr;
  if (ctx->device >= 0 && ctx->device < kMaxNoAllocDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch_ptr[ctx->device];
  }
  if (!p) return fail(LUCENT_E_HIP, "no-allocation scratch area was not set up for this device");
  SpillArea a;
  a.fixed = static_cast<uint32_t*>(p);
  a.fixed_bytes = kNoAllocScratchBytes;
  return launch_intersect_in(ctx, bvh, q, static_cast<hipStream_t>(stream_), a);
}

}  // namespace lucent
