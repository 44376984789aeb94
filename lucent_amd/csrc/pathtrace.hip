// Progressive path tracing on gfx950.  The integrator is pt_device.hpp
// (trace.frag semantics in float32); two launch structures run it:
//
// * k_trace_sample — the reference's structure (Renderer.traceSamples,
//   src/renderer.ts:702-737): one launch per sample over the whole film, one
//   lane per pixel, each wave64 an 8x8 pixel tile so that neighbouring rays
//   share BVH nodes in the caches.  A wave lasts as long as its longest path.
//
// * k_trace_tiles — persistent waves pull 8x8 tile jobs from an atomic queue
//   (one or many films / scenes per launch) and run ALL of their samples:
//   a lane takes one pixel, advances its path one bounce at a time and, as
//   soon as the path ends, folds it into the pixel's running mean and
//   starts the pixel's next sample; after the pixel's last sample it takes
//   the next free pixel of the wave's jobs.  Lanes are refilled instead of
//   idling behind the longest path or the busiest pixel of the wave.  Per
//   pixel the samples run in seed order through the same accumulation step,
//   so the film is bit-identical to spp k_trace_sample launches.  The running means stay in the film
//   (72 B of traffic per sample) rather than in registers: registers, not
//   bandwidth, bound this kernel's occupancy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "pt_device.hpp"
#include "render_host.hpp"

namespace lucent {
namespace {

constexpr int kPtBlock = 64;
#ifndef LUCENT_RING_JOBS
#define LUCENT_RING_JOBS 4
#endif
#ifndef LUCENT_SHADE_BATCH
#define LUCENT_SHADE_BATCH 32
#endif
#ifndef LUCENT_TRACE_TURNS
#define LUCENT_TRACE_TURNS 8
#endif
constexpr int kTile = 8;

__global__ __launch_bounds__(kPtBlock) void k_trace_sample(PtScene sc, PtFrame f, PtFilm film, uint32_t* spill,
                                                          uint32_t spill_depth, uint32_t tiles_x) {
  __shared__ uint32_t lds[pt::kPtStackWords * kPtBlock];
  const uint32_t tile = blockIdx.x;
  const uint32_t x = (tile % tiles_x) * kTile + (threadIdx.x & (kTile - 1));
  const uint32_t y = (tile / tiles_x) * kTile + (threadIdx.x / kTile);
  if (x >= f.width || y >= f.height) return;
  const uint32_t idx = y * f.width + x;
  pt::Rng rng = pt::seed_rng(idx, f.seed);
  pt::v3 ro, rd;
  pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  pt::PtStack<kPtBlock> st{pt::as_lds(lds + threadIdx.x), spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
                                 0, 0};
  pt::Guide g;
  const pt::v3 c = pt::trace_path(sc, f, ro, rd, rng, st, g);
  pt::accumulate(film, idx, c, g);
}

// projectToPrevious (trace.frag:274-282): the continuous pixel position of
// d (a point relative to the previous camera, or a direction) in the
// previous view, if it lies inside it.
__device__ inline bool project_to_previous(const PrevView& pv, pt::v3 d, float& px, float& py) {
  const float z = pt::dot(d, pt::mk(pv.cam_fwd));
  if (z <= 1e-6f) return false;
  const float aspect = float(pv.width) / float(pv.height);
  const float ndx = pt::dot(d, pt::mk(pv.cam_right)) / (z * pv.tan_half_fov * aspect);
  const float ndy = pt::dot(d, pt::mk(pv.cam_up)) / (z * pv.tan_half_fov);
  px = (ndx * 0.5f + 0.5f) * float(pv.width);
  py = (ndy * 0.5f + 0.5f) * float(pv.height);
  return px >= 0.0f && py >= 0.0f && px < float(pv.width) && py < float(pv.height);
}

// reprojectHistory (trace.frag:284-322): bilinear fetch of the previous
// running means from the taps that saw the same surface (or the same sky);
// false when too little of the footprint survives (disocclusion).
__device__ inline bool reproject_history(const PrevView& pv, const pt::Guide& g, pt::v3 rd, pt::PixelAccum& a) {
  a.color = make_float4(0.f, 0.f, 0.f, 0.f);
  a.albedo = a.color;
  a.normal = a.color;
  a.n = 0.0f;
  const bool surface = g.coverage > 0.5f;
  const pt::v3 d = surface ? g.position - pt::mk(pv.cam_pos) : rd;
  float px, py;
  if (!project_to_previous(pv, d, px, py)) return false;
  const float expected = pt::length(d);
  const float fx = px - 0.5f, fy = py - 0.5f;
  const float flx = floorf(fx), fly = floorf(fy);
  const int bx = int(flx), by = int(fly);
  const float tx = fx - flx, ty = fy - fly;
  float wsum = 0.0f;
  for (int i = 0; i < 4; ++i) {
    const int ox = i & 1, oy = i >> 1;
    const int qx = bx + ox, qy = by + oy;
    if (qx < 0 || qy < 0 || qx >= int(pv.width) || qy >= int(pv.height)) continue;
    const pt::PixelAccum q = pt::load_accum(pv.film, uint32_t(qy) * pv.width + uint32_t(qx));
    bool same;
    if (surface) {
      const pt::v3 qn = pt::normalize(pt::mk(q.normal.x + 1e-6f, q.normal.y + 1e-6f, q.normal.z + 1e-6f));
      same = q.albedo.w > 0.5f && fabsf(q.normal.w - expected) < 0.02f * expected && pt::dot(qn, g.normal) > 0.85f;
    } else {
      same = q.albedo.w < 0.5f;
    }
    if (!same) continue;
    const float w = (ox == 1 ? tx : 1.0f - tx) * (oy == 1 ? ty : 1.0f - ty);
    a.color = make_float4(a.color.x + q.color.x * w, a.color.y + q.color.y * w, a.color.z + q.color.z * w,
                          a.color.w + q.color.w * w);
    a.albedo = make_float4(a.albedo.x + q.albedo.x * w, a.albedo.y + q.albedo.y * w, a.albedo.z + q.albedo.z * w,
                           a.albedo.w + q.albedo.w * w);
    a.normal = make_float4(a.normal.x + q.normal.x * w, a.normal.y + q.normal.y * w, a.normal.z + q.normal.z * w,
                           a.normal.w + q.normal.w * w);
    a.n += q.n * w;
    wsum += w;
  }
  if (wsum < 1e-3f) return false;
  a.color = make_float4(a.color.x / wsum, a.color.y / wsum, a.color.z / wsum, a.color.w / wsum);
  a.albedo = make_float4(a.albedo.x / wsum, a.albedo.y / wsum, a.albedo.z / wsum, a.albedo.w / wsum);
  a.normal = make_float4(a.normal.x / wsum, a.normal.y / wsum, a.normal.z / wsum, a.normal.w / wsum);
  a.n = fminf(a.n / wsum, pv.history_cap * g.reuse);
  return true;
}

// The sample after a camera move (trace.frag:324-355 with uReproject = 1):
// one path per pixel; its running means restart from the previous view's,
// reprojected, or from nothing where the pixel was disoccluded.  `film` and
// pv.film must not alias.
__global__ __launch_bounds__(kPtBlock) void k_trace_reprojected(PtScene sc, PtFrame f, PtFilm film, PrevView pv,
                                                                uint32_t* spill, uint32_t spill_depth,
                                                                uint32_t tiles_x) {
  __shared__ uint32_t lds[pt::kPtStackWords * kPtBlock];
  const uint32_t tile = blockIdx.x;
  const uint32_t x = (tile % tiles_x) * kTile + (threadIdx.x & (kTile - 1));
  const uint32_t y = (tile / tiles_x) * kTile + (threadIdx.x / kTile);
  if (x >= f.width || y >= f.height) return;
  const uint32_t idx = y * f.width + x;
  pt::Rng rng = pt::seed_rng(idx, f.seed);
  pt::v3 ro, rd;
  pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  pt::PtStack<kPtBlock> st{pt::as_lds(lds + threadIdx.x), spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
                                 0, 0};
  pt::PrimaryTracer<pt::PtStack<kPtBlock>> tr{{sc, st}};
  pt::PathState p;
  pt::path_begin(p, ro, rd);
  if (int(f.max_bounces) >= 0)
    while (pt::path_step_with(sc, f, p, rng, tr)) {
    }
  pt::PixelAccum a;
  reproject_history(pv, p.guide, rd, a);  // on failure a is all zero: the pixel starts over
  pt::fold_sample(a, p.radiance, p.guide);
  pt::store_accum(film, idx, a);
}

__device__ __forceinline__ void begin_sample(const PtFrame& f, uint32_t x, uint32_t y, uint32_t idx, uint32_t seed,
                                             pt::Rng& rng, pt::PathState& p) {
  rng = pt::seed_rng(idx, seed);
  pt::v3 ro, rd;
  pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  pt::path_begin(p, ro, rd);
}

// Takes the next job from the launch's queue: one atomic by lane 0, the
// result broadcast.  Called with all 64 lanes active.
__device__ __forceinline__ uint32_t take_job(uint32_t* counter) {
  uint32_t v = 0;
  if (threadIdx.x == 0) v = atomicAdd(counter, 1u);
  return __builtin_amdgcn_readlane(v, 0);
}

// Runs one item's jobs on this wave, starting with `job`, already taken.
// Each lane samples one pixel at a time: all of the pixel's spp samples in
// seed order, one bounce per loop turn, then it takes the next pixel slot of
// the job being handed out (slots are the valid pixels of an 8x8 tile; a
// ballot ranks the lanes asking, so a turn hands out as many slots as there
// are idle lanes).  When the job runs out the wave takes another from the
// queue.  A job of a later item (or the end of the queue) stops the hand-out;
// the wave finishes its pixels and returns that job.  So lanes do not idle
// behind pixels whose paths are long (a tile straddling the sky and an
// object, a glass surface) until the queue is empty.
template <class StackT>
__device__ __forceinline__ uint32_t run_item(const TileItem& it, uint32_t job, uint32_t job_base, uint32_t job_end,
                                             uint32_t* counter, StackT& st) {
  const PtFrame& f = it.frame;
  // The first seed, read once per item in uniform control flow (a scalar).
  const uint32_t seed0 = it.seed_src ? __builtin_amdgcn_readfirstlane(*it.seed_src) : f.seed;
  const uint32_t lane = threadIdx.x;
  // The job being handed out (wave-uniform): tile corner, width, slots.
  uint32_t x0 = 0, y0 = 0, tw = 1, slots = 0, handed = 0;
  auto open_job = [&](uint32_t j) {
    const uint32_t t = j - job_base;
    x0 = (t % it.tiles_x) * kTile;
    y0 = (t / it.tiles_x) * kTile;
    tw = min(uint32_t(kTile), f.width - x0);
    slots = tw * min(uint32_t(kTile), f.height - y0);
    handed = 0;
  };
  open_job(job);
  bool taking = true;
  uint32_t next = 0;
  // This lane's pixel and its current sample.
  bool active = false;
  uint32_t x = 0, y = 0, idx = 0, k = 0;
  pt::Rng rng{};
  pt::PathState p;
  for (;;) {
    uint64_t idle = __ballot(!active);
    while (taking && idle != 0) {
      if (handed == slots) {
        const uint32_t j = take_job(counter);
        if (j >= job_end) {
          taking = false;
          next = j;
          break;
        }
        open_job(j);
        continue;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
      const uint32_t avail = slots - handed;
      if (((idle >> lane) & 1u) && rank < avail) {
        const uint32_t s = handed + rank;
        x = x0 + s % tw;
        y = y0 + s / tw;
        idx = y * f.width + x;
        k = 0;
        begin_sample(f, x, y, idx, seed0, rng, p);
        active = true;
      }
      handed += min(uint32_t(__popcll(idle)), avail);
      idle = __ballot(!active);
    }
    if (active && !pt::path_step(it.scene, f, p, rng, st)) {
      if (it.sample_radiance)
        pt::st_global(reinterpret_cast<float4*>(it.sample_radiance) + size_t(k) * f.width * f.height + idx,
            make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f));
      pt::accumulate(it.film, idx, p.radiance, p.guide);
      if (++k >= it.spp)
        active = false;
      else
        begin_sample(f, x, y, idx, seed0 + k, rng, p);
    }
    if (!taking && __ballot(active) == 0) return next;
  }
}

// The fused loop's three light-sample slots per lane: two float4 each, the
// shadow ray's direction and length, and its unoccluded contribution.  (The
// ray's origin is the step's shadow origin, which the path state keeps.)
constexpr uint32_t kSlotFloat4 = 6;
// Per-lane words of the tile kernels' scratch area: the light-sample slots
// (used when they do not live in LDS), then the traversal stack's spill words.
constexpr uint32_t kShadowWords = 4 * kSlotFloat4;
__host__ __device__ inline uint32_t tile_lane_words(uint32_t stack_words) {
  return kShadowWords + ((stack_words + 3u) & ~3u);
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) f32x4 lds_f4;

// Slots in LDS, lane-interleaved (element e of a lane is e * kPtBlock
// float4 after its element 0), so a wave's access is one conflict-free
// ds_read/ds_write_b128.
struct LdsSlots {
  lds_f4* base;
  __device__ __forceinline__ float4 load(uint32_t e) const {
    const f32x4 v = base[e * kPtBlock];
    return make_float4(v.x, v.y, v.z, v.w);
  }
  __device__ __forceinline__ void store(uint32_t e, float4 v) {
    f32x4 w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    w.w = v.w;
    base[e * kPtBlock] = w;
  }
};
// Slots in the lane's own scratch area (occupancies whose LDS share holds
// only the traversal stack).
struct AreaSlots {
  float4* base;
  __device__ __forceinline__ float4 load(uint32_t e) const { return base[e]; }
  __device__ __forceinline__ void store(uint32_t e, float4 v) { base[e] = v; }
};

// LDS of the tile kernels per lane at kWaves waves per SIMD (four SIMDs
// share a CU's 160 KiB: 40 words per lane at 4 waves, 53 at 3, 32 at 5, 26
// at 6): the traversal stack's ring of kStack entries (2 words each) and, up
// to 4 waves, the light-sample slots.
template <int kWaves>
struct TileLds {
  static constexpr bool kSlots = kWaves <= 4;
  static constexpr int kStack = kWaves == 4 ? 8 : 12;
  static constexpr int kWords = 2 * kStack + (kSlots ? 4 * int(kSlotFloat4) : 0);
  using Stack = pt::Stack4<kStack, kPtBlock>;
};

// The fused loop's path-step tracer: the closest hit was traced before the
// step; each light sample goes to slot `which` (its ray and its
// contribution, evaluated here: an unoccluded sample adds exactly this
// value), to be traced after the step in slot order.  Every light sample of
// a step starts at the step's shadow origin, and so does the continuation
// ray (path_step_with sets p.ro to it before returning true), so the origin
// is stored once, in p.ro; a path that ends at this step no longer reads it.
template <class SlotsT>
struct DeferTracer {
  static constexpr bool kRecordPrimary = false;
  const PtScene& sc;
  SlotsT& slots;
  pt::HitRecord rec;
  uint32_t mask;  // slots filled (bit `which`)
  __device__ __forceinline__ bool closest(pt::v3 ro, pt::v3 rd, pt::Hit& hit, pt::Rng&) {
    return pt::build_hit(sc, ro, rd, rec, hit);
  }
  template <class F>
  __device__ __forceinline__ void nee(int which, pt::v3 o, pt::v3 d, float tmax, pt::Rng&, pt::PathState& p,
                                      F contribution) {
    const pt::v3 c = contribution();
    p.ro = o;
    slots.store(2 * which, make_float4(d.x, d.y, d.z, tmax));
    slots.store(2 * which + 1, make_float4(c.x, c.y, c.z, 0.0f));
    mask |= 1u << which;
  }
};

// Lanes whose rays have ended wait for shading until this many of them are
// ready (or no lane is tracing), so that the shading code runs with most
// lanes of the wave: kShadeBatch for the whole-pixel hand-out, the item's
// shade_batch for the per-sample hand-out.
constexpr int kShadeBatch = LUCENT_SHADE_BATCH;

// run_item for scenes whose traversal draws no random numbers (no
// alpha-blended material): every ray of a lane — closest-hit rays and the
// shadow rays of its light samples — goes through ONE traversal loop, one
// node or leaf step per turn, instead of a closest-hit loop and a shadow
// loop per light sample that the lanes without such a ray sit out.  A lane
// moves through: trace the closest hit; wait; shade the step (light samples
// into its slots, the continuation ray into the path state); trace the
// samples' shadow rays in slot order, adding each unoccluded contribution;
// trace the continuation or, at the path's end, wait and fold the sample
// into the film.  The random numbers are drawn in the same order as
// path_step draws them and the radiance is summed in the same order, so the
// film is bit-identical to run_item's.
enum : uint32_t { kLaneIdle = 0, kLaneTrace = 1, kLaneShade = 2, kLaneFinish = 3 };

// One lane's path in the fused loop: its state, the ray being traced and
// the light samples of the current step.
template <class StackT, class SlotsT>
struct FusedPath {
  const PtScene& sc;
  StackT& st;
  SlotsT slots;  // the lane's light-sample slots
  uint32_t state = kLaneIdle;
  pt::Rng rng{};
  pt::PathState p;
  pt::Trav4 t;
  // The analytic part of the closest hit (kind << 30 | prim); its distance
  // is t.tmax until a triangle is accepted.
  uint32_t analytic = 0;
  bool any = false, cont = false;
  uint32_t which = 0, pending = 0;

  __device__ __forceinline__ void start_closest() {
    pt::HitRecord rec;
    pt::search_analytic(sc, p.ro, p.rd, rec);
    analytic = rec.kind << 30 | rec.prim;
    pt::trav4_begin(sc, t, p.ro, p.rd, rec.t, st);
    any = false;
    state = kLaneTrace;
  }
  // A new path (camera ray of pixel idx = (x, y) with this seed).
  __device__ __forceinline__ void start(const PtFrame& f, uint32_t x, uint32_t y, uint32_t idx, uint32_t seed) {
    begin_sample(f, x, y, idx, seed, rng, p);
    start_closest();
  }
  // The next ray of the step: the lowest pending light sample not blocked by
  // a sphere or the ground (occluded's order), else the continuation, else
  // the path is finished.
  __device__ __forceinline__ void next_ray() {
    while (pending != 0) {
      which = uint32_t(__builtin_ctz(pending));
      pending &= pending - 1u;
      const float4 d = slots.load(2 * which);
      const pt::v3 ro = p.ro, rd = pt::mk(d);
      bool blocked = false;
      for (uint32_t i = 0; i < sc.sphere_count && !blocked; ++i)
        blocked = pt::hit_sphere(ro, rd, pt::ld_global(reinterpret_cast<const float4*>(sc.spheres) + i), d.w) < d.w;
      if (!blocked) blocked = pt::hit_ground(sc, ro, rd, d.w) < d.w;
      if (!blocked) {
        pt::trav4_begin(sc, t, ro, rd, d.w, st);
        any = true;
        state = kLaneTrace;
        return;
      }
    }
    if (cont)
      start_closest();
    else
      state = kLaneFinish;
  }
  // One traversal step of the ray being traced; a finished shadow ray adds
  // its contribution when unoccluded and moves on, a finished closest-hit
  // ray waits for shading.
  __device__ __forceinline__ void trace_turn() {
    bool blocked = false;
    if (!pt::trav4_done(t, st)) blocked = pt::trav4_step_mixed(sc, t, st, rng, any);
    if (blocked || pt::trav4_done(t, st)) {
      if (any) {
        if (!blocked) p.radiance = p.radiance + pt::mk(slots.load(2 * which + 1));
        next_ray();
      } else {
        state = kLaneShade;
      }
    }
  }
  // Shades the step at the closest hit: light samples into the slots, the
  // continuation into the path state.
  __device__ __forceinline__ void shade(const PtFrame& f) {
    pt::HitRecord rec;
    rec.t = t.tmax;
    if (t.best >= 0) {
      rec.kind = pt::kHitTri;
      rec.prim = uint32_t(t.best);
      rec.u = t.bu;
      rec.v = t.bv;
    } else {
      rec.kind = analytic >> 30;
      rec.prim = analytic & 0x3FFFFFFFu;
      rec.u = rec.v = 0.0f;
    }
    DeferTracer<SlotsT> tr{sc, slots, rec, 0u};
    cont = pt::path_step_with(sc, f, p, rng, tr);
    pending = tr.mask;
    next_ray();
  }
};

// True when the wave should shade now: enough lanes wait, or none traces.
__device__ __forceinline__ bool shade_now(uint32_t state, uint32_t batch = kShadeBatch) {
  const uint64_t ready = __ballot(state == kLaneShade || state == kLaneFinish);
  return ready != 0 && (uint32_t(__popcll(ready)) >= batch || __ballot(state == kLaneTrace) == 0);
}

// Up to kTraceTurns traversal steps while some lane traces, between the
// wave's checks for shading and new work: the checks cost little, but a
// loop of traversal steps alone keeps the wave on its hot code (measured:
// 8 steps per check, 128 -> 116 ms for 1440x900 at 64 samples).
constexpr int kTraceTurns = LUCENT_TRACE_TURNS;

template <class Lane>
__device__ __forceinline__ void trace_turns(Lane& lp, uint32_t turns = kTraceTurns) {
  for (uint32_t turn = 0; turn < turns && __ballot(lp.state == kLaneTrace) != 0; ++turn)
    if (lp.state == kLaneTrace) lp.trace_turn();
}

template <class StackT, class SlotsT>
__device__ __forceinline__ uint32_t run_item_fused(const TileItem& it, uint32_t job, uint32_t job_base, uint32_t job_end,
                                                   uint32_t* counter, StackT& st, SlotsT slots) {
  const PtFrame& f = it.frame;
  const uint32_t seed0 = it.seed_src ? __builtin_amdgcn_readfirstlane(*it.seed_src) : f.seed;
  const uint32_t lane = threadIdx.x;
  uint32_t x0 = 0, y0 = 0, tw = 1, slots_in_job = 0, handed = 0;
  auto open_job = [&](uint32_t j) {
    const uint32_t t = j - job_base;
    x0 = (t % it.tiles_x) * kTile;
    y0 = (t / it.tiles_x) * kTile;
    tw = min(uint32_t(kTile), f.width - x0);
    slots_in_job = tw * min(uint32_t(kTile), f.height - y0);
    handed = 0;
  };
  open_job(job);
  bool taking = true;
  uint32_t next = 0;
  uint32_t idx = 0, k = 0;
  FusedPath<StackT, SlotsT> lp{it.scene, st, slots};
  for (;;) {
    uint64_t idle = __ballot(lp.state == kLaneIdle);
    while (taking && idle != 0) {
      if (handed == slots_in_job) {
        const uint32_t j = take_job(counter);
        if (j >= job_end) {
          taking = false;
          next = j;
          break;
        }
        open_job(j);
        continue;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
      const uint32_t avail = slots_in_job - handed;
      if (((idle >> lane) & 1u) && rank < avail) {
        const uint32_t s = handed + rank;
        idx = (y0 + s / tw) * f.width + x0 + s % tw;
        k = 0;
        lp.start(f, idx % f.width, idx / f.width, idx, seed0);
      }
      handed += min(uint32_t(__popcll(idle)), avail);
      idle = __ballot(lp.state == kLaneIdle);
    }
    trace_turns(lp);
    if (shade_now(lp.state)) {
      if (lp.state == kLaneShade) lp.shade(f);
      if (lp.state == kLaneFinish) {
        if (it.sample_radiance)
          pt::st_global(reinterpret_cast<float4*>(it.sample_radiance) + size_t(k) * f.width * f.height + idx,
              make_float4(lp.p.radiance.x, lp.p.radiance.y, lp.p.radiance.z, 0.0f));
        pt::accumulate(it.film, idx, lp.p.radiance, lp.p.guide);
        if (++k >= it.spp)
          lp.state = kLaneIdle;
        else
          lp.start(f, idx % f.width, idx / f.width, idx, seed0 + k);
      }
    }
    if (!taking && __ballot(lp.state != kLaneIdle) == 0) return next;
  }
}

// A finished path as the film's accumulation step takes it: three float4
// (colour, depth), (albedo, coverage), (normal, 0).
constexpr uint32_t kRecordFloat4 = 3;

// run_item_fused for one film with the work handed out per pixel SAMPLE
// (the job's pixels at sample 0, then at sample 1, ...) instead of per
// pixel, so that at the end of the queue lanes do not idle behind the last
// pixels' remaining samples (a lane sees only a few pixels of a 1440x900
// film, one of a 512x512 one).  A finished path is stored as a record
// (slot job * 64 + pixel, sample k) and counted; the lane that brings a
// pixel's count to spp folds the pixel's records into the film in sample
// order, with the accumulation step of the other structures, so the film is
// bit-identical.  All samples of a job are handed out by the wave that took
// it, so records and counters are shared within one wave only
// (workgroup-scope ordering).
template <class StackT, class SlotsT>
__device__ __forceinline__ void run_units_fused(const TileItem& it, uint32_t job, uint32_t n_jobs, uint32_t* counter,
                                                StackT& st, SlotsT slots, float4* __restrict__ records,
                                                uint32_t* __restrict__ arrived) {
  const PtFrame& f = it.frame;
  const uint32_t seed0 = it.seed_src ? __builtin_amdgcn_readfirstlane(*it.seed_src) : f.seed;
  const uint32_t spp = it.spp;
  const uint32_t lane = threadIdx.x;
  const uint32_t tile_w = it.tile_w, tile_h = it.tile_h, tile_px = tile_w * tile_h;
  const uint32_t batch = it.shade_batch, turns = it.trace_turns;
  // The job being handed out (wave-uniform): tile corner, width, pixels,
  // and the next of its px * spp samples.
  uint32_t cur = 0, x0 = 0, y0 = 0, tw = 1, px = 0, units = 0, handed = 0;
  auto open_job = [&](uint32_t j) {
    cur = j;
    x0 = (j % it.tiles_x) * tile_w;
    y0 = (j / it.tiles_x) * tile_h;
    tw = min(tile_w, f.width - x0);
    px = tw * min(tile_h, f.height - y0);
    units = px * spp;
    handed = 0;
  };
  open_job(job);
  bool taking = true;
  uint32_t idx = 0, k = 0, rslot = 0;
  FusedPath<StackT, SlotsT> lp{it.scene, st, slots};
  for (;;) {
    uint64_t idle = __ballot(lp.state == kLaneIdle);
    while (taking && idle != 0) {
      if (handed == units) {
        const uint32_t j = take_job(counter);
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
        const uint32_t slot = u % px;
        k = u / px;
        rslot = cur * tile_px + slot;
        idx = (y0 + slot / tw) * f.width + x0 + slot % tw;
        lp.start(f, idx % f.width, idx / f.width, idx, seed0 + k);
      }
      handed += min(uint32_t(__popcll(idle)), avail);
      idle = __ballot(lp.state == kLaneIdle);
    }
    trace_turns(lp, turns);
    if (shade_now(lp.state, batch)) {
      if (lp.state == kLaneShade) lp.shade(f);
      if (lp.state == kLaneFinish) {
        const pt::PathState& p = lp.p;
        if (it.sample_radiance)
          pt::st_global(reinterpret_cast<float4*>(it.sample_radiance) + size_t(k) * f.width * f.height + idx,
              make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f));
        float4* r = records + (size_t(rslot) * spp + k) * kRecordFloat4;
        r[0] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, p.guide.depth);
        r[1] = make_float4(p.guide.albedo.x, p.guide.albedo.y, p.guide.albedo.z, p.guide.coverage);
        r[2] = make_float4(p.guide.normal.x, p.guide.normal.y, p.guide.normal.z, 0.0f);
        const uint32_t before =
            __hip_atomic_fetch_add(arrived + rslot, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (before + 1u == spp) {
          pt::PixelAccum a = pt::load_accum(it.film, idx);
          const float4* q = records + size_t(rslot) * spp * kRecordFloat4;
          for (uint32_t s = 0; s < spp; ++s, q += kRecordFloat4) {
            const float4 c = q[0], al = q[1], n = q[2];
            pt::Guide g = p.guide;
            g.albedo = pt::mk(al.x, al.y, al.z);
            g.coverage = al.w;
            g.normal = pt::mk(n.x, n.y, n.z);
            g.depth = c.w;
            pt::fold_sample(a, pt::mk(c.x, c.y, c.z), g);
          }
          pt::store_accum(it.film, idx, a);
        }
        lp.state = kLaneIdle;
      }
    }
    if (!taking && __ballot(lp.state != kLaneIdle) == 0) return;
  }
}

// Jobs whose records a wave of the batch kernel keeps (a ring).
constexpr uint32_t kRingJobs = LUCENT_RING_JOBS;

// run_units_fused for a launch over many films (a batch): the wave keeps
// the records of its last kRingJobs jobs in its own area (kRingJobs x 64
// pixels x s_max samples), and opens a job into a ring slot only once the
// job that used the slot before has folded all its pixels (its samples were
// all handed out kRingJobs - 1 jobs earlier, so the wave seldom waits).  A change of film waits only for the paths in flight, not for the
// lanes' pixels' remaining samples.  Returns the first job taken that
// belongs to another film (or >= the queue's end).
template <class StackT, class SlotsT>
__device__ __forceinline__ uint32_t run_units_ring(const TileItem& it, uint32_t job, uint32_t job_base, uint32_t job_end,
                                                   uint32_t* counter, StackT& st, SlotsT slots,
                                                   float4* __restrict__ records, uint32_t* __restrict__ arrived,
                                                   uint32_t s_max) {
  const PtFrame& f = it.frame;
  const uint32_t seed0 = it.seed_src ? __builtin_amdgcn_readfirstlane(*it.seed_src) : f.seed;
  const uint32_t spp = it.spp;
  const uint32_t lane = threadIdx.x;
  const uint32_t tile_w = it.tile_w, tile_h = it.tile_h;
  const uint32_t batch = it.shade_batch, turns = it.trace_turns;
  // The job being handed out and the ring (wave-uniform).
  uint32_t x0 = 0, y0 = 0, tw = 1, px = 0, units = 0, handed = 0;
  uint32_t cur_ring = 0, next_ring = 0;
  uint32_t left[kRingJobs] = {};  // pixels not yet folded per ring slot
  auto open_job = [&](uint32_t j) {
    const uint32_t t = j - job_base;
    x0 = (t % it.tiles_x) * tile_w;
    y0 = (t / it.tiles_x) * tile_h;
    tw = min(tile_w, f.width - x0);
    px = tw * min(tile_h, f.height - y0);
    units = px * spp;
    handed = 0;
    cur_ring = next_ring;
    next_ring = next_ring + 1u == kRingJobs ? 0u : next_ring + 1u;
#pragma unroll
    for (uint32_t r = 0; r < kRingJobs; ++r)
      if (r == cur_ring) left[r] = px;
    __hip_atomic_store(arrived + cur_ring * kPtBlock + lane, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  };
  open_job(job);
  bool taking = true;
  uint32_t next = 0;
  uint32_t idx = 0, k = 0, rslot = 0;
  FusedPath<StackT, SlotsT> lp{it.scene, st, slots};
  for (;;) {
    uint64_t idle = __ballot(lp.state == kLaneIdle);
    while (taking && idle != 0) {
      if (handed == units) {
        uint32_t busy = 0;
#pragma unroll
        for (uint32_t r = 0; r < kRingJobs; ++r)
          if (r == next_ring) busy = left[r];
        if (busy != 0) break;  // that slot's job is still folding
        const uint32_t j = take_job(counter);
        if (j >= job_end) {
          taking = false;
          next = j;
          break;
        }
        open_job(j);
        continue;
      }
      const uint32_t rank = __builtin_amdgcn_mbcnt_hi(uint32_t(idle >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(idle), 0u));
      const uint32_t avail = units - handed;
      if (((idle >> lane) & 1u) && rank < avail) {
        const uint32_t u = handed + rank;
        const uint32_t slot = u % px;
        k = u / px;
        rslot = cur_ring * kPtBlock + slot;
        idx = (y0 + slot / tw) * f.width + x0 + slot % tw;
        lp.start(f, idx % f.width, idx / f.width, idx, seed0 + k);
      }
      handed += min(uint32_t(__popcll(idle)), avail);
      idle = __ballot(lp.state == kLaneIdle);
    }
    trace_turns(lp, turns);
    if (shade_now(lp.state, batch)) {
      if (lp.state == kLaneShade) lp.shade(f);
      bool folded = false;
      if (lp.state == kLaneFinish) {
        const pt::PathState& p = lp.p;
        if (it.sample_radiance)
          pt::st_global(reinterpret_cast<float4*>(it.sample_radiance) + size_t(k) * f.width * f.height + idx,
              make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f));
        float4* r = records + (size_t(rslot) * s_max + k) * kRecordFloat4;
        r[0] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, p.guide.depth);
        r[1] = make_float4(p.guide.albedo.x, p.guide.albedo.y, p.guide.albedo.z, p.guide.coverage);
        r[2] = make_float4(p.guide.normal.x, p.guide.normal.y, p.guide.normal.z, 0.0f);
        const uint32_t before =
            __hip_atomic_fetch_add(arrived + rslot, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (before + 1u == spp) {
          pt::PixelAccum a = pt::load_accum(it.film, idx);
          const float4* q = records + size_t(rslot) * s_max * kRecordFloat4;
          for (uint32_t s = 0; s < spp; ++s, q += kRecordFloat4) {
            const float4 c = q[0], al = q[1], n = q[2];
            pt::Guide g = p.guide;
            g.albedo = pt::mk(al.x, al.y, al.z);
            g.coverage = al.w;
            g.normal = pt::mk(n.x, n.y, n.z);
            g.depth = c.w;
            pt::fold_sample(a, pt::mk(c.x, c.y, c.z), g);
          }
          pt::store_accum(it.film, idx, a);
          folded = true;
        }
        lp.state = kLaneIdle;
      }
#pragma unroll
      for (uint32_t r = 0; r < kRingJobs; ++r)
        left[r] -= uint32_t(__popcll(__ballot(folded && rslot / kPtBlock == r)));
    }
    if (!taking && __ballot(lp.state != kLaneIdle) == 0) return next;
  }
}

// A lane's light-sample slots in a tile kernel's LDS (after the stack
// ring's two planes) or in its scratch area.
template <int kWaves>
__device__ __forceinline__ auto tile_slots(uint32_t* lds, uint32_t* area) {
  if constexpr (TileLds<kWaves>::kSlots)
    return LdsSlots{(lds_f4*)(lds + 2 * TileLds<kWaves>::kStack * kPtBlock) + threadIdx.x};
  else
    return AreaSlots{reinterpret_cast<float4*>(area)};
}

// Many films, work handed out per pixel sample (run_units_ring).  Each wave
// owns kRingJobs * 64 * s_max records and kRingJobs * 64 counters.
template <int kWaves>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_trace_batch_units(
    const TileItem* __restrict__ items, const uint32_t* __restrict__ job_begin, uint32_t n_items, uint32_t n_jobs,
    uint32_t* __restrict__ counter, uint32_t* scratch, uint32_t lane_words, float4* __restrict__ records,
    uint32_t* __restrict__ arrived, uint32_t s_max) {
  __shared__ __align__(16) uint32_t lds[TileLds<kWaves>::kWords * kPtBlock];
  uint32_t* area = scratch + size_t(blockIdx.x * kPtBlock + threadIdx.x) * lane_words;
  const auto slots = tile_slots<kWaves>(lds, area);
  typename TileLds<kWaves>::Stack st{pt::as_lds(lds + threadIdx.x), area + kShadowWords, 0, 0};
  float4* my_records = records + size_t(blockIdx.x) * kRingJobs * kPtBlock * s_max * kRecordFloat4;
  uint32_t* my_arrived = arrived + size_t(blockIdx.x) * kRingJobs * kPtBlock;
  uint32_t job = take_job(counter);
  while (job < n_jobs) {
    uint32_t item = 0, hi = n_items;
    while (hi - item > 1) {
      const uint32_t mid = (item + hi) / 2;
      if (job_begin[mid] <= job)
        item = mid;
      else
        hi = mid;
    }
    job = run_units_ring(items[item], job, job_begin[item], job_begin[item + 1], counter, st, slots, my_records,
                         my_arrived, s_max);
  }
}

// One film, work handed out per pixel sample (run_units_fused).  The film's
// descriptor is read through a pointer to device memory: the compiler then
// loads its fields with scalar loads where they are used, where a by-value
// kernel argument kept more of them live (185 instead of 170 VGPRs unbounded;
// 12 B per lane of spill at 3 waves per SIMD, none this way).
template <int kWaves>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_trace_units(
    const TileItem* __restrict__ single, uint32_t n_jobs, uint32_t* __restrict__ counter, uint32_t* scratch,
    uint32_t lane_words, float4* __restrict__ records, uint32_t* __restrict__ arrived) {
  __shared__ __align__(16) uint32_t lds[TileLds<kWaves>::kWords * kPtBlock];
  uint32_t* area = scratch + size_t(blockIdx.x * kPtBlock + threadIdx.x) * lane_words;
  const auto slots = tile_slots<kWaves>(lds, area);
  typename TileLds<kWaves>::Stack st{pt::as_lds(lds + threadIdx.x), area + kShadowWords, 0, 0};
  const uint32_t job = take_job(counter);
  if (job < n_jobs) run_units_fused(*single, job, n_jobs, counter, st, slots, records, arrived);
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Jobs are numbered film by film: film i owns jobs [job_begin[i], job_begin[i+1]).
// kSingle: one film, passed by value (its fields live in the kernel arguments).
// kWaves: the occupancy the register allocation targets (waves per SIMD);
// more waves hide more memory latency at the price of register spills.
// kFused: run_item_fused (no item draws random numbers inside traversal).
// Each lane owns lane_words words of `scratch`: its light-sample slots, then
// its stack's spill words.
template <bool kSingle, int kWaves, bool kFused>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_trace_tiles(const TileItem* __restrict__ single, const TileItem* __restrict__ items,
                                                         const uint32_t* __restrict__ job_begin, uint32_t n_items,
                                                         uint32_t n_jobs, uint32_t* __restrict__ counter,
                                                         uint32_t* scratch, uint32_t lane_words) {
  __shared__ __align__(16) uint32_t lds[TileLds<kWaves>::kWords * kPtBlock];
  uint32_t* area = scratch + size_t(blockIdx.x * kPtBlock + threadIdx.x) * lane_words;
  const auto slots = tile_slots<kWaves>(lds, area);
  typename TileLds<kWaves>::Stack st{pt::as_lds(lds + threadIdx.x), area + kShadowWords, 0, 0};
  uint32_t job = take_job(counter);
  while (job < n_jobs) {
    if constexpr (kSingle) {
      if constexpr (kFused)
        job = run_item_fused(*single, job, 0u, n_jobs, counter, st, slots);
      else
        job = run_item(*single, job, 0u, n_jobs, counter, st);
    } else {
      uint32_t item = 0, hi = n_items;
      while (hi - item > 1) {
        const uint32_t mid = (item + hi) / 2;
        if (job_begin[mid] <= job)
          item = mid;
        else
          hi = mid;
      }
      if constexpr (kFused)
        job = run_item_fused(items[item], job, job_begin[item], job_begin[item + 1], counter, st, slots);
      else
        job = run_item(items[item], job, job_begin[item], job_begin[item + 1], counter, st);
    }
  }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }
// Stores one launch's film descriptor for k_trace_units and resets its job
// counter (stream-ordered, so chunks of one film reuse the same slot).
__global__ void k_begin_units(TileItem* dst, TileItem v, uint32_t* counter) {
  *dst = v;
  *counter = 0u;
}
__global__ void k_add_u32(uint32_t* p, uint32_t v) { *p += v; }
// seeds[i] += items[i].spp: the device seed words of a replayed batch.
__global__ void k_advance_seeds(uint32_t* seeds, const TileItem* items, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) seeds[i] += items[i].spp;
}

}  // namespace

Status launch_trace_sample(Context* ctx, const PtScene& sc, const PtFrame& f, const PtFilm& film, void* stream) {
  const uint32_t tiles_x = (f.width + kTile - 1) / kTile;
  const uint32_t tiles_y = (f.height + kTile - 1) / kTile;
  const uint64_t blocks = uint64_t(tiles_x) * tiles_y;
  if (blocks == 0) return LUCENT_OK;
  const uint32_t depth = pt::pt_spill_words(sc.bvh);
  Status s = ensure_spill(ctx, size_t(blocks) * kPtBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  hipLaunchKernelGGL(k_trace_sample, dim3(uint32_t(blocks)), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream), sc, f,
                     film, static_cast<uint32_t*>(ctx->spill), depth, tiles_x);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_sample launch");
  return LUCENT_OK;
}

Status launch_trace_reprojected(Context* ctx, const PtScene& sc, const PtFrame& f, const PtFilm& film,
                                const PrevView& prev, void* stream) {
  const uint32_t tiles_x = (f.width + kTile - 1) / kTile;
  const uint32_t tiles_y = (f.height + kTile - 1) / kTile;
  const uint64_t blocks = uint64_t(tiles_x) * tiles_y;
  if (blocks == 0) return LUCENT_OK;
  const uint32_t depth = pt::pt_spill_words(sc.bvh);
  Status s = ensure_spill(ctx, size_t(blocks) * kPtBlock * depth * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  hipLaunchKernelGGL(k_trace_reprojected, dim3(uint32_t(blocks)), dim3(kPtBlock), 0, static_cast<hipStream_t>(stream),
                     sc, f, film, prev, static_cast<uint32_t*>(ctx->spill), depth, tiles_x);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_reprojected launch");
  return LUCENT_OK;
}

// The persistent kernels' register variant: 2..6 waves per SIMD as asked,
// else 4 (the fastest on the measured workloads without autotuning).
static uint32_t tile_waves(uint32_t occupancy) { return occupancy >= 2 && occupancy <= 6 ? occupancy : 4; }

size_t tile_spill_bytes(const Context* ctx, const TileItem& item, const TileLaunch& cfg) {
  const uint32_t waves = tile_waves(cfg.occupancy);
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * waves;
  const uint32_t n_jobs = item.spp ? tile_jobs(item.frame.width, item.frame.height) : 0;
  const uint32_t grid = std::min<uint32_t>(n_jobs, uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu);
  return size_t(grid) * kPtBlock * tile_lane_words(pt::pt_spill_words(item.scene.bvh)) * sizeof(uint32_t);
}

Status launch_add_u32(uint32_t* p, uint32_t v, void* stream) {
  hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), p, v);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_add_u32 launch");
  return LUCENT_OK;
}

uint32_t tile_jobs(uint32_t width, uint32_t height) {
  return ((width + kTile - 1) / kTile) * ((height + kTile - 1) / kTile);
}

// A job of tile_px pixels holds tile_px x spp paths, handed out pixel by
// pixel for each sample in turn, so a wave's 64 lanes trace 64 / tile_px
// samples of each pixel together: small tiles keep the lanes' primary rays
// and first hits alike (a 64-spp film ran 27 % faster in 2x2 than in 8x8
// jobs), and enough paths per job keep a wave's lanes fed from the jobs it
// holds (a 1-spp film needs whole 8x8 tiles for that).
void launch_job_tile(const TileLaunch& cfg, uint32_t spp, uint32_t* tile_w, uint32_t* tile_h) {
  if (cfg.tile_w || cfg.tile_h) {
    *tile_w = cfg.tile_w ? cfg.tile_w : kTile;
    *tile_h = cfg.tile_h ? cfg.tile_h : kTile;
    return;
  }
  const uint32_t paths = cfg.job_paths ? cfg.job_paths : kJobPaths;
  const uint32_t px = std::min<uint32_t>(kPtBlock, std::max(1u, paths / std::max(1u, spp)));
  const uint32_t lg = 31u - uint32_t(__builtin_clz(px));  // px >= 1
  *tile_w = 1u << ((lg + 1) / 2);
  *tile_h = 1u << (lg / 2);
}

Status apply_launch_shape(const TileLaunch& cfg, TileItem* item) {
  uint32_t tw = 0, th = 0;
  launch_job_tile(cfg, item->job_spp ? item->job_spp : item->spp, &tw, &th);
  const uint32_t batch = cfg.shade_batch ? cfg.shade_batch : kShadeBatch;
  const uint32_t turns = cfg.trace_turns ? cfg.trace_turns : kTraceTurns;
  if (tw * th > uint32_t(kPtBlock)) return fail(LUCENT_E_INVALID, "launch tile holds more than 64 pixels");
  if (batch > uint32_t(kPtBlock)) return fail(LUCENT_E_INVALID, "shade batch above 64 lanes");
  if (turns > 64) return fail(LUCENT_E_INVALID, "trace turns above 64");
  item->tile_w = tw;
  item->tile_h = th;
  item->shade_batch = batch;
  item->trace_turns = turns;
  item->tiles_x = (item->frame.width + tw - 1) / tw;
  return LUCENT_OK;
}

// Jobs of an item in the per-sample hand-out (its tile shape).
static uint32_t unit_jobs(const TileItem& it) {
  return it.tiles_x * ((it.frame.height + it.tile_h - 1) / it.tile_h);
}

Status launch_trace_units(Context* ctx, TileItem item, const TileLaunch& cfg, hipStream_t stream) {
  Status s = apply_launch_shape(cfg, &item);
  if (s != LUCENT_OK) return s;
  const uint32_t n_jobs = unit_jobs(item);
  const uint32_t waves = tile_waves(cfg.occupancy);
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * waves;
  const uint32_t grid = std::min<uint32_t>(n_jobs, uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu);
  const uint32_t lane_words = tile_lane_words(pt::pt_spill_words(item.scene.bvh));
  s = ensure_spill(ctx, size_t(grid) * kPtBlock * lane_words * sizeof(uint32_t));
  if (s != LUCENT_OK) return s;
  const size_t slots = size_t(n_jobs) * item.tile_w * item.tile_h;
  const size_t per_sample = slots * kRecordFloat4 * sizeof(float4);
  // Path records of one launch at most ctx->record_budget bytes (48 B per
  // path: a 1440x900 film's 64 samples take 4.0 GB of the 288 GB); larger
  // sample counts run as several launches, each continuing the film.
  const uint32_t chunk = uint32_t(std::max<size_t>(1, std::min<size_t>(item.spp, ctx->record_budget / per_sample)));
  const size_t counters_off = (size_t(chunk) * per_sample + 255) / 256 * 256;
  if ((s = ensure_units(ctx, counters_off + slots * sizeof(uint32_t))) != LUCENT_OK) return s;
  // Tile table: [counter | the chunk's TileItem], 256-byte aligned parts.
  if ((s = ensure_tile_table(ctx, 256 + sizeof(TileItem))) != LUCENT_OK) return s;
  float4* records = static_cast<float4*>(ctx->units);
  uint32_t* arrived = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->units) + counters_off);
  uint32_t* counter = static_cast<uint32_t*>(ctx->tile_table);
  TileItem* d_item = reinterpret_cast<TileItem*>(static_cast<char*>(ctx->tile_table) + 256);
  uint32_t* scratch = static_cast<uint32_t*>(ctx->spill);
  const size_t frame_px = size_t(item.frame.width) * item.frame.height;
  const uint32_t spp = item.spp;
  for (uint32_t done = 0; done < spp; done += chunk) {
    TileItem it = item;
    it.spp = std::min(chunk, spp - done);
    it.frame.seed = item.frame.seed + done;
    if (item.sample_radiance) it.sample_radiance = item.sample_radiance + size_t(done) * frame_px * 4;
    hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t((slots + 255) / 256)), dim3(256), 0, stream, arrived, slots, 0u);
    hipLaunchKernelGGL(k_begin_units, dim3(1), dim3(1), 0, stream, d_item, it, counter);
    switch (waves) {
      case 3:
        hipLaunchKernelGGL(k_trace_units<3>, dim3(grid), dim3(kPtBlock), 0, stream, d_item, n_jobs, counter, scratch,
                           lane_words, records, arrived);
        break;
      case 4:
        hipLaunchKernelGGL(k_trace_units<4>, dim3(grid), dim3(kPtBlock), 0, stream, d_item, n_jobs, counter, scratch,
                           lane_words, records, arrived);
        break;
      case 5:
        hipLaunchKernelGGL(k_trace_units<5>, dim3(grid), dim3(kPtBlock), 0, stream, d_item, n_jobs, counter, scratch,
                           lane_words, records, arrived);
        break;
      case 6:
        hipLaunchKernelGGL(k_trace_units<6>, dim3(grid), dim3(kPtBlock), 0, stream, d_item, n_jobs, counter, scratch,
                           lane_words, records, arrived);
        break;
      default:
        hipLaunchKernelGGL(k_trace_units<2>, dim3(grid), dim3(kPtBlock), 0, stream, d_item, n_jobs, counter, scratch,
                           lane_words, records, arrived);
    }
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_units launch");
  return LUCENT_OK;
}

Status plan_batch(const Context* ctx, const TileItem* items, uint32_t n_items, const TileLaunch& cfg, BatchPlan* p) {
  p->items.assign(items, items + n_items);
  p->job_begin.assign(n_items + 1, 0);
  uint32_t depth = 1;
  for (uint32_t i = 0; i < n_items; ++i) depth = std::max(depth, pt::pt_spill_words(items[i].scene.bvh));
  p->waves = tile_waves(cfg.occupancy);
  p->fused = cfg.fused;
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * p->waves;
  const uint32_t slots = uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu;
  p->lane_words = tile_lane_words(depth);
  p->ring = false;
  p->s_max = 0;
  p->records_bytes = p->arrived_bytes = 0;
  if (cfg.fused && cfg.units) {
    // Per-sample hand-out with a record ring of kRingJobs jobs per wave,
    // when the records fit the context's budget (else the pixel hand-out
    // below).  Its jobs are tiles of the launch shape.
    uint32_t s_max = 0;
    for (uint32_t i = 0; i < n_items; ++i) s_max = std::max(s_max, items[i].spp);
    std::vector<TileItem> shaped(items, items + n_items);
    std::vector<uint32_t> ubegin(n_items + 1, 0);
    for (uint32_t i = 0; i < n_items; ++i) {
      Status s = apply_launch_shape(cfg, &shaped[i]);
      if (s != LUCENT_OK) return s;
      ubegin[i + 1] = ubegin[i] + (shaped[i].spp ? unit_jobs(shaped[i]) : 0);
    }
    const uint32_t u_jobs = ubegin[n_items];
    const uint32_t u_grid = std::min<uint32_t>(u_jobs, slots);
    const size_t ring_records = size_t(u_grid) * kRingJobs * kPtBlock * s_max * kRecordFloat4 * sizeof(float4);
    if (s_max > 0 && ring_records <= ctx->record_budget) {
      p->ring = true;
      p->items.swap(shaped);
      p->job_begin.swap(ubegin);
      p->n_jobs = u_jobs;
      p->grid = u_grid;
      p->s_max = s_max;
      p->records_bytes = ring_records;
      p->arrived_bytes = size_t(u_grid) * kRingJobs * kPtBlock * sizeof(uint32_t);
    }
  }
  if (!p->ring) {
    // Whole pixels of 8x8 tiles.
    for (uint32_t i = 0; i < n_items; ++i) {
      TileItem& it = p->items[i];
      it.tiles_x = (it.frame.width + kTile - 1) / kTile;
      p->job_begin[i + 1] = p->job_begin[i] + (it.spp ? tile_jobs(it.frame.width, it.frame.height) : 0);
    }
    p->n_jobs = p->job_begin[n_items];
    p->grid = std::min<uint32_t>(p->n_jobs, slots);
  }
  p->spill_bytes = size_t(p->grid) * kPtBlock * p->lane_words * sizeof(uint32_t);
  return LUCENT_OK;
}

Status launch_batch_planned(const BatchPlan& p, uint32_t* counter, const uint32_t* d_begin, const TileItem* d_items,
                            uint32_t* spill, void* records_, uint32_t* arrived, void* stream_) {
  if (p.n_jobs == 0) return LUCENT_OK;
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  float4* records = static_cast<float4*>(records_);
  const uint32_t n = uint32_t(p.items.size());
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, stream, counter, 0u);
  if (p.ring) {
    auto batch_kernel = p.waves == 3   ? k_trace_batch_units<3>
                        : p.waves == 4 ? k_trace_batch_units<4>
                        : p.waves == 5 ? k_trace_batch_units<5>
                        : p.waves == 6 ? k_trace_batch_units<6>
                                       : k_trace_batch_units<2>;
    hipLaunchKernelGGL(batch_kernel, dim3(p.grid), dim3(kPtBlock), 0, stream, d_items, d_begin, n, p.n_jobs, counter,
                       spill, p.lane_words, records, arrived, p.s_max);
  } else {
    auto pick = [&](auto fused) {
      constexpr bool kF = decltype(fused)::value;
      return p.waves == 3   ? k_trace_tiles<false, 3, kF>
             : p.waves == 4 ? k_trace_tiles<false, 4, kF>
             : p.waves == 5 ? k_trace_tiles<false, 5, kF>
             : p.waves == 6 ? k_trace_tiles<false, 6, kF>
                            : k_trace_tiles<false, 2, kF>;
    };
    auto kernel = p.fused ? pick(std::true_type{}) : pick(std::false_type{});
    hipLaunchKernelGGL(kernel, dim3(p.grid), dim3(kPtBlock), 0, stream, nullptr, d_items, d_begin, n, p.n_jobs, counter,
                       spill, p.lane_words);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, p.ring ? "k_trace_batch_units launch" : "k_trace_tiles launch");
  return LUCENT_OK;
}

size_t batch_table_bytes(uint32_t n_items, size_t* begin_off, size_t* items_off) {
  *begin_off = 256;
  *items_off = (*begin_off + size_t(n_items + 1) * sizeof(uint32_t) + 255) / 256 * 256;
  return *items_off + size_t(n_items) * sizeof(TileItem);
}

Status launch_trace_tiles(Context* ctx, TileItem* items, uint32_t n_items, const TileLaunch& cfg, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  if (n_items == 0) return LUCENT_OK;
  if (n_items == 1 && cfg.fused && cfg.units && !cfg.counter) {
    if (items[0].spp == 0) return LUCENT_OK;
    return launch_trace_units(ctx, items[0], cfg, stream);
  }
  if (n_items > 1) {
    if (cfg.counter) return fail(LUCENT_E_INVALID, "caller-owned tile buffers take one film");
    BatchPlan plan;
    Status s = plan_batch(ctx, items, n_items, cfg, &plan);
    if (s != LUCENT_OK) return s;
    for (uint32_t i = 0; i < n_items; ++i) items[i].tiles_x = plan.items[i].tiles_x;
    if (plan.n_jobs == 0) return LUCENT_OK;
    if ((s = ensure_spill(ctx, plan.spill_bytes)) != LUCENT_OK) return s;
    // Device table: [counter | job_begin (n+1) | items], 256-byte aligned parts.
    size_t begin_off = 0, items_off = 0;
    const size_t bytes = batch_table_bytes(n_items, &begin_off, &items_off);
    if ((s = ensure_tile_table(ctx, bytes)) != LUCENT_OK) return s;
    char* table = static_cast<char*>(ctx->tile_table);
    std::vector<char> host(bytes - begin_off);
    std::memcpy(host.data(), plan.job_begin.data(), plan.job_begin.size() * sizeof(uint32_t));
    std::memcpy(host.data() + (items_off - begin_off), plan.items.data(), size_t(n_items) * sizeof(TileItem));
    hipError_t e = hipMemcpyAsync(table + begin_off, host.data(), host.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(static_cast<hipEvent_t>(ctx->tile_event), stream);
    if (e == hipSuccess) e = hipEventSynchronize(static_cast<hipEvent_t>(ctx->tile_event));  // `host` must outlive the copy
    if (e != hipSuccess) return fail_hip(e, "tile table upload");
    float4* records = nullptr;
    uint32_t* arrived = nullptr;
    if (plan.ring) {
      const size_t counters_off = (plan.records_bytes + 255) / 256 * 256;
      if ((s = ensure_units(ctx, counters_off + plan.arrived_bytes)) != LUCENT_OK) return s;
      records = static_cast<float4*>(ctx->units);
      arrived = reinterpret_cast<uint32_t*>(static_cast<char*>(ctx->units) + counters_off);
    }
    return launch_batch_planned(plan, reinterpret_cast<uint32_t*>(table),
                                reinterpret_cast<const uint32_t*>(table + begin_off),
                                reinterpret_cast<const TileItem*>(table + items_off), static_cast<uint32_t*>(ctx->spill),
                                records, arrived, stream);
  }
  // One film, whole pixels: the context's buffers or the caller's (a
  // captured graph).
  TileItem& it = items[0];
  it.tiles_x = (it.frame.width + kTile - 1) / kTile;
  const uint32_t n_jobs = it.spp ? tile_jobs(it.frame.width, it.frame.height) : 0;
  if (n_jobs == 0) return LUCENT_OK;
  const uint32_t depth = std::max(1u, pt::pt_spill_words(it.scene.bvh));
  const uint32_t waves = tile_waves(cfg.occupancy);
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * waves;
  const uint32_t grid = std::min<uint32_t>(n_jobs, uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu);
  const uint32_t lane_words = tile_lane_words(depth);
  const size_t spill_bytes = size_t(grid) * kPtBlock * lane_words * sizeof(uint32_t);
  uint32_t* spill = cfg.spill;
  uint32_t* counter = cfg.counter;
  TileItem* d_single = cfg.item;
  Status s;
  if (cfg.counter) {
    if (cfg.spill_bytes < spill_bytes) return fail(LUCENT_E_INVALID, "caller-owned spill buffer too small");
    if (!cfg.item) return fail(LUCENT_E_INVALID, "caller-owned tile buffers need a film descriptor slot");
  } else {
    if ((s = ensure_spill(ctx, spill_bytes)) != LUCENT_OK) return s;
    spill = static_cast<uint32_t*>(ctx->spill);
    if ((s = ensure_tile_table(ctx, 256 + sizeof(TileItem))) != LUCENT_OK) return s;
    counter = static_cast<uint32_t*>(ctx->tile_table);
    d_single = reinterpret_cast<TileItem*>(static_cast<char*>(ctx->tile_table) + 256);
  }
  // The film's descriptor goes to device memory with the counter reset (a
  // kernel, so that a captured graph holds it), and the kernel reads it
  // through a pointer (fewer registers than a by-value argument).
  hipLaunchKernelGGL(k_begin_units, dim3(1), dim3(1), 0, stream, d_single, it, counter);
  auto launch_waves = [&](auto fused) {
    constexpr bool kF = decltype(fused)::value;
    auto kernel = waves == 3   ? k_trace_tiles<true, 3, kF>
                  : waves == 4 ? k_trace_tiles<true, 4, kF>
                  : waves == 5 ? k_trace_tiles<true, 5, kF>
                  : waves == 6 ? k_trace_tiles<true, 6, kF>
                               : k_trace_tiles<true, 2, kF>;
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kPtBlock), 0, stream, d_single, nullptr, nullptr, 1u, n_jobs, counter,
                       spill, lane_words);
  };
  if (cfg.fused)
    launch_waves(std::true_type{});
  else
    launch_waves(std::false_type{});
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_tiles launch");
  return LUCENT_OK;
}

Status launch_advance_seeds(uint32_t* seeds, const TileItem* d_items, uint32_t n_items, void* stream) {
  if (n_items == 0) return LUCENT_OK;
  hipLaunchKernelGGL(k_advance_seeds, dim3((n_items + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), seeds,
                     d_items, n_items);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_advance_seeds launch");
  return LUCENT_OK;
}

}  // namespace lucent
