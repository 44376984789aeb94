// Wavefront progressive path tracing on gfx950.
//
// The megakernels (pathtrace.hip) run a whole path step per lane: closest-hit
// traversal, shading, shadow traversals.  Their register budget is set by the
// shading code (2-4 waves per SIMD), and a wave lasts as long as its slowest
// lane in every traversal.  Here a path step is split into kernels over a
// queue of live paths, one path slot per pixel (so a pixel's samples still
// run one after the other, in seed order):
//
//   extend  - persistent waves take queued paths and find the closest hit of
//             each path's ray (spheres, ground, then the BVH), one node visit
//             or one triangle test per loop trip; a lane whose ray is done
//             takes the next queued path as soon as enough lanes of its wave
//             are idle.  Few registers, high occupancy, full waves.
//   shadow  - the same for the up to three shadow rays (sun, environment,
//             light) each queued path's previous vertex produced.
//   shade   - the path step itself (pt_device.hpp's path_step_with, shared with
//             the megakernels): adds the previous vertex's unoccluded light
//             samples, then the hit's emission, the light-sample contributions
//             of the new vertex (deferred to the next shadow pass), BSDF
//             sampling and roulette.  A finished path is folded into the film
//             and the pixel's next sample starts at once.
//
// The traversal visits nodes and triangles in exactly the megakernel's order
// and the shading runs the same code in the same random-number order, so the
// film is bit-identical to the megakernels'.  Alpha-blended materials draw
// random numbers inside traversal and keep the megakernel path.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pt_device.hpp"
#include "render_host.hpp"
#include "wavefront.hpp"

namespace lucent {
namespace {

using namespace pt;

constexpr int kWfBlock = 64;
constexpr int kTile = 8;
constexpr uint32_t kFlagSpecular = 1u, kFlagGuideDone = 2u, kFlagTerminating = 4u;
constexpr int kBounceShift = 8;
// A wave refills its idle lanes once at least this many are idle.
constexpr int kRefill = 16;

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }
__device__ inline uint64_t lanes_below() { return (uint64_t(1) << lane_id()) - 1u; }

// Wave-wide work fetch: `n` consecutive items for the lanes set in `want`;
// each of those lanes gets base + (its rank among them).  Returns base.
__device__ inline uint32_t wave_fetch(uint32_t* cursor, uint64_t want) {
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(cursor, uint32_t(__popcll(want)));
  return __shfl(base, 0);
}

// Slot -> pixel: slots enumerate 8x8 tiles, 64 slots per tile.
__device__ inline bool slot_pixel(const PtFrame& f, uint32_t tiles_x, uint32_t slot, uint32_t& x, uint32_t& y) {
  const uint32_t tile = slot >> 6, l = slot & 63u;
  x = (tile % tiles_x) * kTile + (l & (kTile - 1));
  y = (tile / tiles_x) * kTile + (l / kTile);
  return x < f.width && y < f.height;
}

// ---------------------------------------------------------------------------
// Path state, structure of arrays (one entry per slot).
// ---------------------------------------------------------------------------
__device__ inline void load_state(const WfBuffers& B, uint32_t slot, PathState& p, uint32_t& flags) {
  const float4 o = B.ray_o[slot], d = B.ray_d[slot], be = B.beta[slot], ra = B.rad[slot], ti = B.tint[slot],
               ga = B.galb[slot], gn = B.gnrm[slot];
  flags = B.flags[slot];
  p.ro = mk(o.x, o.y, o.z);
  p.rd = mk(d.x, d.y, d.z);
  p.bsdf_pdf = d.w;
  p.throughput = mk(be.x, be.y, be.z);
  p.radiance = mk(ra.x, ra.y, ra.z);
  p.tint = mk(ti.x, ti.y, ti.z);
  p.guide.depth = ti.w;
  p.guide.albedo = mk(ga.x, ga.y, ga.z);
  p.guide.coverage = ga.w;
  p.guide.normal = mk(gn.x, gn.y, gn.z);
  p.guide.done = (flags & kFlagGuideDone) != 0;
  p.specular = (flags & kFlagSpecular) != 0;
  p.bounce = int(flags >> kBounceShift);
}

__device__ inline void store_state(const WfBuffers& B, uint32_t slot, const PathState& p, uint32_t extra_flags) {
  B.ray_o[slot] = make_float4(p.ro.x, p.ro.y, p.ro.z, 0.0f);
  B.ray_d[slot] = make_float4(p.rd.x, p.rd.y, p.rd.z, p.bsdf_pdf);
  B.beta[slot] = make_float4(p.throughput.x, p.throughput.y, p.throughput.z, 0.0f);
  B.rad[slot] = make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f);
  B.tint[slot] = make_float4(p.tint.x, p.tint.y, p.tint.z, p.guide.depth);
  B.galb[slot] = make_float4(p.guide.albedo.x, p.guide.albedo.y, p.guide.albedo.z, p.guide.coverage);
  B.gnrm[slot] = make_float4(p.guide.normal.x, p.guide.normal.y, p.guide.normal.z, 0.0f);
  B.flags[slot] = (p.specular ? kFlagSpecular : 0u) | (p.guide.done ? kFlagGuideDone : 0u) | extra_flags |
                  (uint32_t(p.bounce) << kBounceShift);
}

__device__ inline void clear_nee(const WfBuffers& B, uint32_t slot) {
  for (int j = 0; j < 3; ++j) B.nee[slot * 3 + j] = make_float4(0.f, 0.f, 0.f, -1.0f);
}

// Starts sample k of the slot's pixel: camera ray, fresh path state.
__device__ inline void begin_sample(const PtFrame& f, const WfBuffers& B, uint32_t slot, uint32_t x, uint32_t y,
                                    uint32_t k) {
  const uint32_t idx = y * f.width + x;
  Rng rng = seed_rng(idx, f.seed + k);
  v3 ro, rd;
  camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  PathState p;
  path_begin(p, ro, rd);
  store_state(B, slot, p, 0u);
  B.rng[slot] = rng.s;
  B.sample[slot] = k;
  clear_nee(B, slot);
}

// The shade pass's tracer: the closest hit comes from the extend pass; light
// samples are stored for the next shadow pass.
struct WfTracer {
  static constexpr bool kRecordPrimary = false;
  const PtScene& sc;
  const WfBuffers& B;
  uint32_t slot;
  HitRecord rec;
  uint32_t mask;  // light samples queued (bit `which`)
  __device__ __forceinline__ bool closest(v3 ro, v3 rd, Hit& hit, Rng&) { return build_hit(sc, ro, rd, rec, hit); }
  template <class F>
  __device__ __forceinline__ void nee(int which, v3 o, v3 d, float tmax, Rng&, PathState&, F contribution) {
    const v3 c = contribution();
    const uint32_t e = slot * 3 + uint32_t(which);
    B.nee[e] = make_float4(c.x, c.y, c.z, 0.0f);
    B.shadow_o[e] = make_float4(o.x, o.y, o.z, tmax);
    B.shadow_d[e] = make_float4(d.x, d.y, d.z, 0.0f);
    mask |= 1u << which;
  }
};

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__global__ void k_wf_init(WfCtl* ctl, uint32_t n_slots) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctl->parity = 1;
    ctl->count[0] = ctl->count[1] = 0;
    ctl->scount[0] = ctl->scount[1] = 0;
    ctl->cursor[0] = ctl->cursor[1] = ctl->cursor[2] = 0;
    ctl->live = 0;
    ctl->slots = n_slots;
  }
}

// Every valid slot starts sample 0 and joins queue 0.
__global__ __launch_bounds__(kWfBlock) void k_wf_raygen(TileItem it, WfBuffers B) {
  const uint32_t slot = blockIdx.x * kWfBlock + threadIdx.x;
  uint32_t x = 0, y = 0;
  const bool ok = slot < B.ctl->slots && slot_pixel(it.frame, it.tiles_x, slot, x, y) && it.spp > 0;
  if (ok) begin_sample(it.frame, B, slot, x, y, 0);
  const uint64_t m = __ballot(ok);
  if (m == 0) return;
  const uint32_t base = wave_fetch(&B.ctl->count[0], m);
  if (ok) B.queue[0][base + __popcll(m & lanes_below())] = slot;
}

// Start of an iteration: flip the queues, clear the counters the shade pass fills.
__global__ void k_wf_advance(WfCtl* ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const uint32_t p = ctl->parity ^ 1u;
    ctl->parity = p;
    ctl->count[p ^ 1u] = 0;
    ctl->scount[p ^ 1u] = 0;
    ctl->cursor[0] = ctl->cursor[1] = ctl->cursor[2] = 0;
    ctl->live = ctl->count[p];
  }
}

template <bool kShadow>
__global__ __launch_bounds__(kWfBlock) void k_wf_trace(TileItem it, WfBuffers B) {
  __shared__ uint32_t lds[kPtStackWords * kWfBlock];
  const PtScene& sc = it.scene;
  PtStack<kWfBlock> st{as_lds(lds + threadIdx.x), B.spill + size_t(blockIdx.x * kWfBlock + threadIdx.x) * B.spill_depth,
                             0, 0};
  const uint32_t p = __builtin_amdgcn_readfirstlane(B.ctl->parity);
  const uint32_t items = __builtin_amdgcn_readfirstlane(kShadow ? B.ctl->scount[p] : B.ctl->count[p]);
  const uint32_t* queue = kShadow ? B.squeue[p] : B.queue[p];
  uint32_t* cursor = &B.ctl->cursor[kShadow ? 1 : 0];
  Rng rng{0};  // alpha-blended materials never reach this kernel
  Trav4 t;
  HitRecord rec;
  uint32_t item = 0;
  bool active = false, exhausted = false;
  for (;;) {
    const uint64_t idle = __ballot(!active);
    if (!exhausted && __popcll(idle) >= kRefill) {
      const uint32_t base = wave_fetch(cursor, idle);
      if (!active) {
        const uint32_t i = base + uint32_t(__popcll(idle & lanes_below()));
        if (i < items) {
          if (kShadow) {
            const uint32_t e = queue[i];
            item = e;
            const float4 o = B.shadow_o[e], d = B.shadow_d[e];
            const v3 ro = mk(o.x, o.y, o.z), rd = mk(d.x, d.y, d.z);
            bool blocked = false;
            for (uint32_t k = 0; k < sc.sphere_count && !blocked; ++k)
              blocked = hit_sphere(ro, rd, reinterpret_cast<const float4*>(sc.spheres)[k], o.w) < o.w;
            if (!blocked) blocked = hit_ground(sc, ro, rd, o.w) < o.w;
            if (!blocked) {
              trav4_begin(sc, t, ro, rd, o.w, st);
              active = true;
            }
          } else {
            const uint32_t slot = queue[i];
            if ((B.flags[slot] & kFlagTerminating) == 0) {
              item = slot;
              const float4 o = B.ray_o[slot], d = B.ray_d[slot];
              const v3 ro = mk(o.x, o.y, o.z), rd = mk(d.x, d.y, d.z);
              search_analytic(sc, ro, rd, rec);
              trav4_begin(sc, t, ro, rd, rec.t, st);
              active = true;
            }
          }
        }
      }
      if (base + uint32_t(__popcll(idle)) >= items) exhausted = true;
    }
    if (!__any(active)) {
      if (exhausted) break;
      continue;
    }
    if (active) {
      // A scene without triangles starts its rays done: nothing to pop.
      const bool blocked = !trav4_done(t, st) && trav4_step<kShadow>(sc, t, st, rng);
      if (kShadow) {
        if (blocked || trav4_done(t, st)) {
          if (!blocked) B.nee[item].w = 1.0f;
          active = false;
        }
      } else if (trav4_done(t, st)) {
        HitRecord r = rec;
        if (t.best >= 0) {
          r.t = t.tmax;
          r.kind = kHitTri;
          r.prim = uint32_t(t.best);
          r.u = t.bu;
          r.v = t.bv;
        }
        B.hit[item] = make_float4(r.t, r.u, r.v, __uint_as_float((r.kind << 30) | r.prim));
        active = false;
      }
    }
  }
  // Shadow rays blocked by a sphere or the ground never started: they stay pending (0), not visible.
}

__device__ inline void shade_slot(const TileItem& it, const WfBuffers& B, uint32_t slot, bool& enqueue,
                                  uint32_t& shadows) {
  const PtScene& sc = it.scene;
  const PtFrame& f = it.frame;
  uint32_t x = 0, y = 0;
  slot_pixel(f, it.tiles_x, slot, x, y);
  PathState p;
  uint32_t flags;
  load_state(B, slot, p, flags);
  Rng rng{B.rng[slot]};
  // The previous vertex's light samples that reached their light, in order.
  for (int j = 0; j < 3; ++j) {
    const float4 e = B.nee[slot * 3 + j];
    if (e.w == 1.0f) p.radiance = p.radiance + mk(e.x, e.y, e.z);
  }
  bool ended = (flags & kFlagTerminating) != 0;
  if (!ended) {
    clear_nee(B, slot);
    const float4 h = B.hit[slot];
    const uint32_t bits = __float_as_uint(h.w);
    WfTracer tr{sc, B, slot, HitRecord{h.x, h.y, h.z, bits >> 30, bits & 0x3FFFFFFFu}, 0u};
    const bool more = path_step_with(sc, f, p, rng, tr);
    shadows = tr.mask;
    if (more) {
      store_state(B, slot, p, 0u);
      B.rng[slot] = rng.s;
      enqueue = true;
      return;
    }
    if (tr.mask != 0) {  // fold after the shadow pass has resolved this vertex's light samples
      store_state(B, slot, p, kFlagTerminating);
      B.rng[slot] = rng.s;
      enqueue = true;
      return;
    }
  }
  accumulate(it.film, y * f.width + x, p.radiance, p.guide);
  const uint32_t k = B.sample[slot] + 1;
  if (k < it.spp) {
    begin_sample(f, B, slot, x, y, k);
    enqueue = true;
  } else {
    clear_nee(B, slot);
  }
}

__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_wf_shade(TileItem it,
                                                                                              WfBuffers B) {
  const uint32_t p = __builtin_amdgcn_readfirstlane(B.ctl->parity);
  const uint32_t count = __builtin_amdgcn_readfirstlane(B.ctl->count[p]);
  const uint32_t* queue = B.queue[p];
  uint32_t* out = B.queue[p ^ 1u];
  for (;;) {
    const uint32_t base = wave_fetch(&B.ctl->cursor[2], ~uint64_t(0));
    if (base >= count) break;
    const uint32_t i = base + lane_id();
    bool enq = false;
    uint32_t slot = 0, shadows = 0;
    if (i < count) {
      slot = queue[i];
      shade_slot(it, B, slot, enq, shadows);
    }
    const uint64_t m = __ballot(enq);
    if (m != 0) {
      const uint32_t ob = wave_fetch(&B.ctl->count[p ^ 1u], m);
      if (enq) out[ob + __popcll(m & lanes_below())] = slot;
    }
    // The new light samples, compacted into the next shadow queue.
    const uint64_t m0 = __ballot(shadows & 1u), m1 = __ballot(shadows & 2u), m2 = __ballot(shadows & 4u);
    const uint32_t n0 = __popcll(m0), n1 = __popcll(m1), n2 = __popcll(m2);
    if (n0 + n1 + n2 != 0) {
      uint32_t sb = 0;
      if (lane_id() == 0) sb = atomicAdd(&B.ctl->scount[p ^ 1u], n0 + n1 + n2);
      sb = __shfl(sb, 0);
      uint32_t* sq = B.squeue[p ^ 1u];
      const uint64_t below = lanes_below();
      if (shadows & 1u) sq[sb + __popcll(m0 & below)] = slot * 3 + 0;
      if (shadows & 2u) sq[sb + n0 + __popcll(m1 & below)] = slot * 3 + 1;
      if (shadows & 4u) sq[sb + n0 + n1 + __popcll(m2 & below)] = slot * 3 + 2;
    }
  }
}

}  // namespace

size_t wavefront_bytes(uint32_t slots) {
  // 8 float4 of path state + 1 hit + 3 light samples x (contribution, origin, direction),
  // 3 u32 (rng, sample, flags), 2 path queues, 2 shadow queues (3 per slot).
  return size_t(slots) * (16 * 8 + 16 + 3 * 48 + 12 + 8 + 24) + 4096 + 1024;
}

Status launch_wavefront(Context* ctx, TileItem& it, const WavefrontLaunch& cfg, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  it.tiles_x = (it.frame.width + kTile - 1) / kTile;
  const uint32_t tiles_y = (it.frame.height + kTile - 1) / kTile;
  const uint32_t slots = it.tiles_x * tiles_y * 64;
  if (slots == 0 || it.spp == 0) return LUCENT_OK;
  Status s = ensure_wavefront(ctx, wavefront_bytes(slots));
  if (s != LUCENT_OK) return s;
  // Persistent widths: as many waves per CU as each pass's registers allow.
  const uint32_t cus = uint32_t(std::max(ctx->compute_units, 1));
  const uint32_t w_extend = cus * (cfg.waves_per_cu ? cfg.waves_per_cu : 20);
  const uint32_t w_shadow = cus * (cfg.waves_per_cu ? cfg.waves_per_cu : 24);
  const uint32_t w_shade = cus * (cfg.waves_per_cu ? cfg.waves_per_cu : 8);
  const uint32_t waves = std::max(w_extend, w_shadow);
  const uint32_t depth = pt_spill_words(it.scene.bvh);
  if ((s = ensure_spill(ctx, size_t(waves) * kWfBlock * depth * sizeof(uint32_t))) != LUCENT_OK) return s;
  // Carve the buffers.
  char* base = static_cast<char*>(ctx->wavefront);
  auto take = [&](size_t bytes) {
    char* p = base;
    base += (bytes + 255) / 256 * 256;
    return p;
  };
  WfBuffers B;
  B.ctl = reinterpret_cast<WfCtl*>(take(sizeof(WfCtl)));
  B.ray_o = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.ray_d = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.beta = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.rad = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.tint = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.galb = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.gnrm = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.hit = reinterpret_cast<float4*>(take(size_t(slots) * 16));
  B.nee = reinterpret_cast<float4*>(take(size_t(slots) * 48));
  B.shadow_o = reinterpret_cast<float4*>(take(size_t(slots) * 48));
  B.shadow_d = reinterpret_cast<float4*>(take(size_t(slots) * 48));
  B.rng = reinterpret_cast<uint32_t*>(take(size_t(slots) * 4));
  B.sample = reinterpret_cast<uint32_t*>(take(size_t(slots) * 4));
  B.flags = reinterpret_cast<uint32_t*>(take(size_t(slots) * 4));
  B.queue[0] = reinterpret_cast<uint32_t*>(take(size_t(slots) * 4));
  B.queue[1] = reinterpret_cast<uint32_t*>(take(size_t(slots) * 4));
  B.squeue[0] = reinterpret_cast<uint32_t*>(take(size_t(slots) * 12));
  B.squeue[1] = reinterpret_cast<uint32_t*>(take(size_t(slots) * 12));
  B.spill = static_cast<uint32_t*>(ctx->spill);
  B.spill_depth = depth;
  if (size_t(base - static_cast<char*>(ctx->wavefront)) > ctx->wavefront_bytes)
    return fail(LUCENT_E_INVALID, "wavefront buffer layout exceeds its allocation");

  hipLaunchKernelGGL(k_wf_init, dim3(1), dim3(64), 0, stream, B.ctl, slots);
  hipLaunchKernelGGL(k_wf_raygen, dim3((slots + kWfBlock - 1) / kWfBlock), dim3(kWfBlock), 0, stream, it, B);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "wavefront set-up launch");
  // Iterations until no path is left.  Queues never grow (a queued path
  // queues at most itself), so the live count read every few iterations
  // bounds the work of the next ones and sizes their grids: the tail, where
  // few pixels still trace, launches few waves.  A path needs at most
  // max_bounces + 2 shade passes per sample, which bounds the loop.
  const uint64_t bound = uint64_t(it.spp) * (it.frame.max_bounces + 2) + 2;
  uint32_t live = slots;
  uint64_t iter = 0;
  while (live != 0) {
    const uint32_t g = (live + kWfBlock - 1) / kWfBlock;
    const uint32_t ge = std::min(w_extend, g), gs = std::min(w_shadow, 3 * g), gh = std::min(w_shade, g);
    const uint32_t check = cfg.check_every ? cfg.check_every : (live > slots / 8 ? 16u : 4u);
    for (uint32_t k = 0; k < check; ++k, ++iter) {
      hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(64), 0, stream, B.ctl);
      hipLaunchKernelGGL(k_wf_trace<false>, dim3(ge), dim3(kWfBlock), 0, stream, it, B);
      hipLaunchKernelGGL(k_wf_trace<true>, dim3(gs), dim3(kWfBlock), 0, stream, it, B);
      hipLaunchKernelGGL(k_wf_shade, dim3(gh), dim3(kWfBlock), 0, stream, it, B);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e, "wavefront launch");
    WfCtl host;
    if ((e = hipMemcpyAsync(&host, B.ctl, sizeof(WfCtl), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
        (e = hipStreamSynchronize(stream)) != hipSuccess)
      return fail_hip(e, "wavefront progress read");
    live = host.count[host.parity ^ 1u];  // paths queued for the next iteration
    if (iter > bound + 16) return fail(LUCENT_E_OVERFLOW, "wavefront did not drain within its iteration bound");
  }
  return LUCENT_OK;
}

}  // namespace lucent
