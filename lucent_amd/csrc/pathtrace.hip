// Progressive path tracing on gfx950.  The integrator is pt_device.hpp
// (trace.frag semantics in float32); two launch structures run it:
//
// * k_trace_sample — the reference's structure (Renderer.traceSamples,
//   src/renderer.ts:702-737): one launch per sample over the whole film, one
//   lane per pixel, each wave64 an 8x8 pixel tile so that neighbouring rays
//   share BVH nodes in the caches.  A wave lasts as long as its longest path.
//
// * k_trace_tiles — persistent waves pull 8x8 tile jobs from an atomic queue
//   (one or many films / scenes per launch) and run ALL of a job's samples:
//   each lane owns one pixel, advances its path one bounce at a time and, as
//   soon as the path ends, folds it into the pixel's running mean and
//   starts the pixel's next sample.  Lanes are refilled instead of idling
//   behind the longest path of the wave.  Per pixel the samples run in seed
//   order through the same accumulation step, so the film is bit-identical
//   to spp k_trace_sample launches.  The running means stay in the film
//   (72 B of traffic per sample) rather than in registers: registers, not
//   bandwidth, bound this kernel's occupancy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "pt_device.hpp"
#include "render_host.hpp"

namespace lucent {
namespace {

constexpr int kPtBlock = 64;
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
  pt::PtStack<kPtBlock> st{lds + threadIdx.x, spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
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
  pt::PtStack<kPtBlock> st{lds + threadIdx.x, spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
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

template <class StackT>
__device__ __forceinline__ void trace_pixel_samples(const TileItem& it, uint32_t x, uint32_t y, StackT& st,
                                                    uint32_t seed0) {
  const PtScene& sc = it.scene;
  const PtFrame& f = it.frame;
  const uint32_t idx = y * f.width + x;
  uint32_t k = 0;
  pt::Rng rng = pt::seed_rng(idx, seed0);
  pt::v3 ro, rd;
  pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
  pt::PathState p;
  pt::path_begin(p, ro, rd);
  for (;;) {
    if (!pt::path_step(sc, f, p, rng, st)) {
      if (it.sample_radiance)
        reinterpret_cast<float4*>(it.sample_radiance)[size_t(k) * f.width * f.height + idx] =
            make_float4(p.radiance.x, p.radiance.y, p.radiance.z, 0.0f);
      pt::accumulate(it.film, idx, p.radiance, p.guide);
      if (++k >= it.spp) break;
      rng = pt::seed_rng(idx, seed0 + k);
      pt::camera_ray(f, float(x) + 0.5f, float(y) + 0.5f, rng, ro, rd);
      pt::path_begin(p, ro, rd);
    }
  }
}

template <class StackT>
__device__ __forceinline__ void run_tile(const TileItem& it, uint32_t tile, StackT& st) {
  // The first seed, read once per job in uniform control flow (a scalar).
  const uint32_t seed0 = it.seed_src ? __builtin_amdgcn_readfirstlane(*it.seed_src) : it.frame.seed;
  const uint32_t x = (tile % it.tiles_x) * kTile + (threadIdx.x & (kTile - 1));
  const uint32_t y = (tile / it.tiles_x) * kTile + (threadIdx.x / kTile);
  if (x < it.frame.width && y < it.frame.height) trace_pixel_samples(it, x, y, st, seed0);
}

// Jobs are numbered film by film: film i owns jobs [job_begin[i], job_begin[i+1]).
// kSingle: one film, passed by value (its fields live in the kernel arguments).
// kWaves: the occupancy the register allocation targets (waves per SIMD);
// more waves hide more memory latency at the price of register spills.
template <bool kSingle, int kWaves>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_trace_tiles(TileItem single, const TileItem* __restrict__ items,
                                                         const uint32_t* __restrict__ job_begin, uint32_t n_items,
                                                         uint32_t n_jobs, uint32_t* __restrict__ counter,
                                                         uint32_t* spill, uint32_t spill_depth) {
  __shared__ uint32_t lds[pt::kPtStackWords * kPtBlock];
  __shared__ uint32_t next_job;
  pt::PtStack<kPtBlock> st{lds + threadIdx.x, spill + size_t(blockIdx.x * kPtBlock + threadIdx.x) * spill_depth,
                                 0, 0};
  for (;;) {
    if (threadIdx.x == 0) next_job = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t job = __builtin_amdgcn_readfirstlane(next_job);
    __syncthreads();
    if (job >= n_jobs) break;
    uint32_t item = 0;
    if (!kSingle) {
      uint32_t hi = n_items;
      while (hi - item > 1) {
        const uint32_t mid = (item + hi) / 2;
        if (job_begin[mid] <= job)
          item = mid;
        else
          hi = mid;
      }
    }
    if constexpr (kSingle)
      run_tile(single, job, st);
    else
      run_tile(items[item], job - job_begin[item], st);
  }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }
__global__ void k_add_u32(uint32_t* p, uint32_t v) { *p += v; }

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

size_t tile_spill_bytes(const Context* ctx, const TileItem& item, const TileLaunch& cfg) {
  const uint32_t waves = (cfg.occupancy >= 2 && cfg.occupancy <= 6) ? cfg.occupancy : 2;
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * waves;
  const uint32_t n_jobs = item.spp ? tile_jobs(item.frame.width, item.frame.height) : 0;
  const uint32_t grid = std::min<uint32_t>(n_jobs, uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu);
  return size_t(grid) * kPtBlock * std::max<uint32_t>(1, pt::pt_spill_words(item.scene.bvh)) * sizeof(uint32_t);
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

Status launch_trace_tiles(Context* ctx, TileItem* items, uint32_t n_items, const TileLaunch& cfg, void* stream_) {
  hipStream_t stream = static_cast<hipStream_t>(stream_);
  if (n_items == 0) return LUCENT_OK;
  std::vector<uint32_t> job_begin(n_items + 1, 0);
  uint32_t depth = 1;
  for (uint32_t i = 0; i < n_items; ++i) {
    items[i].tiles_x = (items[i].frame.width + kTile - 1) / kTile;
    const uint32_t jobs = items[i].spp ? tile_jobs(items[i].frame.width, items[i].frame.height) : 0;
    job_begin[i + 1] = job_begin[i] + jobs;
    depth = std::max(depth, pt::pt_spill_words(items[i].scene.bvh));
  }
  const uint32_t n_jobs = job_begin[n_items];
  if (n_jobs == 0) return LUCENT_OK;
  const uint32_t waves = (cfg.occupancy >= 2 && cfg.occupancy <= 6) ? cfg.occupancy : 2;
  const uint32_t waves_per_cu = cfg.waves_per_cu ? cfg.waves_per_cu : 4 * waves;
  const uint32_t grid = std::min<uint32_t>(n_jobs, uint32_t(std::max(ctx->compute_units, 1)) * waves_per_cu);
  const size_t spill_bytes = size_t(grid) * kPtBlock * depth * sizeof(uint32_t);
  uint32_t* spill = cfg.spill;
  uint32_t* counter = cfg.counter;
  Status s;
  if (cfg.counter) {
    // Caller-owned buffers (a captured graph): one film, nothing to upload.
    if (n_items != 1) return fail(LUCENT_E_INVALID, "caller-owned tile buffers take one film");
    if (cfg.spill_bytes < spill_bytes) return fail(LUCENT_E_INVALID, "caller-owned spill buffer too small");
  } else {
    if ((s = ensure_spill(ctx, spill_bytes)) != LUCENT_OK) return s;
    spill = static_cast<uint32_t*>(ctx->spill);
  }
  // Device table: [counter | job_begin (n+1) | items], 256-byte aligned parts.
  const size_t begin_off = 256;
  const size_t items_off = (begin_off + job_begin.size() * sizeof(uint32_t) + 255) / 256 * 256;
  const size_t bytes = items_off + (n_items > 1 ? size_t(n_items) * sizeof(TileItem) : 0);
  char* table = nullptr;
  if (!cfg.counter) {
    if ((s = ensure_tile_table(ctx, bytes)) != LUCENT_OK) return s;
    table = static_cast<char*>(ctx->tile_table);
    counter = reinterpret_cast<uint32_t*>(table);
  }
  if (n_items > 1) {
    std::vector<char> host(bytes - begin_off);
    std::memcpy(host.data(), job_begin.data(), job_begin.size() * sizeof(uint32_t));
    std::memcpy(host.data() + (items_off - begin_off), items, size_t(n_items) * sizeof(TileItem));
    hipError_t e = hipMemcpyAsync(table + begin_off, host.data(), host.size(), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(static_cast<hipEvent_t>(ctx->tile_event), stream);
    if (e == hipSuccess) e = hipEventSynchronize(static_cast<hipEvent_t>(ctx->tile_event));  // `host` must outlive the copy
    if (e != hipSuccess) return fail_hip(e, "tile table upload");
  }
  hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, stream, counter, 0u);
  const TileItem* d_items = n_items > 1 ? reinterpret_cast<const TileItem*>(table + items_off) : nullptr;
  const uint32_t* d_begin = n_items > 1 ? reinterpret_cast<const uint32_t*>(table + begin_off) : nullptr;
  auto launch = [&](auto single_kernel, auto multi_kernel) {
    if (n_items == 1)
      hipLaunchKernelGGL(single_kernel, dim3(grid), dim3(kPtBlock), 0, stream, items[0], nullptr, nullptr, 1u, n_jobs,
                         counter, spill, depth);
    else
      hipLaunchKernelGGL(multi_kernel, dim3(grid), dim3(kPtBlock), 0, stream, TileItem{}, d_items, d_begin, n_items,
                         n_jobs, counter, spill, depth);
  };
  switch (waves) {
    case 3:
      launch(k_trace_tiles<true, 3>, k_trace_tiles<false, 3>);
      break;
    case 4:
      launch(k_trace_tiles<true, 4>, k_trace_tiles<false, 4>);
      break;
    case 5:
      launch(k_trace_tiles<true, 5>, k_trace_tiles<false, 5>);
      break;
    case 6:
      launch(k_trace_tiles<true, 6>, k_trace_tiles<false, 6>);
      break;
    default:
      launch(k_trace_tiles<true, 2>, k_trace_tiles<false, 2>);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(e, "k_trace_tiles launch");
  return LUCENT_OK;
}

}  // namespace lucent
