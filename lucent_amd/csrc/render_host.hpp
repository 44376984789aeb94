// Host-side rendering helpers (lighting tables, camera) and launchers.
#pragma once

#include <cstdint>
#include <vector>

#include "internal.hpp"
#include "pt_types.hpp"

namespace lucent {

struct EnvSampler {
  uint32_t grid_w = 0, grid_h = 0;
  double total = 0;
  std::vector<float> table;  // 4 floats per cell
};

Status build_alias_table(const double* w, uint32_t n, float* prob, uint32_t* alias, double* pmf, double* total);
Status build_env_sampler(const float* rgb, uint32_t width, uint32_t height, uint32_t max_width, EnvSampler* out);
// The sampler grid of a width x height map (environment.ts:53-54) and the
// sin(theta) weight of grid row cy (environment.ts:58), shared by the host
// and device builds so that both see the same values.
void env_grid(uint32_t width, uint32_t height, uint32_t max_width, uint32_t* gw, uint32_t* gh);
double env_row_sine(uint32_t cy, uint32_t gh);
// buildEnvSampler on the device from an RGBA float map already in device
// memory; the table (4 floats per cell) is written to d_table.  Bit-identical
// to build_env_sampler on the same texels, or with `parallel` the alias
// pairs from FP64 prefix sums (the same distribution, other pairs).
// Synchronous on `stream`.
Status build_env_sampler_gpu(Context* ctx, const float* d_rgba, uint32_t width, uint32_t height, uint32_t max_width,
                             float* d_table, uint32_t* grid_w, uint32_t* grid_h, double* total, void* stream,
                             float* ms_out, bool parallel = false);
void bake_sky(const lucent_sky& sky, const lucent_sun& sun, uint32_t width, std::vector<float>* out);
double sheen_albedo(double cos_v, double roughness, uint32_t steps);
void sheen_lut(uint32_t size, uint32_t steps, float* out);
float round_to_half(float f);
void camera_basis(const lucent_camera& cam, double fwd[3], double right[3], double up[3], double* tan_half_fov);

Status launch_trace_sample(Context* ctx, const PtScene& sc, const PtFrame& f, const PtFilm& film, void* stream);
// One reprojecting sample of every pixel of `film` (seed f.seed); film and prev.film must not alias.
Status launch_trace_reprojected(Context* ctx, const PtScene& sc, const PtFrame& f, const PtFilm& film,
                                const PrevView& prev, void* stream);

// Launch shape of the persistent tile kernel (0 = default).
struct TileLaunch {
  uint32_t waves_per_cu = 0;  // persistent waves per CU (0: 4 x occupancy)
  uint32_t occupancy = 0;     // register-allocation target in waves per SIMD: 2..6 (0: 4)
  // Per-sample hand-out shape (TileItem): tile of a job, shading batch and
  // traversal turns per check; 0 = 32, 8 for the last two.  A tile of 0 x 0
  // is sized per film to its sample count (launch_job_tile): about
  // job_paths paths per job (0 = kJobPaths).
  uint32_t tile_w = 0, tile_h = 0, shade_batch = 0, trace_turns = 0;
  uint32_t job_paths = 0;
  // Optional caller-owned job counter, film descriptor slot and spill buffer
  // (one film only); otherwise the context's buffers are used.
  uint32_t* counter = nullptr;
  TileItem* item = nullptr;
  uint32_t* spill = nullptr;
  size_t spill_bytes = 0;
  // All rays of a lane through one traversal loop (false: each shadow ray
  // traced where the integrator asks for it).  Only for scenes whose
  // traversal draws no random numbers (no alpha-blended material).
  bool fused = true;
  // One film, fused: hand out pixel samples rather than whole pixels (the
  // context's buffers only; launches of up to the context's record budget
  // of path records, lucent_set_record_budget).
  bool units = true;
};
// Spill bytes the launch of one film with `cfg` needs.
size_t tile_spill_bytes(const Context* ctx, const TileItem& item, const TileLaunch& cfg);
// *p += v on the device (a graph node).
Status launch_add_u32(uint32_t* p, uint32_t v, void* stream);
uint32_t tile_jobs(uint32_t width, uint32_t height);
// Paths per job a sized job tile aims at when the launch names none.
constexpr uint32_t kJobPaths = 256;
// Job tile of a film traced with `spp` samples under `cfg`: its explicit
// tile_w x tile_h (a 0 beside a non-zero side is 8), or for 0 x 0 the
// largest tile of the sequence 8x8, 8x4, 4x4, 4x2, 2x2, 2x1, 1x1 whose
// pixels times spp stay within job_paths (at least one pixel).
void launch_job_tile(const TileLaunch& cfg, uint32_t spp, uint32_t* tile_w, uint32_t* tile_h);
// Checks a launch shape (tile of at most 64 pixels, shade batch 1..64,
// trace turns 1..64; 0 = default) and writes it into `item`.
Status apply_launch_shape(const TileLaunch& cfg, TileItem* item);
// Runs items[i].spp samples of every pixel of each item's film, seeds from
// items[i].frame.seed; fills items[i].tiles_x.
Status launch_trace_tiles(Context* ctx, TileItem* items, uint32_t n_items, const TileLaunch& cfg, void* stream);

// A launch over many films (a batch) as data: the film descriptors as the
// kernel reads them, the job table (film i owns jobs [job_begin[i],
// job_begin[i+1])), the grid, and the sizes of the buffers it needs.
// lucent_trace_batch runs it on the context's buffers, a batch graph on
// its own.
struct BatchPlan {
  std::vector<TileItem> items;
  std::vector<uint32_t> job_begin;
  uint32_t n_jobs = 0, grid = 0, waves = 4, lane_words = 0, s_max = 0;
  bool fused = true;
  bool ring = false;  // per-sample hand-out with per-wave record rings, else whole pixels
  size_t spill_bytes = 0, records_bytes = 0, arrived_bytes = 0;
};
Status plan_batch(const Context* ctx, const TileItem* items, uint32_t n_items, const TileLaunch& cfg, BatchPlan* p);
// The device table of a batch: [job counter | job_begin | items], 256-byte
// aligned parts; returns its size.
size_t batch_table_bytes(uint32_t n_items, size_t* begin_off, size_t* items_off);
// Resets the job counter and launches a planned batch whose job table and
// descriptors are in device memory (kernels only: capturable in a graph).
Status launch_batch_planned(const BatchPlan& p, uint32_t* counter, const uint32_t* d_begin, const TileItem* d_items,
                            uint32_t* spill, void* records, uint32_t* arrived, void* stream);
// seeds[i] += d_items[i].spp on the device.
Status launch_advance_seeds(uint32_t* seeds, const TileItem* d_items, uint32_t n_items, void* stream);

}  // namespace lucent
