// Internal declarations shared by the host C++ and the HIP translation units.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lucent_hip.h"

namespace lucent {

using Status = lucent_status;

// Records `msg` as the calling thread's last error and returns `code`.
Status fail(Status code, const std::string& msg);
// Same for a failed HIP call; `what` names the call.
Status fail_hip(int hip_err, const char* what);
void clear_error();

Status build_bvh(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins, float* nodes_out,
                 uint32_t* node_count, uint32_t* tri_indices_out, double bounds[6]);
// The same splits built with subtrees on `threads` host threads (node
// numbering differs from build_bvh's; deterministic).
Status build_bvh_parallel(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                          float* nodes_out, uint32_t* node_count, uint32_t* tri_indices_out, double bounds[6],
                          uint32_t threads);
struct Context;
// Search radius of the clustered GPU build (neighbours each side in Morton
// order) unless the caller names one.
constexpr uint32_t kClusterRadius = 16;
// radius 0: the linear BVH (radix tree of the sorted Morton codes); > 0
// (at most 64): clusters merged by surface area within that many positions
// of the Morton order.
Status build_bvh_gpu(Context* ctx, const float* positions, uint32_t tri_count, uint32_t max_leaf, float* nodes_out,
                     uint32_t* node_count_out, uint32_t* tri_indices_out, double bounds_out[6], uint32_t radius = 0);

// ---------------------------------------------------------------------------
// Device scene layout (all device pointers; plain types so that host C++
// compiled without HIP can fill it in).
//
//   nodes4   : one 64-byte record per 4-wide node (scene_host.cpp,
//              convert_bvh4): the frame origin and per-axis power-of-two
//              scales, 8-bit lower/upper codes of the four child boxes per
//              axis (decoded as fmaf(code, scale, origin), rounded outward),
//              and four links: a node index, a leaf (bit 31 | (count-1) << 26
//              | first slot) or 0xFFFFFFFF for no child.
//   tris     : three float4 per leaf slot: (v0, material) (v1, light)
//              (v2, alpha flag), the w fields as int32 bits.
//   attrs    : four float4 per slot, the reference ATTR texels verbatim.
//   tri_ids  : slot -> original triangle id.
// ---------------------------------------------------------------------------
struct DevBvh {
  const float* tris = nullptr;    // float4[3 * tri_count]
  const float* attrs = nullptr;   // float4[4 * tri_count] or null
  const uint32_t* tri_ids = nullptr;
  // slot -> the caller's slot (position in its triIndices) when the device
  // tree was rebuilt, or null when the slots are the caller's; exact queries
  // break ties by it, as the caller's left-first traversal does.
  const uint32_t* ref_slots = nullptr;
  uint32_t tri_count = 0;
  uint32_t depth = 0;       // binary tree depth (deepest leaf, root = 0)
  // 4-wide nodes of the path tracer: 16 floats per node (frame origin, scale,
  // 8-bit child codes per axis, 4 links); a link is a node index, a leaf
  // (bit 31 | (count - 1) << 26 | first slot) or 0xFFFFFFFF for no child.
  const float* nodes4 = nullptr;
  const float* nodes4f = nullptr;  // the same nodes, float32 child bounds (32 floats per node)
  uint32_t node4_count = 0;
  uint32_t root4 = 0xFFFFFFFFu;   // link of the root (a node or a leaf)
  uint32_t stack4 = 1;            // (link, distance) entries a 4-wide traversal may need
};

struct Context {
  int device = 0;
  void* stream = nullptr;  // hipStream_t
  int compute_units = 0;
  int64_t global_mem = 0;
  // Scratch for host-pointer entry points, grown on demand.
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // Spill area for traversal stacks that outgrow their LDS part.
  void* spill = nullptr;
  size_t spill_bytes = 0;
  // Persistent tile launches: job counter, per-film job ranges and films.
  void* tile_table = nullptr;
  size_t tile_table_bytes = 0;
  void* tile_event = nullptr;   // hipEvent_t: the table upload has completed
  // Wavefront path tracer: path state and queues.
  void* wavefront = nullptr;
  size_t wavefront_bytes = 0;
  // Per-sample tile launches: path records and per-pixel counters, and the
  // most record bytes one launch may use.
  void* units = nullptr;
  size_t units_bytes = 0;
  size_t record_budget = size_t(9) << 29;
  uint32_t* d_flags = nullptr;  // device status word(s)
  float* sheen_lut = nullptr;   // 16x16 half-rounded sheen albedo table
  // Device tree built at scene upload (binned SAH over the caller's
  // triangles); tree_leaf = 0 keeps the caller's tree as the device tree.
  uint32_t tree_leaf = 2, tree_bins = 32;
};

// Host-side checks and conversion of the reference layout into the device layout.
// validate_bvh: leftFirst/count integral and in range, a tree; depth of the deepest leaf.
Status validate_bvh(const float* nodes, uint32_t node_count, uint32_t tri_count, uint32_t* depth);
struct HostBvh4Image {
  std::vector<float> nodes;  // 16 floats per node
  // The same nodes with float32 child bounds, 32 floats per node: the
  // frame origin (xyz, 0), then per axis the four children's lower and
  // upper offsets from the origin (rounded outward and widened by 2^-12 of
  // the node's extent), then the four links.
  std::vector<float> nodes_f32;
  uint32_t node_count = 0;
  uint32_t root = 0xFFFFFFFFu;
  uint32_t depth = 0;        // levels of 4-wide nodes
};
Status convert_bvh4(const float* nodes, uint32_t node_count, uint32_t tri_count, HostBvh4Image* out);

// Device memory helpers (implemented over the HIP runtime).
Status dev_alloc(void** p, size_t bytes);
Status dev_upload(void* dst, const void* src, size_t bytes, void* stream);
Status dev_download(void* dst, const void* src, size_t bytes, void* stream);
Status stream_sync(void* stream);
Status ensure_scratch(Context* ctx, size_t bytes);
Status ensure_spill(Context* ctx, size_t bytes);
Status ensure_tile_table(Context* ctx, size_t bytes);
Status ensure_wavefront(Context* ctx, size_t bytes);
Status ensure_units(Context* ctx, size_t bytes);

// Kernel launchers (traverse.hip).
struct RayQuery {
  const void* org;  // float or double xyz per ray
  const void* dir;
  bool f64_rays;
  uint32_t n;
  double tmin, tmax;
  uint32_t flags;
  double* t;
  uint32_t* id;
  double* u;
  double* v;
};
Status launch_intersect(Context* ctx, const DevBvh& bvh, const RayQuery& q, void* stream);
// The same query with the traversal stacks' spill area in a fixed device
// array of kNoAllocScratchBytes (no device allocation); noalloc_scratch_init
// looks up that array's address on `device` (current) once, at context set-up.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxNoAllocDevices = 64;
Status noalloc_scratch_init(int device);
Status launch_intersect_noalloc(Context* ctx, const DevBvh& bvh, const RayQuery& q, void* stream);

}  // namespace lucent
