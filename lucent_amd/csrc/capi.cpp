// The extern "C" entry points of include/lucent_hip.h: argument checks, error
// reporting, device memory management and dispatch to the kernel launchers.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "internal.hpp"
#include "render_host.hpp"
#include "scene.hpp"

namespace lucent {

namespace {
thread_local std::string g_last_error;
}

Status fail(Status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

Status fail_hip(int hip_err, const char* what) {
  const hipError_t e = static_cast<hipError_t>(hip_err);
  g_last_error = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  if (e == hipErrorOutOfMemory) return LUCENT_E_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return LUCENT_E_NO_DEVICE;
  return LUCENT_E_HIP;
}

void clear_error() { g_last_error.clear(); }

Status dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  const hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail_hip(e, "hipMalloc");
  return LUCENT_OK;
}

Status dev_upload(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return LUCENT_OK;
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync(H2D)");
  return LUCENT_OK;
}

Status dev_download(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return LUCENT_OK;
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync(D2H)");
  return LUCENT_OK;
}

Status stream_sync(void* stream) {
  const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
  return LUCENT_OK;
}

static Status grow(void** buf, size_t* have, size_t bytes, void* stream) {
  if (*have >= bytes) return LUCENT_OK;
  if (*buf) {
    // The old buffer may still be read by queued work on this stream.
    hipStreamSynchronize(static_cast<hipStream_t>(stream));
    hipFree(*buf);
    *buf = nullptr;
    *have = 0;
  }
  size_t want = bytes + bytes / 4;
  Status s = dev_alloc(buf, want);
  if (s != LUCENT_OK) return s;
  *have = want;
  return LUCENT_OK;
}

Status ensure_scratch(Context* ctx, size_t bytes) { return grow(&ctx->scratch, &ctx->scratch_bytes, bytes, ctx->stream); }
Status ensure_spill(Context* ctx, size_t bytes) { return grow(&ctx->spill, &ctx->spill_bytes, bytes, ctx->stream); }
Status ensure_wavefront(Context* ctx, size_t bytes) {
  return grow(&ctx->wavefront, &ctx->wavefront_bytes, bytes, ctx->stream);
}
Status ensure_units(Context* ctx, size_t bytes) { return grow(&ctx->units, &ctx->units_bytes, bytes, ctx->stream); }
Status ensure_tile_table(Context* ctx, size_t bytes) {
  return grow(&ctx->tile_table, &ctx->tile_table_bytes, bytes, ctx->stream);
}

namespace {

Status upload_owned(Scene* sc, const void* src, size_t bytes, const float** dst) {
  void* p = nullptr;
  Status s = dev_alloc(&p, bytes);
  if (s != LUCENT_OK) return s;
  sc->allocations.push_back(p);
  sc->stats.device_bytes += bytes;
  s = dev_upload(p, src, bytes, sc->ctx->stream);
  if (s != LUCENT_OK) return s;
  *dst = static_cast<const float*>(p);
  return LUCENT_OK;
}

void destroy_scene(Scene* sc) {
  if (!sc) return;
  if (sc->ctx) hipStreamSynchronize(static_cast<hipStream_t>(sc->ctx->stream));
  for (void* p : sc->allocations) hipFree(p);
  for (void* p : sc->env_allocations) hipFree(p);
  for (void* p : sc->tex_allocations) hipFree(p);
  delete sc;
}

bool finite_all(const float* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

// Shared tail of both upload forms: BVH conversion plus leaf-order geometry.
// The caller's tree (validated) and its leaf-ordered triangles -> the device
// scene.  Unless disabled on the context, the device tree is rebuilt with a
// finer binned SAH and smaller leaves (measured: fewer node visits and
// triangle tests for every traversal kernel), the per-slot arrays permuted
// into its leaf order, and each device slot remembers the caller's slot.
Status finish_scene(Context* ctx, Scene* sc, const float* nodes, uint32_t node_count, const uint32_t* tri_indices,
                    const std::vector<float>& tris_in, const float* attrs_in) {
  uint32_t depth = 0;
  Status s = validate_bvh(nodes, node_count, sc->bvh.tri_count, &depth);
  if (s != LUCENT_OK) return s;
  const uint32_t n = sc->bvh.tri_count;
  for (uint32_t i = 0; i < n && tri_indices; ++i)
    if (tri_indices[i] >= n) return fail(LUCENT_E_INVALID, "tri_indices entry out of range");
  sc->stats.tri_count = n;
  sc->stats.node_count = node_count;
  const float* dev_nodes = nodes;
  uint32_t dev_node_count = node_count;
  std::vector<float> tris, own_nodes;
  std::vector<float> attrs;
  std::vector<uint32_t> ids(n), ref;
  for (uint32_t i = 0; i < n; ++i) ids[i] = tri_indices ? tri_indices[i] : i;
  const bool rebuild = ctx->tree_leaf > 0 && n > 1 && n <= (1u << 24);
  if (rebuild) {
    std::vector<float> pos(size_t(n) * 9);
    for (uint32_t c = 0; c < n; ++c)
      for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) pos[size_t(c) * 9 + v * 3 + k] = tris_in[size_t(c) * 12 + v * 4 + k];
    own_nodes.resize(size_t(std::max<uint32_t>(1, 2 * n - 1)) * 8);
    ref.resize(n);
    double bounds[6];
    // The host cores build independent subtrees in parallel (any valid tree
    // serves: exact queries break ties by the caller's slots).
    // With tree_bins = 0 the device builds the tree (Morton order, clusters
    // merged by surface area) instead.
    const uint32_t threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (ctx->tree_bins == 0)
      s = build_bvh_gpu(ctx, pos.data(), n, ctx->tree_leaf, own_nodes.data(), &dev_node_count, ref.data(), bounds,
                        kClusterRadius);
    else
      s = build_bvh_parallel(pos.data(), n, ctx->tree_leaf, ctx->tree_bins, own_nodes.data(), &dev_node_count,
                             ref.data(), bounds, threads);
    if (s != LUCENT_OK) return s;
    if ((s = validate_bvh(own_nodes.data(), dev_node_count, n, &depth)) != LUCENT_OK) return s;
    dev_nodes = own_nodes.data();
    tris.resize(tris_in.size());
    std::vector<uint32_t> dev_ids(n);
    for (uint32_t d = 0; d < n; ++d) {
      const uint32_t c = ref[d];
      std::memcpy(&tris[size_t(d) * 12], &tris_in[size_t(c) * 12], 12 * sizeof(float));
      dev_ids[d] = ids[c];
    }
    ids.swap(dev_ids);
    if (attrs_in) {
      attrs.resize(size_t(n) * 16);
      for (uint32_t d = 0; d < n; ++d)
        std::memcpy(&attrs[size_t(d) * 16], attrs_in + size_t(ref[d]) * 16, 16 * sizeof(float));
    }
  }
  const std::vector<float>& t = rebuild ? tris : tris_in;
  const float* a = rebuild ? (attrs_in ? attrs.data() : nullptr) : attrs_in;
  sc->bvh.depth = depth;
  sc->stats.depth = depth;
  HostBvh4Image img4;
  if ((s = convert_bvh4(dev_nodes, dev_node_count, n, &img4)) != LUCENT_OK) return s;
  sc->bvh.node4_count = img4.node_count;
  sc->bvh.root4 = img4.root;
  // Each level pushes at most 3 entries beyond the one it pops.
  sc->bvh.stack4 = 3 * img4.depth + 2;
  sc->stats.node4_count = img4.node_count;
  if (img4.node_count > 0 &&
      (s = upload_owned(sc, img4.nodes.data(), img4.nodes.size() * sizeof(float), &sc->bvh.nodes4)) != LUCENT_OK)
    return s;
  if (img4.node_count > 0 &&
      (s = upload_owned(sc, img4.nodes_f32.data(), img4.nodes_f32.size() * sizeof(float), &sc->bvh.nodes4f)) !=
          LUCENT_OK)
    return s;
  if ((s = upload_owned(sc, t.data(), t.size() * sizeof(float), &sc->bvh.tris)) != LUCENT_OK) return s;
  if (a && (s = upload_owned(sc, a, size_t(n) * 16 * sizeof(float), &sc->bvh.attrs)) != LUCENT_OK) return s;
  const float* idp = nullptr;
  if ((s = upload_owned(sc, ids.data(), size_t(n) * sizeof(uint32_t), &idp)) != LUCENT_OK) return s;
  sc->bvh.tri_ids = reinterpret_cast<const uint32_t*>(idp);
  if (rebuild) {
    const float* rp = nullptr;
    if ((s = upload_owned(sc, ref.data(), size_t(n) * sizeof(uint32_t), &rp)) != LUCENT_OK) return s;
    sc->bvh.ref_slots = reinterpret_cast<const uint32_t*>(rp);
  }
  sc->pt.bvh = sc->bvh;
  return stream_sync(ctx->stream);
}

int32_t float_to_index(float f) { return int32_t(f); }

}  // namespace
}  // namespace lucent

using namespace lucent;


extern "C" {

const char* lucent_version(void) { return "lucent-mi355x 0.1 (gfx950)"; }

const char* lucent_last_error(const lucent_ctx*) { return g_last_error.c_str(); }

lucent_status lucent_create(int device, lucent_ctx** out) {
  clear_error();
  if (!out) return fail(LUCENT_E_INVALID, "out is null");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(LUCENT_E_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(LUCENT_E_NO_DEVICE, "device index out of range");
  if ((e = hipSetDevice(device)) != hipSuccess) return fail_hip(e, "hipSetDevice");
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return fail_hip(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(LUCENT_E_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
  auto* ctx = new lucent_ctx();
  ctx->device = device;
  ctx->compute_units = prop.multiProcessorCount;
  ctx->global_mem = int64_t(prop.totalGlobalMem);
  hipStream_t s = nullptr;
  if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
    delete ctx;
    return fail_hip(e, "hipStreamCreate");
  }
  ctx->stream = s;
  void* flags = nullptr;
  if (dev_alloc(&flags, 256) != LUCENT_OK) {
    hipStreamDestroy(s);
    delete ctx;
    return LUCENT_E_OUT_OF_MEMORY;
  }
  hipMemset(flags, 0, 256);
  ctx->d_flags = static_cast<uint32_t*>(flags);
  hipEvent_t ev = nullptr;
  if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
    hipFree(flags);
    hipStreamDestroy(s);
    delete ctx;
    return fail_hip(e, "hipEventCreate");
  }
  ctx->tile_event = ev;
  // The no-allocation query's scratch array: its address on this device.
  Status st = noalloc_scratch_init(device);
  if (st != LUCENT_OK) {
    hipEventDestroy(ev);
    hipFree(flags);
    hipStreamDestroy(s);
    delete ctx;
    return st;
  }
  *out = ctx;
  return LUCENT_OK;
}

void lucent_destroy(lucent_ctx* ctx) {
  if (!ctx) return;
  hipStreamSynchronize(static_cast<hipStream_t>(ctx->stream));
  if (ctx->scratch) hipFree(ctx->scratch);
  if (ctx->spill) hipFree(ctx->spill);
  if (ctx->d_flags) hipFree(ctx->d_flags);
  if (ctx->tile_table) hipFree(ctx->tile_table);
  if (ctx->wavefront) hipFree(ctx->wavefront);
  if (ctx->units) hipFree(ctx->units);
  if (ctx->tile_event) hipEventDestroy(static_cast<hipEvent_t>(ctx->tile_event));
  hipStreamDestroy(static_cast<hipStream_t>(ctx->stream));
  delete ctx;
}

lucent_status lucent_synchronize(lucent_ctx* ctx) {
  if (!ctx) return fail(LUCENT_E_INVALID, "ctx is null");
  return stream_sync(ctx->stream);
}

lucent_status lucent_device_info(lucent_ctx* ctx, int32_t* cus, int64_t* mem) {
  if (!ctx) return fail(LUCENT_E_INVALID, "ctx is null");
  if (cus) *cus = ctx->compute_units;
  if (mem) *mem = ctx->global_mem;
  return LUCENT_OK;
}

lucent_status lucent_build_bvh(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                               float* nodes_out, uint32_t* node_count_out, uint32_t* tri_indices_out,
                               double* bounds6_out) {
  clear_error();
  double bounds[6];
  Status s = build_bvh(positions, tri_count, max_leaf, bins, nodes_out, node_count_out, tri_indices_out, bounds);
  if (s == LUCENT_OK && bounds6_out) std::memcpy(bounds6_out, bounds, sizeof(bounds));
  return s;
}

lucent_status lucent_build_bvh_parallel(const float* positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                                        uint32_t threads, float* nodes_out, uint32_t* node_count_out,
                                        uint32_t* tri_indices_out, double* bounds6_out) {
  clear_error();
  if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
  double bounds[6];
  Status s = build_bvh_parallel(positions, tri_count, max_leaf, bins, nodes_out, node_count_out, tri_indices_out,
                                bounds, threads);
  if (s == LUCENT_OK && bounds6_out) std::memcpy(bounds6_out, bounds, sizeof(bounds));
  return s;
}

lucent_status lucent_build_bvh_gpu(lucent_ctx* ctx, const float* positions, uint32_t tri_count, uint32_t max_leaf,
                                   float* nodes_out, uint32_t* node_count_out, uint32_t* tri_indices_out,
                                   double* bounds6_out) {
  clear_error();
  if (!ctx || !nodes_out || !node_count_out || (tri_count > 0 && (!positions || !tri_indices_out)))
    return fail(LUCENT_E_INVALID, "null argument");
  if (tri_count >= 0x7FFFFFFFu) return fail(LUCENT_E_INVALID, "too many triangles");
  return build_bvh_gpu(ctx, positions, tri_count, max_leaf, nodes_out, node_count_out, tri_indices_out, bounds6_out);
}

lucent_status lucent_build_bvh_gpu_clustered(lucent_ctx* ctx, const float* positions, uint32_t tri_count,
                                             uint32_t max_leaf, uint32_t radius, float* nodes_out,
                                             uint32_t* node_count_out, uint32_t* tri_indices_out,
                                             double* bounds6_out) {
  clear_error();
  if (!ctx || !nodes_out || !node_count_out || (tri_count > 0 && (!positions || !tri_indices_out)))
    return fail(LUCENT_E_INVALID, "null argument");
  if (tri_count >= 0x7FFFFFFFu) return fail(LUCENT_E_INVALID, "too many triangles");
  if (radius > 64) return fail(LUCENT_E_INVALID, "search radius above 64");
  return build_bvh_gpu(ctx, positions, tri_count, max_leaf, nodes_out, node_count_out, tri_indices_out, bounds6_out,
                       radius ? radius : kClusterRadius);
}

lucent_status lucent_scene_from_bvh(lucent_ctx* ctx, const float* positions, uint32_t tri_count,
                                    const float* bvh_nodes, uint32_t bvh_node_count, const uint32_t* tri_indices,
                                    lucent_scene** out) {
  clear_error();
  if (!ctx || !out) return fail(LUCENT_E_INVALID, "null ctx/out");
  *out = nullptr;
  if (tri_count > 0 && (!positions || !tri_indices)) return fail(LUCENT_E_INVALID, "null positions/tri_indices");
  if (tri_count > 0 && !finite_all(positions, size_t(tri_count) * 9))
    return fail(LUCENT_E_INVALID, "positions contain NaN or infinity");
  auto* sc = new lucent_scene();
  sc->ctx = ctx;
  sc->bvh.tri_count = tri_count;
  std::vector<float> tris(size_t(tri_count) * 12, 0.0f);
  for (uint32_t slot = 0; slot < tri_count; ++slot) {
    const uint32_t src = tri_indices[slot];
    if (src >= tri_count) {
      destroy_scene(sc);
      return fail(LUCENT_E_INVALID, "tri_indices entry out of range");
    }
    for (int v = 0; v < 3; ++v)
      for (int a = 0; a < 3; ++a) tris[size_t(slot) * 12 + v * 4 + a] = positions[size_t(src) * 9 + v * 3 + a];
    // w fields: material 0, light -1, alpha flag 0 (as int32 bits).
    const int32_t mat = 0, light = -1, alpha = 0;
    std::memcpy(&tris[size_t(slot) * 12 + 3], &mat, 4);
    std::memcpy(&tris[size_t(slot) * 12 + 7], &light, 4);
    std::memcpy(&tris[size_t(slot) * 12 + 11], &alpha, 4);
  }
  Status s = finish_scene(ctx, sc, bvh_nodes, bvh_node_count, tri_indices, tris, nullptr);
  if (s != LUCENT_OK) {
    destroy_scene(sc);
    return s;
  }
  *out = sc;
  return LUCENT_OK;
}

lucent_status lucent_set_tree_build(lucent_ctx* ctx, uint32_t max_leaf, uint32_t bins) {
  clear_error();
  if (!ctx) return fail(LUCENT_E_INVALID, "null ctx");
  if (max_leaf > 16) return fail(LUCENT_E_INVALID, "max_leaf must be at most 16 (0 keeps the caller's tree)");
  if (max_leaf > 0 && bins != 0 && (bins < 2 || bins > 4096))
    return fail(LUCENT_E_INVALID, "bins must be 0 (a linear BVH built on the device) or in [2, 4096]");
  ctx->tree_leaf = max_leaf;
  ctx->tree_bins = max_leaf > 0 ? bins : ctx->tree_bins;
  return LUCENT_OK;
}

lucent_status lucent_set_record_budget(lucent_ctx* ctx, uint64_t bytes) {
  clear_error();
  if (!ctx) return fail(LUCENT_E_INVALID, "null ctx");
  ctx->record_budget = bytes ? size_t(bytes) : size_t(9) << 29;
  return LUCENT_OK;
}

lucent_status lucent_scene_upload(lucent_ctx* ctx, const lucent_packed_scene* ps, lucent_scene** out) {
  clear_error();
  if (!ctx || !ps || !out) return fail(LUCENT_E_INVALID, "null ctx/packed/out");
  *out = nullptr;
  const uint32_t n = ps->tri_count;
  if (n > 0 && !ps->tri_geometry) return fail(LUCENT_E_INVALID, "tri_geometry is null");
  if (ps->sphere_count > 64) return fail(LUCENT_E_INVALID, "scene has more than 64 spheres; max is 64");
  if (ps->light_count > 8) return fail(LUCENT_E_INVALID, "scene has more than 8 lights; max is 8");
  if (ps->material_count > 0 && !ps->materials) return fail(LUCENT_E_INVALID, "materials is null");
  auto* sc = new lucent_scene();
  sc->ctx = ctx;
  sc->bvh.tri_count = n;
  std::vector<float> tris(size_t(n) * 12);
  for (uint32_t slot = 0; slot < n; ++slot) {
    const float* g = ps->tri_geometry + size_t(slot) * 12;
    for (int k = 0; k < 12; ++k) tris[size_t(slot) * 12 + k] = g[k];
    const int32_t mat = float_to_index(g[3]), light = float_to_index(g[7]), alpha = g[11] >= 0.5f ? 1 : 0;
    if (mat < 0 || (ps->material_count > 0 && uint32_t(mat) >= ps->material_count)) {
      destroy_scene(sc);
      return fail(LUCENT_E_INVALID, "triangle material index out of range");
    }
    if (light >= int32_t(ps->light_count)) {
      destroy_scene(sc);
      return fail(LUCENT_E_INVALID, "triangle light index out of range");
    }
    std::memcpy(&tris[size_t(slot) * 12 + 3], &mat, 4);
    std::memcpy(&tris[size_t(slot) * 12 + 7], &light, 4);
    std::memcpy(&tris[size_t(slot) * 12 + 11], &alpha, 4);
  }
  for (size_t i = 0; i < tris.size(); ++i)
    if (i % 4 != 3 && !std::isfinite(tris[i])) {
      destroy_scene(sc);
      return fail(LUCENT_E_INVALID, "triangle geometry contains NaN or infinity");
    }
  // Integrator tables: materials, spheres, lights, ground (packScene outputs).
  PtScene& pt = sc->pt;
  Status s = LUCENT_OK;
  if (ps->material_count) {
    const float* m = nullptr;
    if ((s = upload_owned(sc, ps->materials, size_t(ps->material_count) * 40 * sizeof(float), &m)) != LUCENT_OK) {
      destroy_scene(sc);
      return s;
    }
    pt.materials = m;
    sc->blend_alpha = any_blend_alpha(ps->materials, ps->material_count);
  }
  pt.material_count = ps->material_count;
  if (ps->sphere_count) {
    if (!ps->spheres || !ps->sphere_materials || !ps->sphere_lights) {
      destroy_scene(sc);
      return fail(LUCENT_E_INVALID, "sphere arrays are null");
    }
    for (uint32_t i = 0; i < ps->sphere_count; ++i)
      if (ps->sphere_materials[i] < 0 || uint32_t(ps->sphere_materials[i]) >= ps->material_count ||
          ps->sphere_lights[i] >= int32_t(ps->light_count)) {
        destroy_scene(sc);
        return fail(LUCENT_E_INVALID, "sphere material or light index out of range");
      }
    const float *sp = nullptr, *sm = nullptr, *sl = nullptr;
    if ((s = upload_owned(sc, ps->spheres, size_t(ps->sphere_count) * 16, &sp)) != LUCENT_OK ||
        (s = upload_owned(sc, ps->sphere_materials, size_t(ps->sphere_count) * 4, &sm)) != LUCENT_OK ||
        (s = upload_owned(sc, ps->sphere_lights, size_t(ps->sphere_count) * 4, &sl)) != LUCENT_OK) {
      destroy_scene(sc);
      return s;
    }
    pt.spheres = sp;
    pt.sphere_material = reinterpret_cast<const int32_t*>(sm);
    pt.sphere_light = reinterpret_cast<const int32_t*>(sl);
  }
  pt.sphere_count = ps->sphere_count;
  if (ps->light_count) std::memcpy(pt.lights, ps->lights, size_t(ps->light_count) * 16 * sizeof(float));
  pt.light_count = ps->light_count;
  if (ps->ground_material >= 0 && uint32_t(ps->ground_material) >= ps->material_count) {
    destroy_scene(sc);
    return fail(LUCENT_E_INVALID, "ground material index out of range");
  }
  pt.ground_material = ps->ground_material;
  pt.ground_height = ps->ground_height;
  s = finish_scene(ctx, sc, ps->bvh_nodes, ps->bvh_node_count, ps->tri_indices, tris, ps->tri_attributes);
  if (s != LUCENT_OK) {
    destroy_scene(sc);
    return s;
  }
  *out = sc;
  return LUCENT_OK;
}

void lucent_scene_destroy(lucent_scene* scene) { destroy_scene(scene); }

lucent_status lucent_scene_stats_get(const lucent_scene* scene, lucent_scene_stats* out) {
  if (!scene || !out) return fail(LUCENT_E_INVALID, "null scene/out");
  *out = scene->stats;
  return LUCENT_OK;
}

static lucent_status intersect_host(lucent_ctx* ctx, const lucent_scene* scene, const void* org, const void* dir,
                                    bool f64, uint32_t n, double tmin, double tmax, uint32_t flags, double* t,
                                    uint32_t* id, double* u, double* v) {
  clear_error();
  if (!ctx || !scene) return fail(LUCENT_E_INVALID, "null ctx/scene");
  if (n == 0) return LUCENT_OK;
  if (!org || !dir || !t || !id || !u || !v) return fail(LUCENT_E_INVALID, "null array");
  const size_t ray_bytes = size_t(n) * 3 * (f64 ? sizeof(double) : sizeof(float));
  const size_t out_bytes = size_t(n) * (3 * sizeof(double) + sizeof(uint32_t));
  const size_t align = 256;
  auto up = [&](size_t b) { return (b + align - 1) / align * align; };
  const size_t total = 2 * up(ray_bytes) + up(out_bytes);
  Status s = ensure_scratch(ctx, total);
  if (s != LUCENT_OK) return s;
  char* base = static_cast<char*>(ctx->scratch);
  void* d_org = base;
  void* d_dir = base + up(ray_bytes);
  char* o = base + 2 * up(ray_bytes);
  double* d_t = reinterpret_cast<double*>(o);
  double* d_u = d_t + n;
  double* d_v = d_u + n;
  uint32_t* d_id = reinterpret_cast<uint32_t*>(d_v + n);
  if ((s = dev_upload(d_org, org, ray_bytes, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_upload(d_dir, dir, ray_bytes, ctx->stream)) != LUCENT_OK) return s;
  RayQuery q{d_org, d_dir, f64, n, tmin, tmax, flags, d_t, d_id, d_u, d_v};
  if ((s = launch_intersect(ctx, scene->bvh, q, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_download(t, d_t, size_t(n) * 8, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_download(u, d_u, size_t(n) * 8, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_download(v, d_v, size_t(n) * 8, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_download(id, d_id, size_t(n) * 4, ctx->stream)) != LUCENT_OK) return s;
  return stream_sync(ctx->stream);
}

lucent_status lucent_intersect(lucent_ctx* ctx, const lucent_scene* scene, const float* org3, const float* dir3,
                               uint32_t n, double tmin, double tmax, uint32_t flags, double* t_out,
                               uint32_t* tri_id_out, double* u_out, double* v_out) {
  return intersect_host(ctx, scene, org3, dir3, false, n, tmin, tmax, flags, t_out, tri_id_out, u_out, v_out);
}

lucent_status lucent_intersect_f64(lucent_ctx* ctx, const lucent_scene* scene, const double* org3,
                                   const double* dir3, uint32_t n, double tmin, double tmax, uint32_t flags,
                                   double* t_out, uint32_t* tri_id_out, double* u_out, double* v_out) {
  return intersect_host(ctx, scene, org3, dir3, true, n, tmin, tmax, flags, t_out, tri_id_out, u_out, v_out);
}

lucent_status lucent_intersect_device(lucent_ctx* ctx, const lucent_scene* scene, const float* d_org3,
                                      const float* d_dir3, uint32_t n, double tmin, double tmax, uint32_t flags,
                                      double* d_t, uint32_t* d_tri_id, double* d_u, double* d_v, void* stream) {
  clear_error();
  if (!ctx || !scene) return fail(LUCENT_E_INVALID, "null ctx/scene");
  if (n == 0) return LUCENT_OK;
  if (!d_org3 || !d_dir3 || !d_t || !d_tri_id || !d_u || !d_v) return fail(LUCENT_E_INVALID, "null array");
  RayQuery q{d_org3, d_dir3, false, n, tmin, tmax, flags, d_t, d_tri_id, d_u, d_v};
  return launch_intersect(ctx, scene->bvh, q, stream ? stream : ctx->stream);
}

lucent_status lucent_intersect_device_noalloc(lucent_ctx* ctx, const lucent_scene* scene, const float* d_org3,
                                              const float* d_dir3, uint32_t n, double tmin, double tmax,
                                              uint32_t flags, double* d_t, uint32_t* d_tri_id, double* d_u,
                                              double* d_v, void* stream) {
  clear_error();
  if (!ctx || !scene) return fail(LUCENT_E_INVALID, "null ctx/scene");
  if (n == 0) return LUCENT_OK;
  if (!d_org3 || !d_dir3 || !d_t || !d_tri_id || !d_u || !d_v) return fail(LUCENT_E_INVALID, "null array");
  RayQuery q{d_org3, d_dir3, false, n, tmin, tmax, flags, d_t, d_tri_id, d_u, d_v};
  return launch_intersect_noalloc(ctx, scene->bvh, q, stream ? stream : ctx->stream);
}

}  // extern "C"
