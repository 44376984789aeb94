/*
 * lucent_hip.h — C ABI of the MI355X (gfx950) implementation of Lucent's
 * path-tracing hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns a lucent_status: 0 on success, a negative code on failure with a
 * message in lucent_last_error().  Host arrays are borrowed for the duration
 * of the call only.  Entry points suffixed _device take device pointers and
 * enqueue on the given HIP stream (NULL = the context's own stream) without
 * synchronising.  A context is not re-entrant: calls on one context must not
 * overlap.
 *
 * Reference interfaces each entry point replaces are cited as
 * <file>:<line> of clark719664/lucent.
 */
#ifndef LUCENT_HIP_H
#define LUCENT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LUCENT_ABI_VERSION 1

typedef int32_t lucent_status;
enum {
  LUCENT_OK = 0,
  LUCENT_E_INVALID = -1,     /* bad argument; mirrors the reference's thrown Error */
  LUCENT_E_HIP = -2,         /* a HIP runtime call failed */
  LUCENT_E_NO_DEVICE = -3,   /* no usable gfx950 device */
  LUCENT_E_OUT_OF_MEMORY = -4,
  LUCENT_E_UNSUPPORTED = -5, /* input outside what the kernels handle (e.g. tree too deep) */
  LUCENT_E_OVERFLOW = -6     /* a device-side bound was hit; results are not trustworthy */
};

typedef struct lucent_ctx lucent_ctx;
typedef struct lucent_scene lucent_scene;

/* Version string of the library build. */
const char *lucent_version(void);

/* Message of the last failed call on the calling thread ("" if none).
 * ctx may be NULL. */
const char *lucent_last_error(const lucent_ctx *ctx);

/* Opens device `device` and creates the context's stream.
 * Replaces the WebGL2 context set-up of Renderer's constructor
 * (src/renderer.ts:226; src/gl/context.ts:18-29, which throws GLError). */
lucent_status lucent_create(int device, lucent_ctx **out);
/* Destroy the context's scenes, films and frame graphs first: they refer to
 * its stream. */
void lucent_destroy(lucent_ctx *ctx);

/* Waits for all work enqueued on the context's stream. */
lucent_status lucent_synchronize(lucent_ctx *ctx);

/* Number of compute units of the context's device (for launch sizing). */
lucent_status lucent_device_info(lucent_ctx *ctx, int32_t *compute_units, int64_t *global_mem_bytes);

/* ------------------------------------------------------------------------
 * BVH build — buildBVH(positions, {maxLeafSize, bins}) (src/math/bvh.ts:52).
 * positions: 9 float32 per triangle.  nodes_out must hold
 * max(1, 2*tri_count-1)*8 floats in the layout of src/math/bvh.ts:6-16
 * [min.xyz, leftFirst, max.xyz, count]; tri_indices_out tri_count entries;
 * bounds6_out (min.xyz, max.xyz) may be NULL.  Host-only: needs no device.
 * ---------------------------------------------------------------------- */
lucent_status lucent_build_bvh(const float *positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                               float *nodes_out, uint32_t *node_count_out, uint32_t *tri_indices_out,
                               double *bounds6_out);

/* The splits of lucent_build_bvh with independent subtrees built on
 * `threads` host threads (0 = the host's hardware threads): the same node
 * layout and leaf order semantics (root 0, children adjacent, leaves of at
 * most max_leaf triangles), the same set of nodes, numbered subtree by
 * subtree instead of in one depth-first order; deterministic.  No
 * counterpart in the reference (buildBVHAsync moves the one-threaded build
 * to a worker, src/math/bvh-async.ts:8-30). */
lucent_status lucent_build_bvh_parallel(const float *positions, uint32_t tri_count, uint32_t max_leaf, uint32_t bins,
                                        uint32_t threads, float *nodes_out, uint32_t *node_count_out,
                                        uint32_t *tri_indices_out, double *bounds6_out);

/* GPU builder with the same outputs and layout as lucent_build_bvh (root
 * node 0, children of an interior node adjacent, leaves of at most max_leaf
 * triangles, triIndices in leaf order), built on the context's device as a
 * linear BVH: Morton codes of the triangle centroids, radix sort, the
 * radix tree of the sorted codes and bottom-up boxes.  Its trees cost more to
 * traverse than the SAH builder's; its build is O(n log n) on the device.
 * max_leaf 0 means 4.  No counterpart in the reference (it builds on the CPU,
 * in a worker for large scenes: src/math/bvh-async.ts:8-30). */
lucent_status lucent_build_bvh_gpu(lucent_ctx *ctx, const float *positions, uint32_t tri_count, uint32_t max_leaf,
                                   float *nodes_out, uint32_t *node_count_out, uint32_t *tri_indices_out,
                                   double *bounds6_out);

/* GPU builder with the outputs and layout of lucent_build_bvh_gpu whose tree
 * is built by surface area: after the Morton sort, clusters (the triangles at
 * first) are merged in rounds on the device, each cluster pairing with the
 * neighbour within `radius` positions of the Morton order whose union box has
 * the least surface area, mutual pairs merging, until one cluster is left
 * (parallel locally-ordered clustering, Meister and Bittner 2018).  Subtrees
 * of at most max_leaf triangles become leaves.  radius 0 means 16, at most
 * 64; max_leaf 0 means 4.  Its trees cost about what the SAH builder's do to
 * traverse.  No counterpart in the reference. */
lucent_status lucent_build_bvh_gpu_clustered(lucent_ctx *ctx, const float *positions, uint32_t tri_count,
                                             uint32_t max_leaf, uint32_t radius, float *nodes_out,
                                             uint32_t *node_count_out, uint32_t *tri_indices_out,
                                             double *bounds6_out);

/* ------------------------------------------------------------------------
 * Scene upload — PackedScene (src/scene/scene.ts:106-134) as uploaded by
 * Renderer.loadScene (src/renderer.ts:279-312).  Arrays use the reference's
 * texel layouts verbatim:
 *   tri_geometry   12 f32/tri: (v0, material) (v1, light) (v2, alpha flag)  scene.ts:213-235
 *   tri_attributes 16 f32/tri: (n0,u0)(n1,u1)(n2,u2)(v0,v1,v2,0); may be NULL
 *   bvh_nodes       8 f32/node, bvh.ts:6-16
 *   tri_indices    slot -> original triangle id (BVH.triIndices); NULL = identity
 *   materials      40 f32/material, material.ts:117-131
 *   spheres        4 f32/sphere (centre, radius); sphere_materials, sphere_lights int32
 *   lights         16 f32/light, scene.ts:88-95
 * Integer-valued float fields are converted to integer fields on the device.
 * ---------------------------------------------------------------------- */
typedef struct lucent_packed_scene {
  const float *tri_geometry;
  const float *tri_attributes;
  uint32_t tri_count;
  const float *bvh_nodes;
  uint32_t bvh_node_count;
  const uint32_t *tri_indices;
  const float *materials;
  uint32_t material_count;
  const float *spheres;
  const int32_t *sphere_materials;
  const int32_t *sphere_lights;
  uint32_t sphere_count;
  const float *lights;
  uint32_t light_count;
  int32_t ground_material; /* -1 = no ground plane (uGroundMaterial) */
  float ground_height;
} lucent_packed_scene;

lucent_status lucent_scene_upload(lucent_ctx *ctx, const lucent_packed_scene *packed, lucent_scene **out);

/* The device tree of scenes uploaded afterwards on this context.  The
 * caller's BVH (bvh.ts:6-16 layout) is always validated; unless max_leaf is
 * 0 the device tree is then rebuilt from the same triangles with a binned
 * SAH of `bins` bins and leaves of at most max_leaf triangles (default 2
 * and 32: fewer node visits and triangle tests than buildBVH's 4 and 12 for
 * every traversal here), and hit records keep their meaning: original
 * triangle ids, and exact queries still resolve ties to the lowest slot of
 * the caller's tree.  max_leaf = 0 keeps the caller's tree as the device
 * tree.  The SAH rebuild runs its independent subtrees on the host's
 * threads (lucent_build_bvh_parallel); bins = 0 builds the device tree on
 * the GPU instead (lucent_build_bvh_gpu_clustered with radius 16 and leaves
 * of at most max_leaf).  No counterpart in the reference (its shaders walk
 * the tree they are given). */
lucent_status lucent_set_tree_build(lucent_ctx *ctx, uint32_t max_leaf, uint32_t bins);

/* Device memory a tile launch may use for path records (48 B per pixel
 * sample; default 4.5 GiB, which holds 1440x900 at 64 samples): a single
 * film above it runs as several launches, each continuing the film; a batch
 * whose per-wave record rings exceed it hands out whole pixels instead
 * (bit-identical either way).  bytes = 0 restores the default.  No
 * counterpart in the reference. */
lucent_status lucent_set_record_budget(lucent_ctx *ctx, uint64_t bytes);

/* Convenience form for traversal only: the BVH of buildBVH plus the positions
 * it was built over (the arguments of intersectBVH, src/math/bvh.ts:265-271);
 * triangles are gathered into leaf order on upload, as packScene does. */
lucent_status lucent_scene_from_bvh(lucent_ctx *ctx, const float *positions, uint32_t tri_count,
                                    const float *bvh_nodes, uint32_t bvh_node_count,
                                    const uint32_t *tri_indices, lucent_scene **out);

void lucent_scene_destroy(lucent_scene *scene);

typedef struct lucent_scene_stats {
  uint32_t tri_count;
  uint32_t node_count;  /* nodes of the caller's tree */
  uint32_t node4_count; /* 64-byte 4-wide nodes on the device */
  uint32_t depth;       /* deepest leaf of the device's binary tree, root = 0 */
  uint64_t device_bytes;
} lucent_scene_stats;
lucent_status lucent_scene_stats_get(const lucent_scene *scene, lucent_scene_stats *out);

/* ------------------------------------------------------------------------
 * Batched ray queries — intersectBVH (src/math/bvh.ts:265-320), called per
 * ray by pickDistance (src/scene/scene.ts:327-341).
 *
 * flags: LUCENT_PREC_EXACT_FP64 reproduces the CPU traversal's hit records
 * bit for bit (FP64 Möller–Trumbore in the op order of src/math/ray.ts:22-45,
 * ties resolved to the lowest leaf slot as its left-first traversal does);
 * LUCENT_PREC_GLSL_FP32 follows the shader's float32 traversal
 * (src/shaders/scene.glsl:221-238 hitTriangle, EPS = 1e-6 replaces tmin).
 * LUCENT_QUERY_ANY returns any accepted hit (occlusion, scene.glsl:336-377)
 * with t/u/v of that hit.
 * A miss writes tri_id = 0xFFFFFFFF, t = +inf, u = v = 0 (the reference's null).
 * ---------------------------------------------------------------------- */
#define LUCENT_QUERY_CLOSEST 0u
#define LUCENT_QUERY_ANY 1u
#define LUCENT_PREC_EXACT_FP64 0u
#define LUCENT_PREC_GLSL_FP32 2u
#define LUCENT_MISS 0xFFFFFFFFu
/* Testing aid: keeps only 2 stack entries per lane in LDS so that the global
 * spill path of the traversal stack runs on any tree (results unchanged). */
#define LUCENT_FLAG_TINY_LDS_STACK 0x100u

lucent_status lucent_intersect(lucent_ctx *ctx, const lucent_scene *scene, const float *org3, const float *dir3,
                               uint32_t n, double tmin, double tmax, uint32_t flags, double *t_out,
                               uint32_t *tri_id_out, double *u_out, double *v_out);

/* Same with double-precision rays (JavaScript numbers, as pickFocus builds them). */
lucent_status lucent_intersect_f64(lucent_ctx *ctx, const lucent_scene *scene, const double *org3,
                                   const double *dir3, uint32_t n, double tmin, double tmax, uint32_t flags,
                                   double *t_out, uint32_t *tri_id_out, double *u_out, double *v_out);

/* Device-pointer form of lucent_intersect; asynchronous on `stream`. */
lucent_status lucent_intersect_device(lucent_ctx *ctx, const lucent_scene *scene, const float *d_org3,
                                      const float *d_dir3, uint32_t n, double tmin, double tmax, uint32_t flags,
                                      double *d_t, uint32_t *d_tri_id, double *d_u, double *d_v, void *stream);

/* lucent_intersect_device without any device allocation: no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * traversal stacks spill into a fixed 64 MiB device array that the library
 * holds per device for the life of the process (its address is looked up
 * once per device in lucent_create); the launch's grid shrinks to what that
 * array holds, and a tree whose stacks need more than it holds for one
 * workgroup (about 128 * 8 * (3 * depth4 - 10) bytes) is rejected with
 * LUCENT_E_UNSUPPORTED.  Results equal lucent_intersect_device's.  All calls
 * on one device share the array: calls that overlap on one device (other
 * streams, other contexts, other threads) race. */
lucent_status lucent_intersect_device_noalloc(lucent_ctx *ctx, const lucent_scene *scene, const float *d_org3,
                                              const float *d_dir3, uint32_t n, double tmin, double tmax,
                                              uint32_t flags, double *d_t, uint32_t *d_tri_id, double *d_u,
                                              double *d_v, void *stream);

/* ------------------------------------------------------------------------
 * Lighting — the Scene's sun, sky and environment (src/scene/scene.ts:21-38,
 * src/scene/environment.ts:12-22) as uploaded by Renderer.uploadLighting
 * (src/renderer.ts:355-413): the analytic sun is enabled only without an
 * environment map; without a map the procedural sky is baked into a
 * 1024x512 map (environment.ts:106-132); the map's importance sampler is
 * buildEnvSampler (environment.ts:52-85).  Angles in degrees.
 * ---------------------------------------------------------------------- */
typedef struct lucent_sun {
  double elevation, azimuth;
  double color[3];
  double intensity;
  double angular_radius;
} lucent_sun;

typedef struct lucent_sky {
  int32_t enabled;
  double zenith[3], horizon[3], ground[3];
  double intensity;
} lucent_sky;

typedef struct lucent_environment {
  const float *map_rgb; /* NULL = procedural sky; else width*height*3 float32, rows top to bottom */
  uint32_t width, height;
  double rotation; /* about +Y, degrees */
  double intensity;
  int32_t background; /* camera rays that miss see the map (1) or the backdrop (0) */
  double backdrop[3];
} lucent_environment;

typedef struct lucent_lighting {
  lucent_sun sun;
  lucent_sky sky;
  lucent_environment env;
} lucent_lighting;

lucent_status lucent_scene_set_lighting(lucent_ctx *ctx, lucent_scene *scene, const lucent_lighting *lighting);

/* Texture arrays of Renderer.buildTextureArrays (src/renderer.ts:315-335,
 * createTextureArray src/gl/context.ts:224-250): `color` holds the sRGB
 * layers (base colour, emissive; decoded to linear per texel), `data` the
 * linear ones (metal-rough, normal), each layer `size` x `size` RGBA8 texels
 * (R in the low byte), rows in v order, already resampled to the shared
 * size.  Sampling is bilinear with REPEAT wrap at level 0.  Materials refer
 * to layers by index (packScene's layerMap, src/scene/scene.ts:136-153).
 * Replaces the scene's previous arrays; 0 layers clears them. */
lucent_status lucent_scene_set_textures(lucent_ctx *ctx, lucent_scene *scene, uint32_t size,
                                        const uint32_t *color_rgba8, uint32_t color_layers,
                                        const uint32_t *data_rgba8, uint32_t data_layers);

/* buildEnvSampler(img, maxWidth) (environment.ts:52-85), host-only.  table_out
 * must hold 4 floats per grid cell (at most min(width,max_width) x height). */
lucent_status lucent_env_sampler(const float *rgb, uint32_t width, uint32_t height, uint32_t max_width,
                                 float *table_out, uint32_t *grid_w, uint32_t *grid_h, double *total);

/* buildEnvSampler on the device, in float64: d_rgba is an RGBA float map
 * (4 floats per texel, rows top to bottom; the scene's env map layout) in
 * device memory, d_table_out (4 floats per cell, 16-byte aligned) receives
 * the table.  Grid, table and total are bit-identical to lucent_env_sampler
 * on the same texels: the cell weights, pmf and table texels are computed
 * one thread per cell; the weight sum and Vose's pairing (alias.ts:44-63)
 * are sequential float64 chains run by one wave, in the host order.
 * Enqueued on `stream` (NULL = the context's stream) and synchronous;
 * uses the context's scratch memory.  gpu_ms (may be NULL) receives the
 * device time between events around the kernels.  Fails with the host
 * build's message when a weight is negative or not finite, leaving
 * d_table_out untouched.  No counterpart in the reference (it builds the
 * table on the CPU, environment.ts:52-85). */
lucent_status lucent_env_sampler_gpu(lucent_ctx *ctx, const float *d_rgba, uint32_t width, uint32_t height,
                                     uint32_t max_width, float *d_table_out, uint32_t *grid_w, uint32_t *grid_h,
                                     double *total, float *gpu_ms, void *stream);

/* lucent_env_sampler_gpu with both sequential chains replaced by parallel
 * FP64 work: the weights' total is a tree sum, and the alias pairs come from
 * inclusive FP64 prefix sums of the small cells' deficits (1 - pmf n) and
 * the large cells' excesses (pmf n - 1), each cell's pair found by binary
 * search (a small cell takes the first large cell whose excess prefix
 * reaches its deficit prefix; a large cell keeps 1 minus its overdraw and
 * aliases the next large cell).  Same grid and the same distribution as
 * buildEnvSampler (each cell's implied probability is its pmf up to the
 * float32 rounding of the stored probabilities), other alias pairs, so a
 * random number may select another cell than the reference's table does;
 * the total may differ from the sequential sum in its last bits.  For
 * callers that rebuild the sampler often (a map under optimisation).  Same
 * arguments, errors and synchronisation as lucent_env_sampler_gpu.  No
 * counterpart in the reference. */
lucent_status lucent_env_sampler_gpu_parallel(lucent_ctx *ctx, const float *d_rgba, uint32_t width, uint32_t height,
                                              uint32_t max_width, float *d_table_out, uint32_t *grid_w,
                                              uint32_t *grid_h, double *total, float *gpu_ms, void *stream);

/* buildAliasTable (src/math/alias.ts:17-64), host-only. */
lucent_status lucent_alias_table(const double *weights, uint32_t n, float *prob_out, uint32_t *alias_out,
                                 double *pmf_out, double *total_out);

/* sheenAlbedoLUT(size, steps) (src/scene/sheen.ts:52-61), host-only, float64
 * values (the device copy is rounded to half, as the R16F upload is). */
lucent_status lucent_sheen_lut(uint32_t size, uint32_t steps, double *out);

/* ------------------------------------------------------------------------
 * Film — the accumulation targets of Renderer.resize (src/renderer.ts:476-496,
 * formats at :490): colour RGBA32F (mean radiance, mean luminance^2), albedo
 * RGBA16F (guide albedo, coverage), normal RGBA16F (guide normal, primary
 * depth), history R32F (samples behind each pixel).  Rows bottom to top.
 * The RNG stream index `seed` advances every sample and survives resets
 * (renderer.ts:201-202, 728).
 * ---------------------------------------------------------------------- */
typedef struct lucent_film lucent_film;

lucent_status lucent_film_create(lucent_ctx *ctx, uint32_t width, uint32_t height, lucent_film **out);
void lucent_film_destroy(lucent_film *film);
/* Renderer.reset (renderer.ts:463-474): zero the averages and the history.
 * Ordered on the context's stream and synchronous; work the caller enqueued
 * on another stream that touches the film must have completed. */
lucent_status lucent_film_reset(lucent_ctx *ctx, lucent_film *film);
lucent_status lucent_film_set_seed(lucent_film *film, uint32_t seed);
lucent_status lucent_film_info(const lucent_film *film, uint32_t *width, uint32_t *height, uint32_t *samples,
                               uint32_t *seed);
/* Copies the targets to host memory; the half targets are widened to float.
 * Any pointer may be NULL. */
lucent_status lucent_film_read(lucent_ctx *ctx, const lucent_film *film, float *color4, float *albedo4,
                               float *normal4, float *history);
/* Device addresses of the four targets (for zero-copy consumers). */
lucent_status lucent_film_device_ptrs(const lucent_film *film, void **color4, void **albedo_half4,
                                      void **normal_half4, void **history);

/* ------------------------------------------------------------------------
 * Progressive tracing — Renderer.traceSamples (src/renderer.ts:702-737): each
 * of `count` samples is one launch over the film with uSeed = seed++, adding
 * one path per pixel to the running means (trace.frag:324-355).
 * camera: Camera (scene.ts:10-19), basis as cameraBasis (renderer.ts:156-161).
 * settings: the integrator part of RenderSettings (renderer.ts:42-66).
 * ---------------------------------------------------------------------- */
typedef struct lucent_camera {
  double position[3];
  double target[3];
  double fov; /* vertical, degrees */
  double aperture;
  double focus_distance;
} lucent_camera;

/* Launch structure of lucent_trace_samples / lucent_trace_batch.  Both give
 * bit-identical films.  Default: persistent waves take tile jobs (all
 * samples of a tile of pixels, sized by the settings' launch shape) and hand
 * their pixel samples to their lanes, starting a new path as soon as one
 * ends.
 * LUCENT_TRACE_PER_SAMPLE: one launch over the film per sample, as the
 * reference's traceSamples draws one full-screen pass per sample. */
#define LUCENT_TRACE_PER_SAMPLE 0x1u
/* Wavefront launch structure: separate closest-hit, shadow and shading passes
 * over a queue of live paths (one per pixel), the traversal passes persistent
 * with idle lanes refilled from the queue.  Same film bit for bit.  Scenes
 * with alpha-blended materials use the default structure. */
#define LUCENT_TRACE_WAVEFRONT 0x2u
/* Default structure only: trace each shadow ray where the integrator asks for
 * it (trace.frag's order of loops) instead of feeding a lane's closest-hit
 * and shadow rays through one traversal loop.  Same film bit for bit; scenes
 * with alpha-blended materials always trace this way. */
#define LUCENT_TRACE_INLINE_SHADOWS 0x4u

typedef struct lucent_render_settings {
  uint32_t max_bounces;
  float firefly_clamp;   /* 0 disables (uClamp) */
  uint32_t flags;        /* LUCENT_TRACE_* */
  uint32_t waves_per_cu; /* persistent launch width; 0 = 4 x occupancy */
  uint32_t occupancy;    /* persistent kernel variant: registers for 2..6 waves per SIMD; 0 = 4.
                            4..6 spill registers to scratch but hide more memory latency.
                            lucent_backward's replay takes 2..4 (0 = 4, measured fastest).
                            Results do not depend on it (lucent_autotune_tiles picks one). */
  /* Shape of the persistent per-sample hand-out (one film, and batches):
   * a job is a tile_w x tile_h tile of pixels (at most 64) with all its
   * samples; a wave shades once shade_batch of its 64 lanes wait and runs up
   * to trace_turns traversal steps between its checks for shading and new
   * work.  0 = 32, 8 for the last two.  tile_w = tile_h = 0 sizes each
   * film's tile to its sample count: the largest of 8x8, 8x4, 4x4, 4x2,
   * 2x2, 2x1, 1x1 whose pixels x samples stay within job_paths (0 = 256),
   * at least one pixel; a 0 beside a non-zero side is 8.  Results do not
   * depend on them (lucent_autotune_launch picks them with the occupancy).
   * lucent_backward hands out its replays by tiles of tile_w x tile_h
   * (0 = 8) and ignores job_paths. */
  uint32_t tile_w, tile_h, shade_batch, trace_turns;
  uint32_t job_paths;
} lucent_render_settings;

lucent_status lucent_trace_samples(lucent_ctx *ctx, lucent_scene *scene, lucent_film *film,
                                   const lucent_camera *camera, const lucent_render_settings *settings,
                                   uint32_t count, void *stream);

/* The first sample after a camera move with temporal reuse (replaces the
 * uReproject = 1 draw of Renderer.traceSamples, renderer.ts:441-451 and
 * 715-730, and trace.frag:274-355): one sample of every pixel of `film`
 * (seed film->seed, camera `camera`) whose running means restart from
 * `prev`'s, reprojected from `prev_camera`'s view by a bilinear fetch over
 * the previous pixels that saw the same surface (or the same sky), with the
 * pixel's history capped at history_cap times the primary surface's reuse
 * factor (uHistoryCap; the reference uses 8).  Disoccluded pixels start
 * over.  `prev` is typically `film` itself (its targets are first copied to
 * the context's scratch memory) but may be another film of any size, e.g.
 * the previous preview resolution.  Afterwards film's sample count is
 * min(prev's, history_cap) + 1 (renderer.ts:445).  Asynchronous on `stream`. */
lucent_status lucent_trace_reprojected(lucent_ctx *ctx, lucent_scene *scene, lucent_film *film,
                                       const lucent_camera *camera, const lucent_film *prev,
                                       const lucent_camera *prev_camera, float history_cap,
                                       const lucent_render_settings *settings, void *stream);

/* ------------------------------------------------------------------------
 * Batched multi-scene tracing: for every item, `samples` samples of every
 * pixel of its film (seeds continue from the film's seed, history advances by
 * `samples`), all items in ONE persistent launch whose waves pull tile
 * jobs (each film's tile sized by the launch shape) from a shared queue,
 * longest jobs first.  The result for each film is
 * bit-identical to lucent_trace_samples(scene, film, camera, settings,
 * samples).  A film may appear once per batch.  No counterpart in the
 * reference, which renders one scene per Renderer with one draw per sample
 * (renderer.ts:279, 702-737).
 * ---------------------------------------------------------------------- */
typedef struct lucent_batch_item {
  lucent_scene *scene;
  lucent_film *film;
  lucent_camera camera;
  uint32_t samples;
} lucent_batch_item;

lucent_status lucent_trace_batch(lucent_ctx *ctx, const lucent_batch_item *items, uint32_t count,
                                 const lucent_render_settings *settings, void *stream);

/* Autotuner of the persistent launch: runs min(trial_samples, samples) samples
 * of every item (at least 1) twice with each register/occupancy variant
 * (2..6 waves per SIMD) into scratch films of the items' sizes (the items'
 * films are not touched), after one untimed round, and returns the fastest
 * variant by its better time for lucent_render_settings.occupancy (a higher
 * occupancy must be more than 1 % faster than a lower one: its register
 * spills cost memory traffic that a near-tie does not repay).  ms_out, if
 * not NULL, receives the 5 better times (variants 2..6).  The variants give identical films.
 * No counterpart in the reference. */
lucent_status lucent_autotune_tiles(lucent_ctx *ctx, const lucent_batch_item *items, uint32_t count,
                                    const lucent_render_settings *settings, uint32_t trial_samples,
                                    uint32_t *occupancy_out, float *ms_out);

/* The launch autotuner: searches the persistent launch's register variant
 * (occupancy 2..6), then its job tile (sized per film for 128, 256, 512
 * or 1024 paths per job, or one of 8x8, 8x4, 4x4, 4x2, 2x2 for all), shade
 * batch (16, 24, 32, 40, 48 lanes) and traversal turns per check (4, 8,
 * 12, 16), one parameter at a time with the others at their best so far,
 * each candidate timed as lucent_autotune_tiles times a variant (best of two
 * runs of min(trial_samples, samples) samples of every item after an
 * untimed round, in scratch films; sized tiles are chosen for `samples`).
 * A candidate replaces the current choice only when more than 1 % faster.
 * Writes the choice into *settings (occupancy, tile_w, tile_h,
 * shade_batch, trace_turns, job_paths; its other fields
 * are read: flags, depth, clamp, waves_per_cu) and, if trials is not NULL,
 * up to trials_cap timed candidates in order; *n_trials receives their
 * number.  Every choice gives identical films.  No counterpart in the
 * reference. */
typedef struct lucent_autotune_trial {
  uint32_t occupancy, tile_w, tile_h, shade_batch, trace_turns;
  float ms;
  uint32_t job_paths;
} lucent_autotune_trial;
lucent_status lucent_autotune_launch(lucent_ctx *ctx, const lucent_batch_item *items, uint32_t count,
                                     lucent_render_settings *settings, uint32_t trial_samples,
                                     lucent_autotune_trial *trials, uint32_t trials_cap, uint32_t *n_trials);

/* ------------------------------------------------------------------------
 * Denoiser — Renderer.runDenoiser (src/renderer.ts:788-830): variance.frag,
 * then `iterations` à-trous passes (atrous.frag) with steps 1, 2, 4, ...
 * iterations = 0 picks atrousLevels(samples) (renderer.ts:140-142).  The
 * reference's parameters are sigma_l = 4, sigma_n = 64, sigma_p = 1.5.
 * The result (demodulated illumination, variance) stays in the film's filter
 * target; lucent_film_read_denoised copies it out, together with the
 * remodulated radiance illumination * max(albedo, 1e-3) (display.frag:32-40).
 * ---------------------------------------------------------------------- */
typedef struct lucent_denoise_params {
  float sigma_l, sigma_n, sigma_p;
  uint32_t iterations;
} lucent_denoise_params;

lucent_status lucent_denoise(lucent_ctx *ctx, lucent_film *film, const lucent_camera *camera,
                             const lucent_denoise_params *params, void *stream);
/* Either pointer may be NULL; each receives 4 floats per pixel. */
lucent_status lucent_film_read_denoised(lucent_ctx *ctx, const lucent_film *film, float *illum_var4,
                                        float *radiance4);

/* ------------------------------------------------------------------------
 * Presentation (display.frag:1-68, Renderer.render's display draw,
 * renderer.ts:572-593): the film's running mean, blended with the denoised
 * illumination times max(albedo, 1e-3) by denoise_weight where covered,
 * Catmull-Rom upsampled when the view is larger than the film (the moving
 * preview), times exposure, tone mapped, sRGB encoded, dithered with
 * interleaved gradient noise and stored as RGBA8 (clamp, x255, round to
 * nearest even).  Rows bottom-up, like the film.
 * ---------------------------------------------------------------------- */
#define LUCENT_TONEMAP_ACES 0u
#define LUCENT_TONEMAP_REINHARD 1u
#define LUCENT_TONEMAP_NONE 2u

typedef struct lucent_display_params {
  float denoise_weight; /* [0, 1]; > 0 needs lucent_denoise first (denoiseWeight, renderer.ts:143-144) */
  float exposure;       /* settings.exposure, default 1 */
  uint32_t tonemap;     /* LUCENT_TONEMAP_* (settings.tonemap, default ACES) */
} lucent_display_params;

/* Into device memory d_rgba8 (4 * view_w * view_h bytes); asynchronous on `stream`. */
lucent_status lucent_display(lucent_ctx *ctx, const lucent_film *film, const lucent_display_params *params,
                             uint32_t view_w, uint32_t view_h, uint8_t *d_rgba8, void *stream);
/* Renderer.screenshot (renderer.ts:856-862) without the PNG encoding: the
 * displayed image into host memory rgba8; synchronous on the context's
 * stream (uses the context's scratch memory). */
lucent_status lucent_screenshot(lucent_ctx *ctx, const lucent_film *film, const lucent_display_params *params,
                                uint32_t view_w, uint32_t view_h, uint8_t *rgba8);

/* ------------------------------------------------------------------------
 * Asset ingestion.  lucent_decode_hdr replaces parseHDR
 * (src/loaders/hdr.ts:7-93): a Radiance RGBE file (flat or run-length
 * encoded scanlines, any axis order) in memory -> width, height and RGB
 * floats with rows top to bottom, the environment-map layout
 * (HDRImage, environment.ts:6-10).  Pass rgb = NULL to query the size
 * only; otherwise rgb_capacity (floats) must hold width * height * 3.
 * Errors (LUCENT_E_INVALID) carry parseHDR's messages.  Host-only, no
 * context needed.
 * ---------------------------------------------------------------------- */
lucent_status lucent_decode_hdr(const uint8_t *bytes, size_t length, uint32_t *width, uint32_t *height, float *rgb,
                                size_t rgb_capacity);

/* glTF 2.0 / GLB ingestion; together these replace loadGLTF
 * (src/loaders/gltf.ts:188-502).  The reference's loader awaits a fetch per
 * external file; a C caller instead runs three phases:
 *   1. lucent_gltf_parse reads a .glb or .gltf file in memory, checks the
 *      version and the required extensions (KHR_draco_mesh_compression,
 *      EXT_meshopt_compression, KHR_texture_basisu, MSFT_texture_dds and
 *      EXT_mesh_gpu_instancing are refused, gltf.ts:38-44), converts the
 *      materials and lists the external files the asset needs: its buffers,
 *      then the images some material samples, in first-use order
 *      (lucent_gltf_resource_count / _uri; URIs as written, not decoded);
 *   2. the caller hands each one's bytes to lucent_gltf_set_resource;
 *   3. lucent_gltf_build walks the default scene (or, without scenes, every
 *      root node) and produces one world-space triangle soup per triangle
 *      primitive (modes 4, 5 and 6; positions, normals by the inverse
 *      transpose, flat normals where NORMAL is missing, TEXCOORD_0), with
 *      the winding kept under mirroring transforms.
 * Materials come back in the material texel layout (40 values per
 * material, the order of pack_materials / material.ts:132-149) as doubles,
 * so that factors read from the JSON are exact; map fields index the image
 * list.  A primitive without a material uses an appended glTF default
 * material.  Images come back still encoded (PNG, JPEG, WebP), as the
 * reference leaves decoding to the browser.  Warnings (ignored material
 * extensions, skipped point/line primitives, KTX2 textures, TEXCOORD_1 use)
 * are listed once each.  Errors: LUCENT_E_INVALID with gltf.ts's messages
 * for malformed files, LUCENT_E_UNSUPPORTED for refused features.  Host
 * only, no context needed; pointers returned stay valid until
 * lucent_gltf_destroy. */
typedef struct lucent_gltf lucent_gltf;
typedef struct {
  uint32_t mesh_count, material_count, image_count, warning_count;
  uint64_t triangle_count;
  double bounds[6]; /* min xyz, max xyz of the world-space positions */
} lucent_gltf_summary;
lucent_status lucent_gltf_parse(const uint8_t *bytes, size_t length, lucent_gltf **out);
uint32_t lucent_gltf_resource_count(const lucent_gltf *doc);
const char *lucent_gltf_resource_uri(const lucent_gltf *doc, uint32_t i);
lucent_status lucent_gltf_set_resource(lucent_gltf *doc, uint32_t i, const uint8_t *bytes, size_t length);
lucent_status lucent_gltf_build(lucent_gltf *doc);
lucent_status lucent_gltf_info(const lucent_gltf *doc, lucent_gltf_summary *out);
/* Mesh i: tri_count triangles of 3 corners; positions and normals hold
 * 9 floats per triangle, uvs 6 (NULL when the primitive has no TEXCOORD_0). */
lucent_status lucent_gltf_mesh(const lucent_gltf *doc, uint32_t i, uint32_t *tri_count, int32_t *material,
                               const float **positions, const float **normals, const float **uvs);
lucent_status lucent_gltf_material(const lucent_gltf *doc, uint32_t i, double *packed40);
lucent_status lucent_gltf_image(const lucent_gltf *doc, uint32_t i, const uint8_t **data, size_t *length,
                                const char **mime_type, const char **name);
const char *lucent_gltf_warning(const lucent_gltf *doc, uint32_t i);
void lucent_gltf_destroy(lucent_gltf *doc);

/* ------------------------------------------------------------------------
 * Frame graphs: one progressive frame (`samples` samples of every pixel of
 * `film` through the persistent tile launch, then optionally lucent_denoise
 * with *denoise and lucent_display with *display into d_rgba8) captured
 * once into a HIP graph and replayed by lucent_frame_graph_launch with one
 * launch call.  The samples' seeds continue on the device (the graph
 * advances a device copy of the film's seed), so n replays leave the film
 * and the display buffer bit-identical to n calls of lucent_trace_samples
 * and lucent_display, and the denoised image bit-identical to n calls of
 * lucent_denoise with the captured iteration count.  Camera, settings,
 * denoise iterations and display parameters are fixed at capture; with
 * iterations = 0 the count captured is atrousLevels of the sample count
 * after the first replay, and it stays that count on later replays, where
 * eager calls with iterations = 0 would re-evaluate it once the sample
 * count crosses 4, 16, 64 or 256 (pass an explicit count to match eager
 * calls across those thresholds).  The graph
 * holds the scene's, film's and output's device addresses.  Re-create it
 * after moving the camera, after lucent_scene_set_lighting /
 * lucent_scene_set_textures, or when any of them is destroyed.  A replay
 * fails if the film was advanced by other calls since the last replay.
 * No counterpart in the reference (one draw per sample and pass,
 * renderer.ts:552-600, 702-737).
 * ---------------------------------------------------------------------- */
typedef struct lucent_frame_graph lucent_frame_graph;

lucent_status lucent_frame_graph_create(lucent_ctx *ctx, lucent_scene *scene, lucent_film *film,
                                        const lucent_camera *camera, const lucent_render_settings *settings,
                                        uint32_t samples, const lucent_denoise_params *denoise,
                                        const lucent_display_params *display, uint32_t view_w, uint32_t view_h,
                                        uint8_t *d_rgba8, lucent_frame_graph **out);
/* Enqueues one replay on `stream` (NULL: the context's stream). */
lucent_status lucent_frame_graph_launch(lucent_frame_graph *graph, void *stream);
void lucent_frame_graph_destroy(lucent_frame_graph *graph);

/* ------------------------------------------------------------------------
 * Batch graphs: one lucent_trace_batch over `items` (each film `samples`
 * more samples, all films in one persistent launch) captured once into a
 * HIP graph, with its own job table, record rings and spill area in device
 * memory, and replayed by lucent_batch_graph_launch with one launch call:
 * no table upload and no host synchronisation per iteration.  Each film's
 * seed continues on the device (the graph advances a device copy per film),
 * so n replays leave every film bit-identical to n lucent_trace_batch calls
 * (and to lucent_trace_samples per film).  Cameras and settings are fixed at
 * capture; the graph holds the scenes' and films' device addresses:
 * re-create it after a camera move, a change of lighting, textures or
 * materials, or when any of them is destroyed.  A replay fails if a film
 * was advanced by other calls since the last replay.  Settings with
 * LUCENT_TRACE_PER_SAMPLE or LUCENT_TRACE_WAVEFRONT are refused.  No
 * counterpart in the reference (one scene per Renderer, one draw per
 * sample, renderer.ts:279, 702-737).
 * ---------------------------------------------------------------------- */
typedef struct lucent_batch_graph lucent_batch_graph;

lucent_status lucent_batch_graph_create(lucent_ctx *ctx, const lucent_batch_item *items, uint32_t count,
                                        const lucent_render_settings *settings, lucent_batch_graph **out);
/* Enqueues one replay on `stream` (NULL: the context's stream). */
lucent_status lucent_batch_graph_launch(lucent_batch_graph *graph, void *stream);
void lucent_batch_graph_destroy(lucent_batch_graph *graph);

/* ------------------------------------------------------------------------
 * Differentiable rendering.  The forward pass is lucent_trace_samples; the
 * backward pass replays each of the `spp` paths per pixel (seeds seed0 ..
 * seed0+spp-1, the seeds the forward used) from its RNG seed and returns the
 * gradient of sum_pixels dL_dI · mean_radiance with respect to each
 * material's base colour (material texel 0 rgb), roughness (texel 1 x) and
 * metallic (texel 0 w) and to every environment-map texel channel.
 * Sampling probabilities, MIS weights, Russian roulette and the environment
 * sampler are held constant (the estimator is unbiased for any of them).
 * No counterpart in the reference (which has no differentiation); the
 * parameters are those of material.ts:117-131 and lights.glsl:45-47.
 * d_sample_radiance: the record of lucent_forward for these seeds, or NULL
 * to re-trace each primal path.
 * All pointers are device pointers; grads are overwritten (not accumulated).
 * ---------------------------------------------------------------------- */
lucent_status lucent_backward(lucent_ctx *ctx, lucent_scene *scene, const lucent_camera *camera,
                              const lucent_render_settings *settings, uint32_t width, uint32_t height,
                              uint32_t spp, uint32_t seed0, const float *d_dl_di, const float *d_sample_radiance,
                              double *d_grad_materials, double *d_grad_env, void *stream);

/* The forward pass for lucent_backward: resets the film, sets its seed to
 * seed0 and traces spp samples of every pixel (lucent_trace_samples with the
 * persistent tile kernel, clamp 0), and, if d_sample_radiance is not NULL,
 * records each path's radiance there (device, float4 per pixel per sample,
 * sample k at k * width * height; 16 * spp * width * height bytes).  Passing
 * that record to lucent_backward (same scene parameters, seeds and size)
 * spares it re-tracing every primal path.  Asynchronous on `stream`. */
lucent_status lucent_forward(lucent_ctx *ctx, lucent_scene *scene, lucent_film *film, const lucent_camera *camera,
                             const lucent_render_settings *settings, uint32_t spp, uint32_t seed0,
                             float *d_sample_radiance, void *stream);

/* The estimator lucent_backward differentiates, evaluated at other material
 * parameters: the samples of seeds seed0 .. seed0+spp-1 take every random
 * decision, direction, pdf, MIS weight and roulette probability from the
 * scene's current materials, and BSDF values, refraction tints, absorption
 * and emission from `value_materials` (host, 40 floats per material, `count`
 * = the scene's material count).  With value_materials equal to the scene's
 * table this is the lucent_trace_samples image of those seeds (clamp 0) up to
 * rounding; its derivative with respect to value_materials at that point is
 * what lucent_backward returns, so central differences of it check the
 * adjoint sample by sample.  d_out_rgba: device, float4 per pixel (rows
 * bottom-up), receives the mean over the spp samples.  No counterpart in the
 * reference. */
lucent_status lucent_detached_forward(lucent_ctx *ctx, lucent_scene *scene, const lucent_camera *camera,
                                      const lucent_render_settings *settings, uint32_t width, uint32_t height,
                                      uint32_t spp, uint32_t seed0, const float *value_materials, uint32_t count,
                                      float *d_out_rgba, void *stream);

/* Replaces the scene's material table (40 floats per material, the layout
 * of material.ts:117-131; the count must not change). */
lucent_status lucent_scene_set_materials(lucent_ctx *ctx, lucent_scene *scene, const float *materials,
                                         uint32_t material_count);

/* Replaces the environment map's texels (same size), keeping the importance
 * sampler built from the map given to lucent_scene_set_lighting (rebuild it
 * with lucent_scene_rebuild_env_sampler). */
lucent_status lucent_scene_update_env_texels(lucent_ctx *ctx, lucent_scene *scene, const float *rgb,
                                             uint32_t width, uint32_t height);

/* Rebuilds the scene's environment importance sampler from the map now on
 * the device (after lucent_scene_update_env_texels), in place; light
 * sampling switches off if the map became black (total 0).  on_device = 0
 * copies the map to the host, runs the host build (lucent_env_sampler) and
 * uploads the table: 7 ms for a 1024x512 map.  on_device = 1 runs
 * lucent_env_sampler_gpu on the map in HBM, same table bits, no host
 * involvement, but bound by its sequential FP64 chains (112 ms at
 * 1024x512).  on_device = 2 runs lucent_env_sampler_gpu_parallel: the same
 * distribution with other alias pairs, in parallel.  Synchronous on the context's stream; traces on other streams
 * must not overlap it.  total_out may be NULL.  The reference rebuilds on
 * the CPU when a map is uploaded (renderer.ts:389-406). */
lucent_status lucent_scene_rebuild_env_sampler(lucent_ctx *ctx, lucent_scene *scene, uint32_t on_device,
                                               double *total_out);

/* Environment map size of the scene's lighting (0 before lighting is set). */
lucent_status lucent_scene_env_size(const lucent_scene *scene, uint32_t *width, uint32_t *height);

#ifdef __cplusplus
}
#endif
#endif /* LUCENT_HIP_H */
