// C-ABI entry points of the rendering path: lighting upload, lighting tables,
// the accumulation film and progressive tracing.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "adjoint.hpp"
#include "denoise.hpp"
#include "display.hpp"
#include "render_host.hpp"
#include "scene.hpp"
#include "wavefront.hpp"

using namespace lucent;

struct lucent_film {
  Context* ctx = nullptr;
  uint32_t width = 0, height = 0;
  uint32_t samples = 0;  // Renderer.frame
  uint32_t seed = 0;     // Renderer.seed
  void* base = nullptr;
  PtFilm f;
  void* filter = nullptr;       // two RGBA32F filter targets, the remodulated RGBA32F target, 16 B/px of guides, two variance planes
  float* denoised = nullptr;    // the target holding the last denoise result
  float* remodulated = nullptr;
};

namespace {

Status ensure_sheen(Context* ctx) {
  if (ctx->sheen_lut) return LUCENT_OK;
  std::vector<float> lut(kSheenLutSize * kSheenLutSize);
  sheen_lut(kSheenLutSize, 96, lut.data());
  for (float& v : lut) v = round_to_half(v);
  void* p = nullptr;
  Status s = dev_alloc(&p, lut.size() * sizeof(float));
  if (s != LUCENT_OK) return s;
  if ((s = dev_upload(p, lut.data(), lut.size() * sizeof(float), ctx->stream)) != LUCENT_OK) return s;
  if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
  ctx->sheen_lut = static_cast<float*>(p);
  return LUCENT_OK;
}

Status upload_env(Scene* sc, const void* src, size_t bytes, const float** dst) {
  void* p = nullptr;
  Status s = dev_alloc(&p, bytes);
  if (s != LUCENT_OK) return s;
  sc->env_allocations.push_back(p);
  if ((s = dev_upload(p, src, bytes, sc->ctx->stream)) != LUCENT_OK) return s;
  *dst = static_cast<const float*>(p);
  return LUCENT_OK;
}

double deg(double d) { return d * M_PI / 180; }

}  // namespace

extern "C" {

lucent_status lucent_alias_table(const double* weights, uint32_t n, float* prob_out, uint32_t* alias_out,
                                 double* pmf_out, double* total_out) {
  clear_error();
  if (n > 0 && (!weights || !prob_out || !alias_out || !pmf_out)) return fail(LUCENT_E_INVALID, "null array");
  return build_alias_table(weights, n, prob_out, alias_out, pmf_out, total_out);
}

lucent_status lucent_env_sampler(const float* rgb, uint32_t width, uint32_t height, uint32_t max_width,
                                 float* table_out, uint32_t* grid_w, uint32_t* grid_h, double* total) {
  clear_error();
  EnvSampler es;
  Status s = build_env_sampler(rgb, width, height, max_width, &es);
  if (s != LUCENT_OK) return s;
  if (table_out) std::memcpy(table_out, es.table.data(), es.table.size() * sizeof(float));
  if (grid_w) *grid_w = es.grid_w;
  if (grid_h) *grid_h = es.grid_h;
  if (total) *total = es.total;
  return LUCENT_OK;
}

lucent_status lucent_sheen_lut(uint32_t size, uint32_t steps, double* out) {
  clear_error();
  if (size < 2 || steps < 2 || !out) return fail(LUCENT_E_INVALID, "size and steps must be >= 2");
  for (uint32_t y = 0; y < size; ++y)
    for (uint32_t x = 0; x < size; ++x)
      out[y * size + x] = sheen_albedo(std::max(1e-3, double(x) / (size - 1)), double(y) / (size - 1), steps);
  return LUCENT_OK;
}

lucent_status lucent_scene_set_lighting(lucent_ctx* ctx, lucent_scene* scene, const lucent_lighting* L) {
  clear_error();
  if (!ctx || !scene || !L) return fail(LUCENT_E_INVALID, "null ctx/scene/lighting");
  Status s = ensure_sheen(ctx);
  if (s != LUCENT_OK) return s;
  PtScene& pt = scene->pt;
  const lucent_sun& sun = L->sun;
  const lucent_environment& env = L->env;
  // renderer.ts:355-370: sun uniforms; the analytic sun only without a map.
  const double el = deg(sun.elevation), az = deg(sun.azimuth);
  const double c = std::cos(el);
  const double dir[3] = {c * std::sin(az), std::sin(el), c * std::cos(az)};
  const double cos_angle = std::cos(deg(sun.angular_radius));
  const double solid = 2 * M_PI * (1 - cos_angle);
  const bool has_map = env.map_rgb != nullptr;
  pt.sun_enabled = (!has_map && sun.intensity > 0 && sun.elevation > -sun.angular_radius) ? 1 : 0;
  for (int i = 0; i < 3; ++i) {
    pt.sun_dir[i] = float(dir[i]);
    pt.sun_radiance[i] = float(sun.color[i] * (sun.intensity / solid));
  }
  pt.sun_cos = float(cos_angle);
  pt.sun_pdf = float(1 / solid);
  // Environment map: the HDRI, or the procedural sky baked at 1024x512.
  std::vector<float> baked;
  const float* rgb = env.map_rgb;
  uint32_t w = env.width, h = env.height;
  if (!has_map) {
    bake_sky(L->sky, sun, 1024, &baked);
    rgb = baked.data();
    w = 1024;
    h = 512;
  } else if (w == 0 || h == 0) {
    return fail(LUCENT_E_INVALID, "environment map has zero size");
  }
  EnvSampler es;
  if ((s = build_env_sampler(rgb, w, h, 1024, &es)) != LUCENT_OK) return s;
  hipStreamSynchronize(static_cast<hipStream_t>(ctx->stream));
  for (void* p : scene->env_allocations) hipFree(p);
  scene->env_allocations.clear();
  std::vector<float> map4(size_t(w) * h * 4);
  for (size_t i = 0; i < size_t(w) * h; ++i) {
    map4[i * 4] = rgb[i * 3];
    map4[i * 4 + 1] = rgb[i * 3 + 1];
    map4[i * 4 + 2] = rgb[i * 3 + 2];
    map4[i * 4 + 3] = 0.0f;
  }
  const float *dm = nullptr, *dt = nullptr;
  if ((s = upload_env(scene, map4.data(), map4.size() * sizeof(float), &dm)) != LUCENT_OK) return s;
  if ((s = upload_env(scene, es.table.data(), es.table.size() * sizeof(float), &dt)) != LUCENT_OK) return s;
  if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
  pt.env_map = dm;
  pt.env_w = int32_t(w);
  pt.env_h = int32_t(h);
  pt.env_table = dt;
  pt.grid_w = int32_t(es.grid_w);
  pt.grid_h = int32_t(es.grid_h);
  pt.env_sampling = es.total > 0 ? 1 : 0;
  // renderer.ts:372-379: the baked sky is in world orientation at unit intensity.
  const double rot = has_map ? deg(env.rotation) : 0.0;
  pt.env_cos = float(std::cos(rot));
  pt.env_sin = float(std::sin(rot));
  pt.env_intensity = float(has_map ? env.intensity : 1.0);
  pt.env_background = env.background ? 1 : 0;
  for (int i = 0; i < 3; ++i) pt.backdrop[i] = float(env.backdrop[i]);
  pt.sheen_lut = ctx->sheen_lut;
  scene->lighting_set = true;
  return LUCENT_OK;
}

lucent_status lucent_scene_set_textures(lucent_ctx* ctx, lucent_scene* scene, uint32_t size,
                                       const uint32_t* color_rgba8, uint32_t color_layers, const uint32_t* data_rgba8,
                                       uint32_t data_layers) {
  clear_error();
  if (!ctx || !scene) return fail(LUCENT_E_INVALID, "null ctx/scene");
  if ((color_layers && !color_rgba8) || (data_layers && !data_rgba8)) return fail(LUCENT_E_INVALID, "null texel array");
  if ((color_layers || data_layers) && (size == 0 || size > 16384))
    return fail(LUCENT_E_INVALID, "texture size must be in 1..16384");
  const size_t layer = size_t(size) * size * sizeof(uint32_t);
  if ((size_t(color_layers) + data_layers) * layer > (size_t(1) << 34))
    return fail(LUCENT_E_INVALID, "texture arrays larger than 16 GiB");
  Status s = stream_sync(ctx->stream);
  if (s != LUCENT_OK) return s;
  for (void* p : scene->tex_allocations) hipFree(p);
  scene->tex_allocations.clear();
  PtScene& pt = scene->pt;
  pt.color_maps = pt.data_maps = nullptr;
  pt.tex_size = 0;
  pt.color_layers = pt.data_layers = 0;
  auto upload = [&](const uint32_t* src, uint32_t layers, const uint32_t** dst) -> Status {
    if (layers == 0) return LUCENT_OK;
    void* p = nullptr;
    Status st = dev_alloc(&p, layer * layers);
    if (st != LUCENT_OK) return st;
    scene->tex_allocations.push_back(p);
    if ((st = dev_upload(p, src, layer * layers, ctx->stream)) != LUCENT_OK) return st;
    *dst = static_cast<const uint32_t*>(p);
    return LUCENT_OK;
  };
  const uint32_t *dc = nullptr, *dd = nullptr;
  if ((s = upload(color_rgba8, color_layers, &dc)) != LUCENT_OK) return s;
  if ((s = upload(data_rgba8, data_layers, &dd)) != LUCENT_OK) return s;
  if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
  pt.color_maps = dc;
  pt.data_maps = dd;
  pt.color_layers = int32_t(color_layers);
  pt.data_layers = int32_t(data_layers);
  pt.tex_size = (color_layers || data_layers) ? int32_t(size) : 0;
  return LUCENT_OK;
}

lucent_status lucent_film_create(lucent_ctx* ctx, uint32_t width, uint32_t height, lucent_film** out) {
  clear_error();
  if (!ctx || !out) return fail(LUCENT_E_INVALID, "null ctx/out");
  *out = nullptr;
  if (width == 0 || height == 0 || uint64_t(width) * height > (1ull << 28))
    return fail(LUCENT_E_INVALID, "film size must be positive and at most 2^28 pixels");
  const size_t px = size_t(width) * height;
  const size_t bytes = px * (16 + 8 + 8 + 4);
  void* base = nullptr;
  Status s = dev_alloc(&base, bytes);
  if (s != LUCENT_OK) return s;
  auto* f = new lucent_film();
  f->ctx = ctx;
  f->width = width;
  f->height = height;
  f->base = base;
  char* b = static_cast<char*>(base);
  f->f.color = reinterpret_cast<float*>(b);
  f->f.albedo = reinterpret_cast<uint16_t*>(b + px * 16);
  f->f.normal = reinterpret_cast<uint16_t*>(b + px * 24);
  f->f.history = reinterpret_cast<float*>(b + px * 32);
  if ((s = lucent_film_reset(ctx, f)) != LUCENT_OK) {
    lucent_film_destroy(f);
    return s;
  }
  *out = f;
  return LUCENT_OK;
}

void lucent_film_destroy(lucent_film* film) {
  if (!film) return;
  hipStreamSynchronize(static_cast<hipStream_t>(film->ctx->stream));
  hipFree(film->base);
  if (film->filter) hipFree(film->filter);
  delete film;
}

lucent_status lucent_film_reset(lucent_ctx* ctx, lucent_film* film) {
  clear_error();
  if (!ctx || !film) return fail(LUCENT_E_INVALID, "null ctx/film");
  const size_t bytes = size_t(film->width) * film->height * 36;
  const hipError_t e = hipMemsetAsync(film->base, 0, bytes, static_cast<hipStream_t>(ctx->stream));
  if (e != hipSuccess) return fail_hip(e, "hipMemsetAsync(film)");
  film->samples = 0;
  return stream_sync(ctx->stream);
}

lucent_status lucent_film_set_seed(lucent_film* film, uint32_t seed) {
  if (!film) return fail(LUCENT_E_INVALID, "null film");
  film->seed = seed;
  return LUCENT_OK;
}

lucent_status lucent_film_info(const lucent_film* film, uint32_t* w, uint32_t* h, uint32_t* samples, uint32_t* seed) {
  if (!film) return fail(LUCENT_E_INVALID, "null film");
  if (w) *w = film->width;
  if (h) *h = film->height;
  if (samples) *samples = film->samples;
  if (seed) *seed = film->seed;
  return LUCENT_OK;
}

static float half_to_float(uint16_t h) {
  const uint32_t sign = uint32_t(h & 0x8000u) << 16;
  const uint32_t exp = (h >> 10) & 0x1Fu, man = h & 0x3FFu;
  float out;
  if (exp == 0) {
    out = std::ldexp(float(man), -24);
    if (sign) out = -out;
    return out;
  }
  uint32_t bits;
  if (exp == 31)
    bits = sign | 0x7F800000u | (man << 13);
  else
    bits = sign | ((exp + 112) << 23) | (man << 13);
  std::memcpy(&out, &bits, 4);
  return out;
}

lucent_status lucent_film_read(lucent_ctx* ctx, const lucent_film* film, float* color4, float* albedo4, float* normal4,
                               float* history) {
  clear_error();
  if (!ctx || !film) return fail(LUCENT_E_INVALID, "null ctx/film");
  const size_t px = size_t(film->width) * film->height;
  std::vector<uint16_t> tmp;
  Status s;
  if (color4 && (s = dev_download(color4, film->f.color, px * 16, ctx->stream)) != LUCENT_OK) return s;
  if (history && (s = dev_download(history, film->f.history, px * 4, ctx->stream)) != LUCENT_OK) return s;
  for (int k = 0; k < 2; ++k) {
    float* dst = k == 0 ? albedo4 : normal4;
    if (!dst) continue;
    tmp.resize(px * 4);
    if ((s = dev_download(tmp.data(), k == 0 ? film->f.albedo : film->f.normal, px * 8, ctx->stream)) != LUCENT_OK)
      return s;
    if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
    for (size_t i = 0; i < px * 4; ++i) dst[i] = half_to_float(tmp[i]);
  }
  return stream_sync(ctx->stream);
}

lucent_status lucent_film_device_ptrs(const lucent_film* film, void** color4, void** albedo, void** normal,
                                      void** history) {
  if (!film) return fail(LUCENT_E_INVALID, "null film");
  if (color4) *color4 = film->f.color;
  if (albedo) *albedo = film->f.albedo;
  if (normal) *normal = film->f.normal;
  if (history) *history = film->f.history;
  return LUCENT_OK;
}

static Status fill_frame(const lucent_camera* cam, const lucent_render_settings* settings, uint32_t w, uint32_t h,
                         PtFrame* f);

static Status check_traceable(const lucent_scene* scene) {
  if (!scene->lighting_set) return fail(LUCENT_E_INVALID, "call lucent_scene_set_lighting before tracing");
  if (scene->pt.material_count == 0 && scene->bvh.tri_count > 0)
    return fail(LUCENT_E_INVALID, "scene has triangles but no materials (use lucent_scene_upload)");
  return LUCENT_OK;
}

lucent_status lucent_trace_reprojected(lucent_ctx* ctx, lucent_scene* scene, lucent_film* film,
                                       const lucent_camera* cam, const lucent_film* prev,
                                       const lucent_camera* prev_cam, float history_cap,
                                       const lucent_render_settings* settings, void* stream) {
  clear_error();
  if (!ctx || !scene || !film || !cam || !prev || !prev_cam || !settings) return fail(LUCENT_E_INVALID, "null argument");
  if (!(history_cap >= 0.0f) || !std::isfinite(history_cap)) return fail(LUCENT_E_INVALID, "history_cap must be finite and >= 0");
  Status s = check_traceable(scene);
  if (s != LUCENT_OK) return s;
  void* st = stream ? stream : ctx->stream;
  PrevView pv;
  pv.width = prev->width;
  pv.height = prev->height;
  pv.history_cap = history_cap;
  pv.film = prev->f;
  if (prev == film) {
    // The sample reads neighbours of the previous means: trace from a copy.
    const size_t px = size_t(film->width) * film->height, bytes = px * (16 + 8 + 8 + 4);
    if ((s = ensure_scratch(ctx, bytes)) != LUCENT_OK) return s;
    const hipError_t e = hipMemcpyAsync(ctx->scratch, film->base, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(st));
    if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync(previous film)");
    char* b = static_cast<char*>(ctx->scratch);
    pv.film.color = reinterpret_cast<float*>(b);
    pv.film.albedo = reinterpret_cast<uint16_t*>(b + px * 16);
    pv.film.normal = reinterpret_cast<uint16_t*>(b + px * 24);
    pv.film.history = reinterpret_cast<float*>(b + px * 32);
  }
  double fwd[3], right[3], up[3], tan_half;
  camera_basis(*prev_cam, fwd, right, up, &tan_half);
  for (int i = 0; i < 3; ++i) {
    pv.cam_pos[i] = float(prev_cam->position[i]);
    pv.cam_fwd[i] = float(fwd[i]);
    pv.cam_right[i] = float(right[i]);
    pv.cam_up[i] = float(up[i]);
  }
  pv.tan_half_fov = float(tan_half);
  PtFrame f;
  fill_frame(cam, settings, film->width, film->height, &f);
  f.seed = film->seed;
  if ((s = launch_trace_reprojected(ctx, scene->pt, f, film->f, pv, st)) != LUCENT_OK) return s;
  film->seed++;
  const uint32_t cap = history_cap >= 4294967295.0f ? 4294967295u : uint32_t(history_cap);
  film->samples = std::min<uint32_t>(prev->samples, cap) + 1;
  return LUCENT_OK;
}

lucent_status lucent_trace_samples(lucent_ctx* ctx, lucent_scene* scene, lucent_film* film, const lucent_camera* cam,
                                   const lucent_render_settings* settings, uint32_t count, void* stream) {
  clear_error();
  if (!ctx || !scene || !film || !cam || !settings) return fail(LUCENT_E_INVALID, "null argument");
  Status s = check_traceable(scene);
  if (s != LUCENT_OK) return s;
  PtFrame f;
  fill_frame(cam, settings, film->width, film->height, &f);
  void* st = stream ? stream : ctx->stream;
  if (count == 0) return LUCENT_OK;
  if (settings->flags & LUCENT_TRACE_PER_SAMPLE) {
    for (uint32_t i = 0; i < count; ++i) {
      f.seed = film->seed++;
      if ((s = launch_trace_sample(ctx, scene->pt, f, film->f, st)) != LUCENT_OK) return s;
      film->samples++;
    }
    return LUCENT_OK;
  }
  TileItem it;
  it.scene = scene->pt;
  it.frame = f;
  it.frame.seed = film->seed;
  it.film = film->f;
  it.spp = count;
  if ((settings->flags & LUCENT_TRACE_WAVEFRONT) && !scene->blend_alpha) {
    WavefrontLaunch wl;
    wl.waves_per_cu = settings->waves_per_cu;
    if ((s = launch_wavefront(ctx, it, wl, st)) != LUCENT_OK) return s;
    film->seed += count;
    film->samples += count;
    return LUCENT_OK;
  }
  TileLaunch cfg;
  cfg.fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS) && !scene->blend_alpha;
  cfg.waves_per_cu = settings->waves_per_cu;
  cfg.occupancy = settings->occupancy;
  cfg.tile_w = settings->tile_w;
  cfg.tile_h = settings->tile_h;
  cfg.shade_batch = settings->shade_batch;
  cfg.trace_turns = settings->trace_turns;
  cfg.job_paths = settings->job_paths;
  if ((s = launch_trace_tiles(ctx, &it, 1, cfg, st)) != LUCENT_OK) return s;
  film->seed += count;
  film->samples += count;
  return LUCENT_OK;
}

}  // extern "C"

// Checks a batch (scenes traceable, no film twice) and turns it into launch
// items, longest jobs first (a job costs about its paths, tile pixels x
// samples, x the log of the scene size), with the launch shape of
// `settings`.  films[k] is the film of tiles[k]; items without samples are
// left out.
static Status batch_tiles(const lucent_batch_item* items, uint32_t count, const lucent_render_settings* settings,
                          TileLaunch* cfg, std::vector<TileItem>* tiles, std::vector<lucent_film*>* films) {
  for (uint32_t i = 0; i < count; ++i) {
    const lucent_batch_item& b = items[i];
    if (!b.scene || !b.film) return fail(LUCENT_E_INVALID, "batch item without scene or film");
    Status s = check_traceable(b.scene);
    if (s != LUCENT_OK) return s;
    for (uint32_t j = 0; j < i; ++j)
      if (items[j].film == b.film) return fail(LUCENT_E_INVALID, "a film appears twice in one batch");
  }
  bool fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS);
  for (uint32_t i = 0; i < count; ++i) fused = fused && !items[i].scene->blend_alpha;
  cfg->fused = fused;
  cfg->waves_per_cu = settings->waves_per_cu;
  cfg->occupancy = settings->occupancy;
  cfg->tile_w = settings->tile_w;
  cfg->tile_h = settings->tile_h;
  cfg->shade_batch = settings->shade_batch;
  cfg->trace_turns = settings->trace_turns;
  cfg->job_paths = settings->job_paths;
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; ++i) order[i] = i;
  auto job_cost = [&](uint32_t i) {
    const double tris = double(items[i].scene->bvh.tri_count) + 2.0;
    uint32_t tw = 0, th = 0;
    launch_job_tile(*cfg, items[i].samples, &tw, &th);
    return double(tw) * th * items[i].samples * std::log2(tris);
  };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return job_cost(a) > job_cost(b); });
  tiles->clear();
  films->clear();
  for (uint32_t i : order) {
    const lucent_batch_item& b = items[i];
    if (b.samples == 0) continue;
    TileItem it;
    it.scene = b.scene->pt;
    fill_frame(&b.camera, settings, b.film->width, b.film->height, &it.frame);
    it.frame.seed = b.film->seed;
    it.film = b.film->f;
    it.spp = b.samples;
    tiles->push_back(it);
    films->push_back(b.film);
  }
  return LUCENT_OK;
}

extern "C" {

lucent_status lucent_trace_batch(lucent_ctx* ctx, const lucent_batch_item* items, uint32_t count,
                                 const lucent_render_settings* settings, void* stream) {
  clear_error();
  if (!ctx || (!items && count) || !settings) return fail(LUCENT_E_INVALID, "null argument");
  void* st = stream ? stream : ctx->stream;
  TileLaunch cfg;
  std::vector<TileItem> tiles;
  std::vector<lucent_film*> films;
  Status s = batch_tiles(items, count, settings, &cfg, &tiles, &films);
  if (s != LUCENT_OK) return s;
  if (settings->flags & LUCENT_TRACE_PER_SAMPLE) {
    for (uint32_t i = 0; i < count; ++i) {
      lucent_status r = lucent_trace_samples(ctx, items[i].scene, items[i].film, &items[i].camera, settings,
                                             items[i].samples, st);
      if (r != LUCENT_OK) return r;
    }
    return LUCENT_OK;
  }
  if ((s = launch_trace_tiles(ctx, tiles.data(), uint32_t(tiles.size()), cfg, st)) != LUCENT_OK) return s;
  for (uint32_t i = 0; i < count; ++i) {
    items[i].film->seed += items[i].samples;
    items[i].film->samples += items[i].samples;
  }
  return LUCENT_OK;
}

lucent_status lucent_autotune_tiles(lucent_ctx* ctx, const lucent_batch_item* items, uint32_t count,
                                    const lucent_render_settings* settings, uint32_t trial_samples,
                                    uint32_t* occupancy_out, float* ms_out) {
  clear_error();
  if (!ctx || !items || count == 0 || !settings || !occupancy_out) return fail(LUCENT_E_INVALID, "null argument");
  if (trial_samples == 0) return fail(LUCENT_E_INVALID, "trial_samples must be positive");
  for (uint32_t i = 0; i < count; ++i) {
    if (!items[i].scene || !items[i].film) return fail(LUCENT_E_INVALID, "batch item without scene or film");
    Status s = check_traceable(items[i].scene);
    if (s != LUCENT_OK) return s;
  }
  hipStream_t st = static_cast<hipStream_t>(ctx->stream);
  // Scratch films of the items' sizes: the callers' films are not touched.
  std::vector<lucent_film*> films(count, nullptr);
  std::vector<TileItem> tiles;
  Status s = LUCENT_OK;
  hipEvent_t ev[2] = {nullptr, nullptr};
  float best_ms = 0.0f;
  uint32_t best = 0;
  for (uint32_t i = 0; i < count && s == LUCENT_OK; ++i) {
    s = lucent_film_create(ctx, items[i].film->width, items[i].film->height, &films[i]);
    if (s != LUCENT_OK) break;
    TileItem it;
    it.scene = items[i].scene->pt;
    fill_frame(&items[i].camera, settings, films[i]->width, films[i]->height, &it.frame);
    it.frame.seed = items[i].film->seed;
    it.film = films[i]->f;
    it.spp = std::max(1u, std::min(trial_samples, items[i].samples ? items[i].samples : trial_samples));
    it.job_spp = items[i].samples;  // sized tiles as for the real sample count
    tiles.push_back(it);
  }
  if (s == LUCENT_OK) {
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e != hipSuccess) s = fail_hip(e, "hipEventCreate");
  }
  bool fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS);
  for (uint32_t i = 0; i < count; ++i) fused = fused && !items[i].scene->blend_alpha;
  // Pass 0 loads every code object; each variant keeps the best of passes
  // 1 and 2 (a single short trial varies by a few percent).
  float trial_ms[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (int pass = 0; pass < 3 && s == LUCENT_OK; ++pass) {
    for (uint32_t occ = 2; occ <= 6 && s == LUCENT_OK; ++occ) {
      TileLaunch cfg;
      cfg.fused = fused;
      cfg.occupancy = occ;
      cfg.waves_per_cu = settings->waves_per_cu;
      hipError_t e = hipEventRecord(ev[0], st);
      if (e != hipSuccess) {
        s = fail_hip(e, "hipEventRecord");
        break;
      }
      if ((s = launch_trace_tiles(ctx, tiles.data(), uint32_t(tiles.size()), cfg, st)) != LUCENT_OK) break;
      e = hipEventRecord(ev[1], st);
      if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
      float ms = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
      if (e != hipSuccess) {
        s = fail_hip(e, "autotune timing");
        break;
      }
      if (pass == 1 || (pass == 2 && ms < trial_ms[occ - 2])) trial_ms[occ - 2] = ms;
    }
  }
  if (s == LUCENT_OK) {
    for (uint32_t occ = 2; occ <= 6; ++occ) {
      const float ms = trial_ms[occ - 2];
      if (ms_out) ms_out[occ - 2] = ms;
      // A higher occupancy must win by more than 1 %: its register spills
      // cost memory traffic that a near-tie does not repay.
      if (best == 0 || ms < best_ms * 0.99f) {
        best = occ;
        best_ms = ms;
      }
    }
  }
  for (hipEvent_t e : ev)
    if (e) hipEventDestroy(e);
  for (lucent_film* f : films) lucent_film_destroy(f);
  if (s != LUCENT_OK) return s;
  *occupancy_out = best;
  return LUCENT_OK;
}

lucent_status lucent_autotune_launch(lucent_ctx* ctx, const lucent_batch_item* items, uint32_t count,
                                     lucent_render_settings* settings, uint32_t trial_samples,
                                     lucent_autotune_trial* trials, uint32_t trials_cap, uint32_t* n_trials) {
  clear_error();
  if (!ctx || !items || count == 0 || !settings) return fail(LUCENT_E_INVALID, "null argument");
  if (trial_samples == 0) return fail(LUCENT_E_INVALID, "trial_samples must be positive");
  for (uint32_t i = 0; i < count; ++i) {
    if (!items[i].scene || !items[i].film) return fail(LUCENT_E_INVALID, "batch item without scene or film");
    Status s = check_traceable(items[i].scene);
    if (s != LUCENT_OK) return s;
  }
  if (n_trials) *n_trials = 0;
  hipStream_t st = static_cast<hipStream_t>(ctx->stream);
  std::vector<lucent_film*> films(count, nullptr);
  std::vector<TileItem> tiles;
  Status s = LUCENT_OK;
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (uint32_t i = 0; i < count && s == LUCENT_OK; ++i) {
    s = lucent_film_create(ctx, items[i].film->width, items[i].film->height, &films[i]);
    if (s != LUCENT_OK) break;
    TileItem it;
    it.scene = items[i].scene->pt;
    fill_frame(&items[i].camera, settings, films[i]->width, films[i]->height, &it.frame);
    it.frame.seed = items[i].film->seed;
    it.film = films[i]->f;
    it.spp = std::max(1u, std::min(trial_samples, items[i].samples ? items[i].samples : trial_samples));
    it.job_spp = items[i].samples;  // sized tiles as for the real sample count
    tiles.push_back(it);
  }
  if (s == LUCENT_OK) {
    hipError_t e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e != hipSuccess) s = fail_hip(e, "hipEventCreate");
  }
  bool fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS);
  for (uint32_t i = 0; i < count; ++i) fused = fused && !items[i].scene->blend_alpha;
  // One timed run of a candidate (the launch rewrites a copy of the items).
  auto run = [&](const TileLaunch& cfg, float* ms) -> Status {
    std::vector<TileItem> t = tiles;
    hipError_t e = hipEventRecord(ev[0], st);
    if (e != hipSuccess) return fail_hip(e, "hipEventRecord");
    Status r = launch_trace_tiles(ctx, t.data(), uint32_t(t.size()), cfg, st);
    if (r != LUCENT_OK) return r;
    e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ev[0], ev[1]);
    if (e != hipSuccess) return fail_hip(e, "autotune timing");
    return LUCENT_OK;
  };
  TileLaunch best;
  best.fused = fused;
  best.waves_per_cu = settings->waves_per_cu;
  best.occupancy = 4;
  best.tile_w = best.tile_h = 0;  // sized per film
  best.job_paths = kJobPaths;
  best.shade_batch = 32;
  best.trace_turns = 8;
  float best_ms = 0.0f;
  uint32_t logged = 0;
  // Best of two runs, after an untimed one the first time a variant runs.
  auto measure = [&](const TileLaunch& cfg, bool warm, float* out) -> Status {
    float a = 0.0f, b = 0.0f;
    Status r = LUCENT_OK;
    if (warm && (r = run(cfg, &a)) != LUCENT_OK) return r;
    if ((r = run(cfg, &a)) != LUCENT_OK) return r;
    if ((r = run(cfg, &b)) != LUCENT_OK) return r;
    *out = std::min(a, b);
    if (trials && logged < trials_cap)
      trials[logged] = lucent_autotune_trial{cfg.occupancy, cfg.tile_w, cfg.tile_h, cfg.shade_batch,
                                             cfg.trace_turns, *out, cfg.job_paths};
    ++logged;
    return LUCENT_OK;
  };
  // Stage 0: the register variant (each a code object of its own).
  for (uint32_t occ = 2; occ <= 6 && s == LUCENT_OK; ++occ) {
    TileLaunch c = best;
    c.occupancy = occ;
    float ms = 0.0f;
    if ((s = measure(c, true, &ms)) != LUCENT_OK) break;
    // A higher occupancy must win by more than 1 %: its register spills
    // cost memory traffic that a near-tie does not repay.
    if (occ == 2 || ms < best_ms * 0.99f) {
      best = c;
      best_ms = ms;
    }
  }
  // Stages 1-3: the hand-out shape, one parameter at a time.  Tiles: one
  // shape for every film, or (0 x 0) each film's tile sized to its sample
  // count for a number of paths per job.
  const uint32_t tiles_wh[][3] = {{0, 0, 128}, {0, 0, 256}, {0, 0, 512}, {0, 0, 1024}, {8, 8, 0},
                                  {8, 4, 0},   {4, 4, 0},   {4, 2, 0},   {2, 2, 0}};
  const uint32_t batches[] = {16, 24, 32, 40, 48};
  const uint32_t turns[] = {4, 8, 12, 16};
  auto consider = [&](const TileLaunch& c) -> Status {
    float ms = 0.0f;
    Status r = measure(c, false, &ms);
    if (r == LUCENT_OK && ms < best_ms * 0.99f) {
      best = c;
      best_ms = ms;
    }
    return r;
  };
  const uint32_t tile0[3] = {best.tile_w, best.tile_h, best.job_paths};
  for (const auto& wh : tiles_wh) {
    if (s != LUCENT_OK) break;
    if (wh[0] == tile0[0] && wh[1] == tile0[1] && wh[2] == tile0[2]) continue;  // timed already
    TileLaunch c = best;
    c.tile_w = wh[0];
    c.tile_h = wh[1];
    c.job_paths = wh[2];
    s = consider(c);
  }
  const uint32_t batch0 = best.shade_batch;
  for (uint32_t b : batches) {
    if (s != LUCENT_OK) break;
    if (b == batch0) continue;
    TileLaunch c = best;
    c.shade_batch = b;
    s = consider(c);
  }
  const uint32_t turns0 = best.trace_turns;
  for (uint32_t t : turns) {
    if (s != LUCENT_OK) break;
    if (t == turns0) continue;
    TileLaunch c = best;
    c.trace_turns = t;
    s = consider(c);
  }
  for (hipEvent_t e : ev)
    if (e) hipEventDestroy(e);
  for (lucent_film* f : films) lucent_film_destroy(f);
  if (s != LUCENT_OK) return s;
  settings->occupancy = best.occupancy;
  settings->tile_w = best.tile_w;
  settings->tile_h = best.tile_h;
  settings->shade_batch = best.shade_batch;
  settings->trace_turns = best.trace_turns;
  settings->job_paths = best.job_paths;
  if (n_trials) *n_trials = std::min(logged, trials ? trials_cap : 0u);
  return LUCENT_OK;
}

static Status fill_frame(const lucent_camera* cam, const lucent_render_settings* settings, uint32_t w, uint32_t h,
                         PtFrame* f) {
  f->width = w;
  f->height = h;
  f->max_bounces = settings->max_bounces;
  f->clamp = settings->firefly_clamp;
  double fwd[3], right[3], up[3], tan_half;
  camera_basis(*cam, fwd, right, up, &tan_half);
  for (int i = 0; i < 3; ++i) {
    f->cam_pos[i] = float(cam->position[i]);
    f->cam_fwd[i] = float(fwd[i]);
    f->cam_right[i] = float(right[i]);
    f->cam_up[i] = float(up[i]);
  }
  f->tan_half_fov = float(tan_half);
  f->aperture = float(cam->aperture);
  f->focus_dist = float(cam->focus_distance);
  return LUCENT_OK;
}

lucent_status lucent_forward(lucent_ctx* ctx, lucent_scene* scene, lucent_film* film, const lucent_camera* cam,
                             const lucent_render_settings* settings, uint32_t spp, uint32_t seed0,
                             float* d_sample_radiance, void* stream) {
  clear_error();
  if (!ctx || !scene || !film || !cam || !settings) return fail(LUCENT_E_INVALID, "null argument");
  if (spp == 0) return fail(LUCENT_E_INVALID, "zero samples");
  if (settings->firefly_clamp > 0.0f)
    return fail(LUCENT_E_UNSUPPORTED, "the firefly clamp is not differentiable; use firefly_clamp = 0");
  Status s = check_traceable(scene);
  if (s != LUCENT_OK) return s;
  film->seed = seed0;
  if ((s = lucent_film_reset(ctx, film)) != LUCENT_OK) return s;
  TileItem it;
  it.scene = scene->pt;
  fill_frame(cam, settings, film->width, film->height, &it.frame);
  it.frame.seed = seed0;
  it.film = film->f;
  it.spp = spp;
  it.sample_radiance = d_sample_radiance;
  TileLaunch cfg;
  cfg.fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS) && !scene->blend_alpha;
  cfg.waves_per_cu = settings->waves_per_cu;
  cfg.occupancy = settings->occupancy;
  cfg.tile_w = settings->tile_w;
  cfg.tile_h = settings->tile_h;
  cfg.shade_batch = settings->shade_batch;
  cfg.trace_turns = settings->trace_turns;
  cfg.job_paths = settings->job_paths;
  if ((s = launch_trace_tiles(ctx, &it, 1, cfg, stream ? stream : ctx->stream)) != LUCENT_OK) return s;
  film->seed += spp;
  film->samples += spp;
  return LUCENT_OK;
}

lucent_status lucent_backward(lucent_ctx* ctx, lucent_scene* scene, const lucent_camera* cam,
                              const lucent_render_settings* settings, uint32_t width, uint32_t height, uint32_t spp,
                              uint32_t seed0, const float* d_dl_di, const float* d_sample_radiance,
                              double* d_grad_materials, double* d_grad_env, void* stream) {
  clear_error();
  if (!ctx || !scene || !cam || !settings || !d_dl_di || !d_grad_materials) return fail(LUCENT_E_INVALID, "null argument");
  if (!scene->lighting_set) return fail(LUCENT_E_INVALID, "call lucent_scene_set_lighting before lucent_backward");
  if (width == 0 || height == 0 || spp == 0) return fail(LUCENT_E_INVALID, "empty image or zero samples");
  if (settings->firefly_clamp > 0.0f)
    return fail(LUCENT_E_UNSUPPORTED, "the firefly clamp is not differentiable; use firefly_clamp = 0");
  const PtScene& pt = scene->pt;
  const uint32_t env_texels = uint32_t(pt.env_w) * uint32_t(pt.env_h);
  const size_t need = adjoint_scratch_bytes(width, height, pt.material_count, env_texels);
  Status s = ensure_scratch(ctx, need);
  if (s != LUCENT_OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream ? stream : ctx->stream);
  hipError_t e = hipMemsetAsync(d_grad_materials, 0, size_t(pt.material_count) * 5 * sizeof(double), st);
  if (e == hipSuccess && d_grad_env) e = hipMemsetAsync(d_grad_env, 0, size_t(env_texels) * 3 * sizeof(double), st);
  if (e != hipSuccess) return fail_hip(e, "hipMemsetAsync(gradients)");
  AdjointArgs a;
  a.scene = &pt;
  fill_frame(cam, settings, width, height, &a.frame);
  a.spp = spp;
  a.seed0 = seed0;
  a.dl_di = d_dl_di;
  a.recorded_radiance = d_sample_radiance;
  a.occupancy = settings->occupancy;
  a.tile_w = settings->tile_w;
  a.tile_h = settings->tile_h;
  a.grad_materials = d_grad_materials;
  a.grad_env = d_grad_env;
  a.scratch = ctx->scratch;
  return launch_adjoint(ctx, a, st);
}

lucent_status lucent_detached_forward(lucent_ctx* ctx, lucent_scene* scene, const lucent_camera* cam,
                                      const lucent_render_settings* settings, uint32_t width, uint32_t height,
                                      uint32_t spp, uint32_t seed0, const float* value_materials, uint32_t count,
                                      float* d_out_rgba, void* stream) {
  clear_error();
  if (!ctx || !scene || !cam || !settings || !value_materials || !d_out_rgba) return fail(LUCENT_E_INVALID, "null argument");
  if (!scene->lighting_set) return fail(LUCENT_E_INVALID, "call lucent_scene_set_lighting before lucent_detached_forward");
  if (width == 0 || height == 0 || spp == 0) return fail(LUCENT_E_INVALID, "empty image or zero samples");
  if (count != scene->pt.material_count) return fail(LUCENT_E_INVALID, "material count differs from the uploaded scene");
  if (settings->firefly_clamp > 0.0f)
    return fail(LUCENT_E_UNSUPPORTED, "the detached estimator has no firefly clamp; use firefly_clamp = 0");
  const size_t bytes = size_t(count) * kMaterialTexels * 4 * sizeof(float);
  Status s = ensure_scratch(ctx, bytes);
  if (s != LUCENT_OK) return s;
  hipStream_t st = static_cast<hipStream_t>(stream ? stream : ctx->stream);
  if ((s = dev_upload(ctx->scratch, value_materials, bytes, st)) != LUCENT_OK) return s;
  DetachedArgs a;
  a.scene = &scene->pt;
  a.value_materials = static_cast<const float*>(ctx->scratch);
  fill_frame(cam, settings, width, height, &a.frame);
  a.spp = spp;
  a.seed0 = seed0;
  a.out = d_out_rgba;
  if ((s = launch_detached(ctx, a, st)) != LUCENT_OK) return s;
  // The host table must outlive the copy.
  return stream_sync(st);
}

lucent_status lucent_scene_set_materials(lucent_ctx* ctx, lucent_scene* scene, const float* materials, uint32_t n) {
  clear_error();
  if (!ctx || !scene || !materials) return fail(LUCENT_E_INVALID, "null argument");
  if (n != scene->pt.material_count) return fail(LUCENT_E_INVALID, "material count differs from the uploaded scene");
  Status s = stream_sync(ctx->stream);
  if (s != LUCENT_OK) return s;
  if ((s = dev_upload(const_cast<float*>(scene->pt.materials), materials, size_t(n) * 40 * sizeof(float), ctx->stream)) !=
      LUCENT_OK)
    return s;
  scene->blend_alpha = any_blend_alpha(materials, n);
  return stream_sync(ctx->stream);
}

lucent_status lucent_scene_update_env_texels(lucent_ctx* ctx, lucent_scene* scene, const float* rgb, uint32_t w,
                                             uint32_t h) {
  clear_error();
  if (!ctx || !scene || !rgb) return fail(LUCENT_E_INVALID, "null argument");
  if (!scene->lighting_set || int32_t(w) != scene->pt.env_w || int32_t(h) != scene->pt.env_h)
    return fail(LUCENT_E_INVALID, "environment size differs from the scene's map");
  std::vector<float> map4(size_t(w) * h * 4);
  for (size_t i = 0; i < size_t(w) * h; ++i) {
    map4[i * 4] = rgb[i * 3];
    map4[i * 4 + 1] = rgb[i * 3 + 1];
    map4[i * 4 + 2] = rgb[i * 3 + 2];
    map4[i * 4 + 3] = 0.0f;
  }
  Status s = stream_sync(ctx->stream);
  if (s != LUCENT_OK) return s;
  if ((s = dev_upload(const_cast<float*>(scene->pt.env_map), map4.data(), map4.size() * sizeof(float), ctx->stream)) !=
      LUCENT_OK)
    return s;
  return stream_sync(ctx->stream);
}

lucent_status lucent_env_sampler_gpu(lucent_ctx* ctx, const float* d_rgba, uint32_t width, uint32_t height,
                                     uint32_t max_width, float* d_table_out, uint32_t* grid_w, uint32_t* grid_h,
                                     double* total, float* gpu_ms, void* stream) {
  clear_error();
  if (!ctx) return fail(LUCENT_E_INVALID, "null ctx");
  return build_env_sampler_gpu(ctx, d_rgba, width, height, max_width, d_table_out, grid_w, grid_h, total, stream,
                               gpu_ms);
}

lucent_status lucent_env_sampler_gpu_parallel(lucent_ctx* ctx, const float* d_rgba, uint32_t width, uint32_t height,
                                              uint32_t max_width, float* d_table_out, uint32_t* grid_w,
                                              uint32_t* grid_h, double* total, float* gpu_ms, void* stream) {
  clear_error();
  if (!ctx) return fail(LUCENT_E_INVALID, "null ctx");
  return build_env_sampler_gpu(ctx, d_rgba, width, height, max_width, d_table_out, grid_w, grid_h, total, stream,
                               gpu_ms, true);
}

lucent_status lucent_scene_rebuild_env_sampler(lucent_ctx* ctx, lucent_scene* scene, uint32_t on_device,
                                               double* total_out) {
  clear_error();
  if (!ctx || !scene) return fail(LUCENT_E_INVALID, "null ctx/scene");
  PtScene& pt = scene->pt;
  if (!scene->lighting_set || !pt.env_map || !pt.env_table) return fail(LUCENT_E_INVALID, "scene has no lighting");
  const uint32_t w = uint32_t(pt.env_w), h = uint32_t(pt.env_h);
  uint32_t gw = 0, gh = 0;
  double total = 0;
  Status s;
  if (on_device > 2) return fail(LUCENT_E_INVALID, "on_device must be 0, 1 or 2");
  if (on_device) {
    // Writes the scene's own table in place (same grid: the map keeps its
    // size); on failure the table is left as it was.
    s = build_env_sampler_gpu(ctx, pt.env_map, w, h, 1024, const_cast<float*>(pt.env_table), &gw, &gh, &total,
                              ctx->stream, nullptr, on_device == 2);
    if (s != LUCENT_OK) return s;
  } else {
    // The map comes back to the host, the host build runs, the table goes up.
    std::vector<float> map4(size_t(w) * h * 4), rgb(size_t(w) * h * 3);
    if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
    if ((s = dev_download(map4.data(), pt.env_map, map4.size() * sizeof(float), ctx->stream)) != LUCENT_OK) return s;
    if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
    for (size_t i = 0; i < size_t(w) * h; ++i)
      for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = map4[i * 4 + c];
    EnvSampler es;
    if ((s = build_env_sampler(rgb.data(), w, h, 1024, &es)) != LUCENT_OK) return s;
    gw = es.grid_w;
    gh = es.grid_h;
    total = es.total;
    if (int32_t(gw) == pt.grid_w && int32_t(gh) == pt.grid_h) {
      if ((s = dev_upload(const_cast<float*>(pt.env_table), es.table.data(), es.table.size() * sizeof(float),
                          ctx->stream)) != LUCENT_OK)
        return s;
      if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
    }
  }
  if (int32_t(gw) != pt.grid_w || int32_t(gh) != pt.grid_h) return fail(LUCENT_E_INVALID, "sampler grid changed");
  pt.env_sampling = total > 0 ? 1 : 0;
  if (total_out) *total_out = total;
  return LUCENT_OK;
}

lucent_status lucent_scene_env_size(const lucent_scene* scene, uint32_t* w, uint32_t* h) {
  if (!scene) return fail(LUCENT_E_INVALID, "null scene");
  if (w) *w = uint32_t(scene->pt.env_w);
  if (h) *h = uint32_t(scene->pt.env_h);
  return LUCENT_OK;
}

static uint32_t atrous_levels(uint32_t spp) {
  return spp <= 4 ? 5 : spp <= 16 ? 4 : spp <= 64 ? 3 : spp <= 256 ? 2 : 1;
}

// DenoiseArgs over the film's targets; iterations 0 = atrousLevels(samples).
static Status denoise_args(lucent_film* film, const lucent_camera* cam, const lucent_denoise_params* p, uint32_t samples,
                           DenoiseArgs* a) {
  if (p->iterations > 12) return fail(LUCENT_E_INVALID, "at most 12 a-trous iterations");
  const size_t px = size_t(film->width) * film->height;
  if (!film->filter) {
    Status s = dev_alloc(&film->filter, px * (16 * 4 + 8));
    if (s != LUCENT_OK) return s;
  }
  a->compute_units = std::max(film->ctx->compute_units, 1);
  a->color = film->f.color;
  a->albedo = film->f.albedo;
  a->normal = film->f.normal;
  a->history = film->f.history;
  a->ping = static_cast<float*>(film->filter);
  a->pong = a->ping + px * 4;
  a->remodulated = a->pong + px * 4;
  a->guides = a->remodulated + px * 4;
  a->var_a = a->guides + px * 4;
  a->var_b = a->var_a + px;
  a->result = &film->denoised;
  a->width = int(film->width);
  a->height = int(film->height);
  a->iterations = p->iterations ? p->iterations : atrous_levels(samples);
  a->sigma_l = p->sigma_l;
  a->sigma_n = p->sigma_n;
  a->sigma_p = p->sigma_p;
  double fwd[3], right[3], up[3], tan_half;
  camera_basis(*cam, fwd, right, up, &tan_half);
  for (int i = 0; i < 3; ++i) {
    a->cam_pos[i] = float(cam->position[i]);
    a->cam_fwd[i] = float(fwd[i]);
    a->cam_right[i] = float(right[i]);
    a->cam_up[i] = float(up[i]);
  }
  a->tan_half_fov = float(tan_half);
  film->remodulated = a->remodulated;
  return LUCENT_OK;
}

lucent_status lucent_denoise(lucent_ctx* ctx, lucent_film* film, const lucent_camera* cam,
                             const lucent_denoise_params* p, void* stream) {
  clear_error();
  if (!ctx || !film || !cam || !p) return fail(LUCENT_E_INVALID, "null argument");
  DenoiseArgs a;
  Status s = denoise_args(film, cam, p, film->samples, &a);
  if (s != LUCENT_OK) return s;
  return launch_denoise(a, stream ? stream : ctx->stream);
}

static Status display_args(const lucent_film* film, const lucent_display_params* p, uint32_t view_w, uint32_t view_h,
                           DisplayArgs* a) {
  if (view_w == 0 || view_h == 0 || uint64_t(view_w) * view_h > (1ull << 28))
    return fail(LUCENT_E_INVALID, "view size must be positive and at most 2^28 pixels");
  if (p->tonemap > 2) return fail(LUCENT_E_INVALID, "tonemap must be 0 (ACES), 1 (Reinhard) or 2 (none)");
  if (!(p->denoise_weight >= 0.0f && p->denoise_weight <= 1.0f))
    return fail(LUCENT_E_INVALID, "denoise_weight must lie in [0, 1]");
  if (p->denoise_weight > 0.0f && !film->denoised)
    return fail(LUCENT_E_INVALID, "denoise_weight > 0 needs a lucent_denoise of the film first");
  a->color = film->f.color;
  a->albedo = film->f.albedo;
  a->denoised = p->denoise_weight > 0.0f ? film->denoised : nullptr;
  a->render_w = int(film->width);
  a->render_h = int(film->height);
  a->view_w = int(view_w);
  a->view_h = int(view_h);
  a->denoise_weight = p->denoise_weight;
  a->exposure = p->exposure;
  a->tonemap = int(p->tonemap);
  return LUCENT_OK;
}

lucent_status lucent_display(lucent_ctx* ctx, const lucent_film* film, const lucent_display_params* p, uint32_t view_w,
                             uint32_t view_h, uint8_t* d_rgba8, void* stream) {
  clear_error();
  if (!ctx || !film || !p || !d_rgba8) return fail(LUCENT_E_INVALID, "null argument");
  DisplayArgs a;
  Status s = display_args(film, p, view_w, view_h, &a);
  if (s != LUCENT_OK) return s;
  a.out = d_rgba8;
  return launch_display(a, stream ? stream : ctx->stream);
}

lucent_status lucent_screenshot(lucent_ctx* ctx, const lucent_film* film, const lucent_display_params* p,
                                uint32_t view_w, uint32_t view_h, uint8_t* rgba8) {
  clear_error();
  if (!ctx || !film || !p || !rgba8) return fail(LUCENT_E_INVALID, "null argument");
  DisplayArgs a;
  Status s = display_args(film, p, view_w, view_h, &a);
  if (s != LUCENT_OK) return s;
  const size_t bytes = size_t(view_w) * view_h * 4;
  if ((s = ensure_scratch(ctx, bytes)) != LUCENT_OK) return s;
  a.out = static_cast<uint8_t*>(ctx->scratch);
  if ((s = launch_display(a, ctx->stream)) != LUCENT_OK) return s;
  if ((s = dev_download(rgba8, ctx->scratch, bytes, ctx->stream)) != LUCENT_OK) return s;
  return stream_sync(ctx->stream);
}

lucent_status lucent_film_read_denoised(lucent_ctx* ctx, const lucent_film* film, float* illum_var4,
                                        float* radiance4) {
  clear_error();
  if (!ctx || !film) return fail(LUCENT_E_INVALID, "null ctx/film");
  if (!film->denoised) return fail(LUCENT_E_INVALID, "film has not been denoised");
  const size_t px = size_t(film->width) * film->height;
  hipDeviceSynchronize();
  Status s;
  if (illum_var4 && (s = dev_download(illum_var4, film->denoised, px * 16, ctx->stream)) != LUCENT_OK) return s;
  if (radiance4 && (s = dev_download(radiance4, film->remodulated, px * 16, ctx->stream)) != LUCENT_OK) return s;
  return stream_sync(ctx->stream);
}


struct lucent_frame_graph {
  lucent_ctx* ctx = nullptr;
  lucent_film* film = nullptr;
  uint32_t samples = 0;
  uint32_t expected_seed = 0;
  void* buf = nullptr;  // [seed word, job counter] in the first 256 bytes, then the spill area
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
};

void lucent_frame_graph_destroy(lucent_frame_graph* g) {
  if (!g) return;
  if (g->ctx) hipStreamSynchronize(static_cast<hipStream_t>(g->ctx->stream));
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->capture) hipStreamDestroy(g->capture);
  if (g->buf) {
    hipDeviceSynchronize();
    hipFree(g->buf);
  }
  delete g;
}

lucent_status lucent_frame_graph_create(lucent_ctx* ctx, lucent_scene* scene, lucent_film* film,
                                        const lucent_camera* cam, const lucent_render_settings* settings,
                                        uint32_t samples, const lucent_denoise_params* denoise,
                                        const lucent_display_params* display, uint32_t view_w, uint32_t view_h,
                                        uint8_t* d_rgba8, lucent_frame_graph** out) {
  clear_error();
  if (!ctx || !scene || !film || !cam || !settings || !out) return fail(LUCENT_E_INVALID, "null argument");
  *out = nullptr;
  if (samples == 0) return fail(LUCENT_E_INVALID, "a frame graph traces at least one sample");
  if (settings->flags & (LUCENT_TRACE_PER_SAMPLE | LUCENT_TRACE_WAVEFRONT))
    return fail(LUCENT_E_UNSUPPORTED, "frame graphs capture the persistent tile launch only");
  if (display && !d_rgba8) return fail(LUCENT_E_INVALID, "display needs an output buffer");
  Status s = check_traceable(scene);
  if (s != LUCENT_OK) return s;
  // Finish queued work on the film before its counters are read.
  if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
  TileItem it;
  it.scene = scene->pt;
  fill_frame(cam, settings, film->width, film->height, &it.frame);
  it.frame.seed = film->seed;
  it.film = film->f;
  it.spp = samples;
  TileLaunch cfg;
  cfg.fused = !(settings->flags & LUCENT_TRACE_INLINE_SHADOWS) && !scene->blend_alpha;
  cfg.waves_per_cu = settings->waves_per_cu;
  cfg.occupancy = settings->occupancy;
  cfg.tile_w = settings->tile_w;
  cfg.tile_h = settings->tile_h;
  cfg.shade_batch = settings->shade_batch;
  cfg.trace_turns = settings->trace_turns;
  cfg.job_paths = settings->job_paths;
  DenoiseArgs da;
  if (denoise && (s = denoise_args(film, cam, denoise, film->samples + samples, &da)) != LUCENT_OK) return s;
  DisplayArgs disp;
  if (display) {
    if (display->denoise_weight > 0.0f && !denoise && !film->denoised)
      return fail(LUCENT_E_INVALID, "denoise_weight > 0 needs a denoise step in the graph or a lucent_denoise first");
    lucent_film probe = *film;
    if (denoise) probe.denoised = static_cast<float*>(film->filter);  // replaced after the denoise is captured
    if ((s = display_args(&probe, display, view_w, view_h, &disp)) != LUCENT_OK) return s;
    disp.out = d_rgba8;
  }
  auto* g = new lucent_frame_graph();
  g->ctx = ctx;
  g->film = film;
  g->samples = samples;
  g->expected_seed = film->seed;
  const size_t spill = tile_spill_bytes(ctx, it, cfg);
  auto bail = [&](Status st) {
    lucent_frame_graph_destroy(g);
    return st;
  };
  // [seed word | job counter | film descriptor slot | spill], 256-byte aligned parts.
  const size_t item_bytes = (sizeof(TileItem) + 255) / 256 * 256;
  if ((s = dev_alloc(&g->buf, 256 + item_bytes + spill)) != LUCENT_OK) return bail(s);
  uint32_t* seed_word = static_cast<uint32_t*>(g->buf);
  cfg.counter = seed_word + 1;
  cfg.item = reinterpret_cast<TileItem*>(static_cast<char*>(g->buf) + 256);
  cfg.spill = reinterpret_cast<uint32_t*>(static_cast<char*>(g->buf) + 256 + item_bytes);
  cfg.spill_bytes = spill;
  it.seed_src = seed_word;
  hipError_t e = hipMemcpy(seed_word, &film->seed, sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return bail(fail_hip(e, "frame graph capture start"));
  // Trace, advance the device seed, denoise, display.
  Status cs = launch_trace_tiles(ctx, &it, 1, cfg, g->capture);
  if (cs == LUCENT_OK) cs = launch_add_u32(seed_word, samples, g->capture);
  if (cs == LUCENT_OK && denoise) cs = launch_denoise(da, g->capture);
  if (denoise && disp.denoised) disp.denoised = film->denoised;  // the target the captured iterations end in
  if (cs == LUCENT_OK && display) cs = launch_display(disp, g->capture);
  e = hipStreamEndCapture(g->capture, &g->graph);
  if (cs != LUCENT_OK) return bail(cs);
  if (e != hipSuccess) return bail(fail_hip(e, "frame graph capture"));
  if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
    return bail(fail_hip(e, "hipGraphInstantiate"));
  *out = g;
  return LUCENT_OK;
}

struct lucent_batch_graph {
  lucent_ctx* ctx = nullptr;
  std::vector<lucent_film*> films;      // in launch order
  std::vector<uint32_t> samples;        // per film
  std::vector<uint32_t> expected_seed;  // per film, the host seed a replay expects
  void* buf = nullptr;  // [job table | seed words | arrival counters | path records | spill]
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture = nullptr;
};

void lucent_batch_graph_destroy(lucent_batch_graph* g) {
  if (!g) return;
  if (g->ctx) hipStreamSynchronize(static_cast<hipStream_t>(g->ctx->stream));
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->capture) hipStreamDestroy(g->capture);
  if (g->buf) {
    hipDeviceSynchronize();
    hipFree(g->buf);
  }
  delete g;
}

lucent_status lucent_batch_graph_create(lucent_ctx* ctx, const lucent_batch_item* items, uint32_t count,
                                        const lucent_render_settings* settings, lucent_batch_graph** out) {
  clear_error();
  if (!ctx || !items || count == 0 || !settings || !out) return fail(LUCENT_E_INVALID, "null argument");
  *out = nullptr;
  if (settings->flags & (LUCENT_TRACE_PER_SAMPLE | LUCENT_TRACE_WAVEFRONT))
    return fail(LUCENT_E_UNSUPPORTED, "batch graphs capture the persistent batch launch only");
  TileLaunch cfg;
  std::vector<TileItem> tiles;
  std::vector<lucent_film*> films;
  Status s = batch_tiles(items, count, settings, &cfg, &tiles, &films);
  if (s != LUCENT_OK) return s;
  if (tiles.empty()) return fail(LUCENT_E_INVALID, "a batch graph traces at least one sample");
  // Finish queued work on the films before their seeds are read.
  if ((s = stream_sync(ctx->stream)) != LUCENT_OK) return s;
  for (size_t k = 0; k < tiles.size(); ++k) tiles[k].frame.seed = films[k]->seed;
  BatchPlan plan;
  const uint32_t n = uint32_t(tiles.size());
  if ((s = plan_batch(ctx, tiles.data(), n, cfg, &plan)) != LUCENT_OK) return s;
  auto align = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t begin_off = 0, items_off = 0;
  const size_t table = align(batch_table_bytes(n, &begin_off, &items_off));
  const size_t seeds_off = table, arrived_off = seeds_off + align(size_t(n) * sizeof(uint32_t));
  const size_t records_off = arrived_off + align(plan.arrived_bytes);
  const size_t spill_off = records_off + align(plan.records_bytes);
  const size_t bytes = spill_off + align(plan.spill_bytes);
  auto* g = new lucent_batch_graph();
  g->ctx = ctx;
  auto bail = [&](Status st) {
    lucent_batch_graph_destroy(g);
    return st;
  };
  if ((s = dev_alloc(&g->buf, bytes)) != LUCENT_OK) return bail(s);
  char* base = static_cast<char*>(g->buf);
  uint32_t* seeds = reinterpret_cast<uint32_t*>(base + seeds_off);
  std::vector<uint32_t> seed0(n);
  for (uint32_t k = 0; k < n; ++k) {
    plan.items[k].seed_src = seeds + k;
    seed0[k] = films[k]->seed;
  }
  std::vector<char> host(table - begin_off, 0);
  std::memcpy(host.data(), plan.job_begin.data(), plan.job_begin.size() * sizeof(uint32_t));
  std::memcpy(host.data() + (items_off - begin_off), plan.items.data(), size_t(n) * sizeof(TileItem));
  hipError_t e = hipMemcpy(base + begin_off, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(seeds, seed0.data(), size_t(n) * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return bail(fail_hip(e, "batch graph capture start"));
  // Trace every film, then advance the device seeds.
  const TileItem* d_items = reinterpret_cast<const TileItem*>(base + items_off);
  Status cs = launch_batch_planned(plan, reinterpret_cast<uint32_t*>(base),
                                   reinterpret_cast<const uint32_t*>(base + begin_off), d_items,
                                   reinterpret_cast<uint32_t*>(base + spill_off), base + records_off,
                                   reinterpret_cast<uint32_t*>(base + arrived_off), g->capture);
  if (cs == LUCENT_OK) cs = launch_advance_seeds(seeds, d_items, n, g->capture);
  e = hipStreamEndCapture(g->capture, &g->graph);
  if (cs != LUCENT_OK) return bail(cs);
  if (e != hipSuccess) return bail(fail_hip(e, "batch graph capture"));
  if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
    return bail(fail_hip(e, "hipGraphInstantiate"));
  g->films = films;
  g->samples.resize(n);
  g->expected_seed.resize(n);
  for (uint32_t k = 0; k < n; ++k) {
    g->samples[k] = tiles[k].spp;
    g->expected_seed[k] = films[k]->seed;
  }
  *out = g;
  return LUCENT_OK;
}

lucent_status lucent_batch_graph_launch(lucent_batch_graph* g, void* stream) {
  clear_error();
  if (!g) return fail(LUCENT_E_INVALID, "null graph");
  for (size_t k = 0; k < g->films.size(); ++k)
    if (g->films[k]->seed != g->expected_seed[k])
      return fail(LUCENT_E_INVALID, "a film was advanced outside this graph; create a new graph");
  const hipError_t e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream ? stream : g->ctx->stream));
  if (e != hipSuccess) return fail_hip(e, "hipGraphLaunch");
  for (size_t k = 0; k < g->films.size(); ++k) {
    g->films[k]->seed += g->samples[k];
    g->films[k]->samples += g->samples[k];
    g->expected_seed[k] = g->films[k]->seed;
  }
  return LUCENT_OK;
}

lucent_status lucent_frame_graph_launch(lucent_frame_graph* g, void* stream) {
  clear_error();
  if (!g) return fail(LUCENT_E_INVALID, "null graph");
  if (g->film->seed != g->expected_seed)
    return fail(LUCENT_E_INVALID, "the film was advanced outside this graph; create a new graph");
  const hipError_t e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream ? stream : g->ctx->stream));
  if (e != hipSuccess) return fail_hip(e, "hipGraphLaunch");
  g->film->seed += g->samples;
  g->film->samples += g->samples;
  g->expected_seed = g->film->seed;
  return LUCENT_OK;
}

}  // extern "C"
