/*
 * N-API addon over the C ABI of include/lucent_hip.h: the binding a Node
 * host (Lucent's TypeScript, run as JavaScript) uses to drive the gfx950
 * kernels.  TypedArray arguments are passed to the library zero-copy
 * (napi_get_typedarray_info); results come back as new TypedArrays.  A
 * failed call throws a JavaScript Error carrying lucent_last_error(), as the
 * reference throws Error/GLError.  Handles are externals whose finalisers
 * release the device objects; a scene or film keeps its context alive.
 */
#include <node_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lucent_hip.h"

#define CHECK_NAPI(call)                                  \
  do {                                                    \
    if ((call) != napi_ok) {                              \
      napi_throw_error(env, NULL, "N-API call failed: " #call); \
      return NULL;                                        \
    }                                                     \
  } while (0)

static napi_value throw_status(napi_env env, lucent_status st) {
  char code[32];
  snprintf(code, sizeof code, "LUCENT_%d", (int)st);
  napi_throw_error(env, code, lucent_last_error(NULL));
  return NULL;
}
#define CHECK_LUCENT(call)                      \
  do {                                          \
    lucent_status st_ = (call);                 \
    if (st_ != LUCENT_OK) return throw_status(env, st_); \
  } while (0)

/* ---- handles ------------------------------------------------------------ */
typedef struct {
  lucent_ctx *ctx;
} ctx_holder;
typedef struct {
  lucent_scene *scene;
  napi_ref ctx_ref;
} scene_holder;
typedef struct {
  lucent_film *film;
  napi_ref ctx_ref;
} film_holder;

static void ctx_finalize(napi_env env, void *data, void *hint) {
  ctx_holder *h = (ctx_holder *)data;
  if (h->ctx) lucent_destroy(h->ctx);
  free(h);
}
static void scene_finalize(napi_env env, void *data, void *hint) {
  scene_holder *h = (scene_holder *)data;
  if (h->scene) lucent_scene_destroy(h->scene);
  napi_delete_reference(env, h->ctx_ref);
  free(h);
}
static void film_finalize(napi_env env, void *data, void *hint) {
  film_holder *h = (film_holder *)data;
  if (h->film) lucent_film_destroy(h->film);
  napi_delete_reference(env, h->ctx_ref);
  free(h);
}

static void *get_external(napi_env env, napi_value v) {
  void *p = NULL;
  if (napi_get_value_external(env, v, &p) != napi_ok) return NULL;
  return p;
}
static lucent_ctx *get_ctx(napi_env env, napi_value v) {
  ctx_holder *h = (ctx_holder *)get_external(env, v);
  return h ? h->ctx : NULL;
}

/* ---- argument helpers --------------------------------------------------- */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *len) {
  bool is = false;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return 0;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, len, data, &ab, &off) != napi_ok) return 0;
  return t == want;
}
static napi_value prop(napi_env env, napi_value obj, const char *name) {
  napi_value v = NULL;
  napi_get_named_property(env, obj, name, &v);
  return v;
}
static double num(napi_env env, napi_value v, double dflt) {
  napi_valuetype t;
  if (!v || napi_typeof(env, v, &t) != napi_ok) return dflt;
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1 : 0;
  }
  if (t != napi_number) return dflt;
  double d;
  napi_get_value_double(env, v, &d);
  return d;
}
static void vec3(napi_env env, napi_value arr, double *out, double dflt) {
  for (int i = 0; i < 3; i++) {
    napi_value e = NULL;
    out[i] = dflt;
    if (arr && napi_get_element(env, arr, i, &e) == napi_ok) out[i] = num(env, e, dflt);
  }
}
static napi_value new_typed(napi_env env, napi_typedarray_type t, size_t n, size_t esz, void **data) {
  napi_value ab, ta;
  if (napi_create_arraybuffer(env, n * esz, data, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, t, n, ab, 0, &ta) != napi_ok) return NULL;
  return ta;
}
static napi_value make_number(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

/* ---- exports ------------------------------------------------------------ */
static napi_value js_version(napi_env env, napi_callback_info info) {
  napi_value v;
  CHECK_NAPI(napi_create_string_utf8(env, lucent_version(), NAPI_AUTO_LENGTH, &v));
  return v;
}

/* buildBVH(positions: Float32Array, {maxLeafSize, bins}) — src/math/bvh.ts:52. */
/* {nodes: Float32Array, nodeCount, triIndices, bounds: {min, max}} of buildBVH (bvh.ts:16-24). */
static napi_value bvh_result(napi_env env, const float *nodes, uint32_t count, napi_value idx, const double bounds[6]) {
  void *nd;
  napi_value nodes_ta = new_typed(env, napi_float32_array, (size_t)count * 8, 4, &nd);
  if (!nodes_ta) return NULL;
  memcpy(nd, nodes, (size_t)count * 8 * sizeof(float));
  napi_value out, bmin, bmax, b;
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_set_named_property(env, out, "nodes", nodes_ta));
  CHECK_NAPI(napi_set_named_property(env, out, "nodeCount", make_number(env, count)));
  CHECK_NAPI(napi_set_named_property(env, out, "triIndices", idx));
  CHECK_NAPI(napi_create_array_with_length(env, 3, &bmin));
  CHECK_NAPI(napi_create_array_with_length(env, 3, &bmax));
  for (int i = 0; i < 3; i++) {
    napi_set_element(env, bmin, i, make_number(env, bounds[i]));
    napi_set_element(env, bmax, i, make_number(env, bounds[3 + i]));
  }
  CHECK_NAPI(napi_create_object(env, &b));
  CHECK_NAPI(napi_set_named_property(env, b, "min", bmin));
  CHECK_NAPI(napi_set_named_property(env, b, "max", bmax));
  CHECK_NAPI(napi_set_named_property(env, out, "bounds", b));
  return out;
}

static napi_value js_build_bvh(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *pos;
  size_t len;
  if (argc < 1 || !typed(env, argv[0], napi_float32_array, &pos, &len)) {
    napi_throw_type_error(env, NULL, "buildBVH: positions must be a Float32Array");
    return NULL;
  }
  if (len % 9 != 0) {
    char msg[96];
    snprintf(msg, sizeof msg, "positions length %zu is not a multiple of 9", len);
    napi_throw_error(env, NULL, msg);
    return NULL;
  }
  uint32_t max_leaf = 4, bins = 12;
  if (argc > 1) {
    max_leaf = (uint32_t)num(env, prop(env, argv[1], "maxLeafSize"), 4);
    bins = (uint32_t)num(env, prop(env, argv[1], "bins"), 12);
  }
  uint32_t tris = (uint32_t)(len / 9);
  size_t cap = (tris > 0 ? 2 * (size_t)tris - 1 : 1) * 8;
  float *nodes = (float *)malloc(cap * sizeof(float));
  void *idx_data;
  napi_value idx = new_typed(env, napi_uint32_array, tris, 4, &idx_data);
  uint32_t count = 0;
  double bounds[6];
  lucent_status st = lucent_build_bvh((const float *)pos, tris, max_leaf, bins, nodes, &count, (uint32_t *)idx_data,
                                      bounds);
  if (st != LUCENT_OK) {
    free(nodes);
    return throw_status(env, st);
  }
  napi_value out = bvh_result(env, nodes, count, idx, bounds);
  free(nodes);
  return out;
}

static napi_value js_create_context(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  int device = argc > 0 ? (int)num(env, argv[0], 0) : 0;
  lucent_ctx *ctx = NULL;
  CHECK_LUCENT(lucent_create(device, &ctx));
  ctx_holder *h = (ctx_holder *)calloc(1, sizeof *h);
  h->ctx = ctx;
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, ctx_finalize, NULL, &ext));
  return ext;
}

static napi_value wrap_scene(napi_env env, napi_value ctx_v, lucent_scene *s) {
  scene_holder *h = (scene_holder *)calloc(1, sizeof *h);
  h->scene = s;
  napi_create_reference(env, ctx_v, 1, &h->ctx_ref);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, scene_finalize, NULL, &ext));
  return ext;
}

/* sceneFromBVH(ctx, positions, bvh) — the arguments of intersectBVH. */
static napi_value js_scene_from_bvh(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  void *pos, *nodes, *idx;
  size_t plen, nlen, ilen;
  if (!ctx || argc < 3 || !typed(env, argv[1], napi_float32_array, &pos, &plen) ||
      !typed(env, prop(env, argv[2], "nodes"), napi_float32_array, &nodes, &nlen) ||
      !typed(env, prop(env, argv[2], "triIndices"), napi_uint32_array, &idx, &ilen)) {
    napi_throw_type_error(env, NULL, "sceneFromBVH(ctx, positions: Float32Array, bvh: BVH)");
    return NULL;
  }
  lucent_scene *s = NULL;
  CHECK_LUCENT(lucent_scene_from_bvh(ctx, (const float *)pos, (uint32_t)(plen / 9), (const float *)nodes,
                                     (uint32_t)(nlen / 8), (const uint32_t *)idx, &s));
  return wrap_scene(env, argv[0], s);
}

/* uploadScene(ctx, packed: PackedScene) — Renderer.loadScene (renderer.ts:279-312). */
static napi_value js_upload_scene(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  if (!ctx || argc < 2) {
    napi_throw_type_error(env, NULL, "uploadScene(ctx, packed)");
    return NULL;
  }
  napi_value p = argv[1];
  lucent_packed_scene ps;
  memset(&ps, 0, sizeof ps);
  void *d;
  size_t n;
  if (!typed(env, prop(env, p, "triGeometry"), napi_float32_array, &d, &n)) goto bad;
  ps.tri_geometry = (const float *)d;
  ps.tri_count = (uint32_t)num(env, prop(env, p, "triCount"), (double)(n / 12));
  if (typed(env, prop(env, p, "triAttributes"), napi_float32_array, &d, &n)) ps.tri_attributes = (const float *)d;
  if (!typed(env, prop(env, p, "bvhNodes"), napi_float32_array, &d, &n)) goto bad;
  ps.bvh_nodes = (const float *)d;
  ps.bvh_node_count = (uint32_t)num(env, prop(env, p, "bvhNodeCount"), (double)(n / 8));
  {
    napi_value cpu = prop(env, p, "cpu"), bvh = cpu ? prop(env, cpu, "bvh") : NULL;
    if (bvh && typed(env, prop(env, bvh, "triIndices"), napi_uint32_array, &d, &n)) ps.tri_indices = (const uint32_t *)d;
  }
  if (!typed(env, prop(env, p, "materials"), napi_float32_array, &d, &n)) goto bad;
  ps.materials = (const float *)d;
  ps.material_count = (uint32_t)num(env, prop(env, p, "materialCount"), (double)(n / 40));
  if (typed(env, prop(env, p, "spheres"), napi_float32_array, &d, &n)) ps.spheres = (const float *)d;
  if (typed(env, prop(env, p, "sphereMaterials"), napi_int32_array, &d, &n)) ps.sphere_materials = (const int32_t *)d;
  if (typed(env, prop(env, p, "sphereLights"), napi_int32_array, &d, &n)) ps.sphere_lights = (const int32_t *)d;
  ps.sphere_count = (uint32_t)num(env, prop(env, p, "sphereCount"), 0);
  if (typed(env, prop(env, p, "lights"), napi_float32_array, &d, &n)) ps.lights = (const float *)d;
  ps.light_count = (uint32_t)num(env, prop(env, p, "lightCount"), 0);
  ps.ground_material = (int32_t)num(env, prop(env, p, "groundMaterial"), -1);
  ps.ground_height = (float)num(env, prop(env, p, "groundHeight"), 0);
  {
    lucent_scene *s = NULL;
    CHECK_LUCENT(lucent_scene_upload(ctx, &ps, &s));
    return wrap_scene(env, argv[0], s);
  }
bad:
  napi_throw_type_error(env, NULL, "uploadScene: packed scene arrays missing or of the wrong type");
  return NULL;
}

/* setLighting(ctx, scene, {sun, sky, env}) — Renderer.uploadLighting (renderer.ts:355-413). */
static napi_value js_set_lighting(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  scene_holder *sh = argc > 1 ? (scene_holder *)get_external(env, argv[1]) : NULL;
  if (!ctx || !sh || argc < 3) {
    napi_throw_type_error(env, NULL, "setLighting(ctx, scene, {sun, sky, env})");
    return NULL;
  }
  lucent_lighting L;
  memset(&L, 0, sizeof L);
  napi_value sun = prop(env, argv[2], "sun"), sky = prop(env, argv[2], "sky"), e = prop(env, argv[2], "env");
  if (sun) {
    L.sun.elevation = num(env, prop(env, sun, "elevation"), 45);
    L.sun.azimuth = num(env, prop(env, sun, "azimuth"), 0);
    vec3(env, prop(env, sun, "color"), L.sun.color, 1);
    L.sun.intensity = num(env, prop(env, sun, "intensity"), 0);
    L.sun.angular_radius = num(env, prop(env, sun, "angularRadius"), 0.27);
  }
  if (sky) {
    L.sky.enabled = (int32_t)num(env, prop(env, sky, "enabled"), 0);
    vec3(env, prop(env, sky, "zenith"), L.sky.zenith, 0);
    vec3(env, prop(env, sky, "horizon"), L.sky.horizon, 0);
    vec3(env, prop(env, sky, "ground"), L.sky.ground, 0);
    L.sky.intensity = num(env, prop(env, sky, "intensity"), 0);
  }
  if (e) {
    napi_value map = prop(env, e, "map");
    napi_valuetype t = napi_null;
    if (map) napi_typeof(env, map, &t);
    if (t == napi_object) {
      void *d;
      size_t n;
      if (!typed(env, prop(env, map, "data"), napi_float32_array, &d, &n)) {
        napi_throw_type_error(env, NULL, "env.map.data must be a Float32Array");
        return NULL;
      }
      L.env.map_rgb = (const float *)d;
      L.env.width = (uint32_t)num(env, prop(env, map, "width"), 0);
      L.env.height = (uint32_t)num(env, prop(env, map, "height"), 0);
      if ((size_t)L.env.width * L.env.height * 3 != n) {
        napi_throw_error(env, NULL, "env.map: width * height * 3 != data.length");
        return NULL;
      }
    }
    L.env.rotation = num(env, prop(env, e, "rotation"), 0);
    L.env.intensity = num(env, prop(env, e, "intensity"), 1);
    L.env.background = (int32_t)num(env, prop(env, e, "background"), 1);
    vec3(env, prop(env, e, "backdrop"), L.env.backdrop, 0);
  }
  CHECK_LUCENT(lucent_scene_set_lighting(ctx, sh->scene, &L));
  return NULL;
}

/* intersect(ctx, scene, origins, dirs, tMin, tMax, flags) — intersectBVH over a batch. */
static napi_value js_intersect(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value argv[7];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  scene_holder *sh = argc > 1 ? (scene_holder *)get_external(env, argv[1]) : NULL;
  void *o, *d;
  size_t on, dn;
  int f64 = 0;
  if (!ctx || !sh || argc < 4) goto bad;
  if (typed(env, argv[2], napi_float64_array, &o, &on) && typed(env, argv[3], napi_float64_array, &d, &dn))
    f64 = 1;
  else if (!(typed(env, argv[2], napi_float32_array, &o, &on) && typed(env, argv[3], napi_float32_array, &d, &dn)))
    goto bad;
  if (on != dn || on % 3) goto bad;
  {
    double tmin = argc > 4 ? num(env, argv[4], 1e-5) : 1e-5;
    double tmax = argc > 5 ? num(env, argv[5], 1.0 / 0.0) : 1.0 / 0.0;
    uint32_t flags = argc > 6 ? (uint32_t)num(env, argv[6], 0) : 0;
    size_t n = on / 3;
    void *t, *u, *v, *id;
    napi_value ta = new_typed(env, napi_float64_array, n, 8, &t), ua = new_typed(env, napi_float64_array, n, 8, &u),
               va = new_typed(env, napi_float64_array, n, 8, &v), ia = new_typed(env, napi_uint32_array, n, 4, &id);
    if (f64)
      CHECK_LUCENT(lucent_intersect_f64(ctx, sh->scene, (const double *)o, (const double *)d, (uint32_t)n, tmin, tmax,
                                        flags, (double *)t, (uint32_t *)id, (double *)u, (double *)v));
    else
      CHECK_LUCENT(lucent_intersect(ctx, sh->scene, (const float *)o, (const float *)d, (uint32_t)n, tmin, tmax, flags,
                                    (double *)t, (uint32_t *)id, (double *)u, (double *)v));
    napi_value out;
    CHECK_NAPI(napi_create_object(env, &out));
    napi_set_named_property(env, out, "t", ta);
    napi_set_named_property(env, out, "triIndex", ia);
    napi_set_named_property(env, out, "u", ua);
    napi_set_named_property(env, out, "v", va);
    return out;
  }
bad:
  napi_throw_type_error(env, NULL,
                        "intersect(ctx, scene, origins, dirs, tMin, tMax, flags): origins/dirs are Float32Array or "
                        "Float64Array of equal length 3n");
  return NULL;
}

static napi_value js_create_film(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  if (!ctx || argc < 3) {
    napi_throw_type_error(env, NULL, "createFilm(ctx, width, height)");
    return NULL;
  }
  lucent_film *f = NULL;
  CHECK_LUCENT(lucent_film_create(ctx, (uint32_t)num(env, argv[1], 0), (uint32_t)num(env, argv[2], 0), &f));
  film_holder *h = (film_holder *)calloc(1, sizeof *h);
  h->film = f;
  napi_create_reference(env, argv[0], 1, &h->ctx_ref);
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, film_finalize, NULL, &ext));
  return ext;
}

static napi_value js_reset_film(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  film_holder *fh = argc > 1 ? (film_holder *)get_external(env, argv[1]) : NULL;
  if (!ctx || !fh) {
    napi_throw_type_error(env, NULL, "resetFilm(ctx, film[, seed])");
    return NULL;
  }
  CHECK_LUCENT(lucent_film_reset(ctx, fh->film));
  if (argc > 2) CHECK_LUCENT(lucent_film_set_seed(fh->film, (uint32_t)num(env, argv[2], 0)));
  return NULL;
}

static napi_value js_film_info(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  film_holder *fh = argc > 0 ? (film_holder *)get_external(env, argv[0]) : NULL;
  if (!fh) {
    napi_throw_type_error(env, NULL, "filmInfo(film)");
    return NULL;
  }
  uint32_t w, h, n, s;
  CHECK_LUCENT(lucent_film_info(fh->film, &w, &h, &n, &s));
  napi_value out;
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "width", make_number(env, w));
  napi_set_named_property(env, out, "height", make_number(env, h));
  napi_set_named_property(env, out, "samples", make_number(env, n));
  napi_set_named_property(env, out, "seed", make_number(env, s));
  return out;
}

static int read_camera(napi_env env, napi_value c, lucent_camera *cam) {
  if (!c) return 0;
  vec3(env, prop(env, c, "position"), cam->position, 0);
  vec3(env, prop(env, c, "target"), cam->target, 0);
  cam->fov = num(env, prop(env, c, "fov"), 45);
  cam->aperture = num(env, prop(env, c, "aperture"), 0);
  cam->focus_distance = num(env, prop(env, c, "focusDistance"), 1);
  return 1;
}

/* traceSamples(ctx, scene, film, camera, {maxBounces, fireflyClamp}, count) — traceSamples (renderer.ts:702). */
/* decodeHDR(bytes: Uint8Array | Buffer) -> {width, height, data: Float32Array} (parseHDR, hdr.ts:7). */
static napi_value js_decode_hdr(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  bool is_ta = false;
  if (argc < 1 || napi_is_typedarray(env, argv[0], &is_ta) != napi_ok || !is_ta) {
    napi_throw_type_error(env, NULL, "decodeHDR(bytes: Uint8Array)");
    return NULL;
  }
  napi_typedarray_type t;
  size_t len;
  void *data;
  napi_value ab;
  size_t off;
  CHECK_NAPI(napi_get_typedarray_info(env, argv[0], &t, &len, &data, &ab, &off));
  if (t != napi_uint8_array) {
    napi_throw_type_error(env, NULL, "decodeHDR(bytes: Uint8Array)");
    return NULL;
  }
  uint32_t w = 0, h = 0;
  CHECK_LUCENT(lucent_decode_hdr((const uint8_t *)data, len, &w, &h, NULL, 0));
  void *rgb;
  napi_value arr = new_typed(env, napi_float32_array, (size_t)w * h * 3, 4, &rgb);
  CHECK_LUCENT(lucent_decode_hdr((const uint8_t *)data, len, &w, &h, (float *)rgb, (size_t)w * h * 3));
  napi_value out;
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "width", make_number(env, w));
  napi_set_named_property(env, out, "height", make_number(env, h));
  napi_set_named_property(env, out, "data", arr);
  return out;
}

/* glTF ingestion (loadGLTF, gltf.ts:188-502) in the C ABI's three phases:
 * gltfParse(bytes) -> {handle, uris}; gltfSetResource(handle, i, bytes);
 * gltfBuild(handle) -> {meshes: [{positions, normals, uvs, material}],
 * materials: [Float64Array(40)], images: [{data, mimeType, name}], bounds,
 * triangleCount, warnings}.  The handle's finaliser frees the document. */
static void gltf_finalize(napi_env env, void *data, void *hint) {
  lucent_gltf_destroy((lucent_gltf *)data);
}
static int bytes_arg(napi_env env, napi_value v, const uint8_t **data, size_t *len) {
  void *p = NULL;
  if (!typed(env, v, napi_uint8_array, &p, len)) return 0;
  *data = (const uint8_t *)p;
  return 1;
}
static napi_value js_gltf_parse(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  const uint8_t *data;
  size_t len;
  if (argc < 1 || !bytes_arg(env, argv[0], &data, &len)) {
    napi_throw_type_error(env, NULL, "gltfParse(bytes: Uint8Array)");
    return NULL;
  }
  lucent_gltf *doc = NULL;
  CHECK_LUCENT(lucent_gltf_parse(data, len, &doc));
  napi_value ext, uris, out;
  if (napi_create_external(env, doc, gltf_finalize, NULL, &ext) != napi_ok) {
    lucent_gltf_destroy(doc);
    napi_throw_error(env, NULL, "N-API call failed: napi_create_external");
    return NULL;
  }
  const uint32_t n = lucent_gltf_resource_count(doc);
  CHECK_NAPI(napi_create_array_with_length(env, n, &uris));
  for (uint32_t i = 0; i < n; i++) {
    napi_value u;
    CHECK_NAPI(napi_create_string_utf8(env, lucent_gltf_resource_uri(doc, i), NAPI_AUTO_LENGTH, &u));
    CHECK_NAPI(napi_set_element(env, uris, i, u));
  }
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_set_named_property(env, out, "handle", ext));
  CHECK_NAPI(napi_set_named_property(env, out, "uris", uris));
  return out;
}
static napi_value js_gltf_set_resource(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_gltf *doc = argc >= 1 ? (lucent_gltf *)get_external(env, argv[0]) : NULL;
  const uint8_t *data;
  size_t len;
  if (argc < 3 || !doc || !bytes_arg(env, argv[2], &data, &len)) {
    napi_throw_type_error(env, NULL, "gltfSetResource(handle, index, bytes: Uint8Array)");
    return NULL;
  }
  CHECK_LUCENT(lucent_gltf_set_resource(doc, (uint32_t)num(env, argv[1], -1), data, len));
  return NULL;
}
static napi_value copy_floats(napi_env env, const float *src, size_t n) {
  void *dst;
  napi_value ta = new_typed(env, napi_float32_array, n, 4, &dst);
  if (ta && n) memcpy(dst, src, n * sizeof(float));
  return ta;
}
static napi_value js_gltf_build(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_gltf *doc = argc >= 1 ? (lucent_gltf *)get_external(env, argv[0]) : NULL;
  if (!doc) {
    napi_throw_type_error(env, NULL, "gltfBuild(handle)");
    return NULL;
  }
  CHECK_LUCENT(lucent_gltf_build(doc));
  lucent_gltf_summary sum;
  CHECK_LUCENT(lucent_gltf_info(doc, &sum));
  napi_value out, meshes, materials, images, warnings, bmin, bmax, bounds;
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_create_array_with_length(env, sum.mesh_count, &meshes));
  for (uint32_t i = 0; i < sum.mesh_count; i++) {
    uint32_t tris;
    int32_t mat;
    const float *p, *n, *uv;
    CHECK_LUCENT(lucent_gltf_mesh(doc, i, &tris, &mat, &p, &n, &uv));
    napi_value m, nulls;
    CHECK_NAPI(napi_create_object(env, &m));
    CHECK_NAPI(napi_set_named_property(env, m, "positions", copy_floats(env, p, (size_t)tris * 9)));
    CHECK_NAPI(napi_set_named_property(env, m, "normals", copy_floats(env, n, (size_t)tris * 9)));
    CHECK_NAPI(napi_get_null(env, &nulls));
    CHECK_NAPI(napi_set_named_property(env, m, "uvs", uv ? copy_floats(env, uv, (size_t)tris * 6) : nulls));
    CHECK_NAPI(napi_set_named_property(env, m, "material", make_number(env, mat)));
    CHECK_NAPI(napi_set_element(env, meshes, i, m));
  }
  CHECK_NAPI(napi_create_array_with_length(env, sum.material_count, &materials));
  for (uint32_t i = 0; i < sum.material_count; i++) {
    void *t;
    napi_value ta = new_typed(env, napi_float64_array, 40, 8, &t);
    if (!ta) return NULL;
    CHECK_LUCENT(lucent_gltf_material(doc, i, (double *)t));
    CHECK_NAPI(napi_set_element(env, materials, i, ta));
  }
  CHECK_NAPI(napi_create_array_with_length(env, sum.image_count, &images));
  for (uint32_t i = 0; i < sum.image_count; i++) {
    const uint8_t *d;
    size_t len;
    const char *mime, *name;
    CHECK_LUCENT(lucent_gltf_image(doc, i, &d, &len, &mime, &name));
    void *dst;
    napi_value ta = new_typed(env, napi_uint8_array, len, 1, &dst), o, sv;
    if (!ta) return NULL;
    if (len) memcpy(dst, d, len);
    CHECK_NAPI(napi_create_object(env, &o));
    CHECK_NAPI(napi_set_named_property(env, o, "data", ta));
    CHECK_NAPI(napi_create_string_utf8(env, mime, NAPI_AUTO_LENGTH, &sv));
    CHECK_NAPI(napi_set_named_property(env, o, "mimeType", sv));
    CHECK_NAPI(napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &sv));
    CHECK_NAPI(napi_set_named_property(env, o, "name", sv));
    CHECK_NAPI(napi_set_element(env, images, i, o));
  }
  CHECK_NAPI(napi_create_array_with_length(env, sum.warning_count, &warnings));
  for (uint32_t i = 0; i < sum.warning_count; i++) {
    napi_value w;
    CHECK_NAPI(napi_create_string_utf8(env, lucent_gltf_warning(doc, i), NAPI_AUTO_LENGTH, &w));
    CHECK_NAPI(napi_set_element(env, warnings, i, w));
  }
  CHECK_NAPI(napi_create_array_with_length(env, 3, &bmin));
  CHECK_NAPI(napi_create_array_with_length(env, 3, &bmax));
  for (int i = 0; i < 3; i++) {
    napi_set_element(env, bmin, i, make_number(env, sum.bounds[i]));
    napi_set_element(env, bmax, i, make_number(env, sum.bounds[3 + i]));
  }
  CHECK_NAPI(napi_create_object(env, &bounds));
  CHECK_NAPI(napi_set_named_property(env, bounds, "min", bmin));
  CHECK_NAPI(napi_set_named_property(env, bounds, "max", bmax));
  CHECK_NAPI(napi_set_named_property(env, out, "meshes", meshes));
  CHECK_NAPI(napi_set_named_property(env, out, "materials", materials));
  CHECK_NAPI(napi_set_named_property(env, out, "images", images));
  CHECK_NAPI(napi_set_named_property(env, out, "bounds", bounds));
  CHECK_NAPI(napi_set_named_property(env, out, "triangleCount", make_number(env, (double)sum.triangle_count)));
  CHECK_NAPI(napi_set_named_property(env, out, "warnings", warnings));
  return out;
}

/* RenderSettings fields of the integrator (renderer.ts:42-84). */
static void read_settings(napi_env env, napi_value o, lucent_render_settings *rs) {
  rs->max_bounces = (uint32_t)num(env, prop(env, o, "maxBounces"), 8);
  rs->firefly_clamp = (float)num(env, prop(env, o, "fireflyClamp"), 8);
  rs->flags = (num(env, prop(env, o, "perSampleLaunch"), 0) != 0 ? LUCENT_TRACE_PER_SAMPLE : 0u) |
              (num(env, prop(env, o, "wavefront"), 0) != 0 ? LUCENT_TRACE_WAVEFRONT : 0u) |
              (num(env, prop(env, o, "inlineShadows"), 0) != 0 ? LUCENT_TRACE_INLINE_SHADOWS : 0u);
  rs->waves_per_cu = (uint32_t)num(env, prop(env, o, "wavesPerCU"), 0);
  rs->occupancy = (uint32_t)num(env, prop(env, o, "occupancy"), 0);
  rs->tile_w = (uint32_t)num(env, prop(env, o, "tileWidth"), 0);
  rs->tile_h = (uint32_t)num(env, prop(env, o, "tileHeight"), 0);
  rs->shade_batch = (uint32_t)num(env, prop(env, o, "shadeBatch"), 0);
  rs->trace_turns = (uint32_t)num(env, prop(env, o, "traceTurns"), 0);
  rs->job_paths = (uint32_t)num(env, prop(env, o, "jobPaths"), 0);
}

/* traceReprojected(ctx, scene, film, camera, prevFilm, prevCamera, historyCap, settings):
 * the uReproject = 1 sample of traceSamples (renderer.ts:715-730). */
static napi_value js_trace_reprojected(napi_env env, napi_callback_info info) {
  size_t argc = 8;
  napi_value argv[8];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  scene_holder *sh = argc > 1 ? (scene_holder *)get_external(env, argv[1]) : NULL;
  film_holder *fh = argc > 2 ? (film_holder *)get_external(env, argv[2]) : NULL;
  film_holder *ph = argc > 4 ? (film_holder *)get_external(env, argv[4]) : NULL;
  lucent_camera cam, prev;
  if (!ctx || !sh || !fh || !ph || argc < 8 || !read_camera(env, argv[3], &cam) || !read_camera(env, argv[5], &prev)) {
    napi_throw_type_error(env, NULL, "traceReprojected(ctx, scene, film, camera, prevFilm, prevCamera, historyCap, settings)");
    return NULL;
  }
  lucent_render_settings rs;
  read_settings(env, argv[7], &rs);
  CHECK_LUCENT(lucent_trace_reprojected(ctx, sh->scene, fh->film, &cam, ph->film, &prev, (float)num(env, argv[6], 8), &rs,
                                        NULL));
  CHECK_LUCENT(lucent_synchronize(ctx));
  return NULL;
}

/* screenshot(ctx, film, {denoiseWeight, exposure, tonemap}, viewW, viewH) -> Uint8Array RGBA8
 * (rows bottom to top): the display pass (display.frag). */
static napi_value js_screenshot(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  film_holder *fh = argc > 1 ? (film_holder *)get_external(env, argv[1]) : NULL;
  if (!ctx || !fh || argc < 5) {
    napi_throw_type_error(env, NULL, "screenshot(ctx, film, params, viewW, viewH)");
    return NULL;
  }
  lucent_display_params p;
  p.denoise_weight = (float)num(env, prop(env, argv[2], "denoiseWeight"), 0);
  p.exposure = (float)num(env, prop(env, argv[2], "exposure"), 1);
  p.tonemap = (uint32_t)num(env, prop(env, argv[2], "tonemap"), 0);
  uint32_t vw = (uint32_t)num(env, argv[3], 0), vh = (uint32_t)num(env, argv[4], 0);
  void *px;
  napi_value out = new_typed(env, napi_uint8_array, (size_t)vw * vh * 4, 1, &px);
  CHECK_LUCENT(lucent_screenshot(ctx, fh->film, &p, vw, vh, (uint8_t *)px));
  return out;
}

static napi_value js_trace_samples(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  scene_holder *sh = argc > 1 ? (scene_holder *)get_external(env, argv[1]) : NULL;
  film_holder *fh = argc > 2 ? (film_holder *)get_external(env, argv[2]) : NULL;
  lucent_camera cam;
  if (!ctx || !sh || !fh || argc < 6 || !read_camera(env, argv[3], &cam)) {
    napi_throw_type_error(env, NULL, "traceSamples(ctx, scene, film, camera, settings, count)");
    return NULL;
  }
  lucent_render_settings rs;
  read_settings(env, argv[4], &rs);
  CHECK_LUCENT(lucent_trace_samples(ctx, sh->scene, fh->film, &cam, &rs, (uint32_t)num(env, argv[5], 1), NULL));
  CHECK_LUCENT(lucent_synchronize(ctx));
  return NULL;
}

/* readFilm(ctx, film) -> {color, albedo, normal, history} (rows bottom to top). */
static napi_value js_read_film(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  film_holder *fh = argc > 1 ? (film_holder *)get_external(env, argv[1]) : NULL;
  if (!ctx || !fh) {
    napi_throw_type_error(env, NULL, "readFilm(ctx, film)");
    return NULL;
  }
  uint32_t w, h, n, s;
  CHECK_LUCENT(lucent_film_info(fh->film, &w, &h, &n, &s));
  size_t px = (size_t)w * h;
  void *c, *a, *nm, *hi;
  napi_value ca = new_typed(env, napi_float32_array, px * 4, 4, &c), aa = new_typed(env, napi_float32_array, px * 4, 4, &a),
             na = new_typed(env, napi_float32_array, px * 4, 4, &nm), ha = new_typed(env, napi_float32_array, px, 4, &hi);
  CHECK_LUCENT(lucent_film_read(ctx, fh->film, (float *)c, (float *)a, (float *)nm, (float *)hi));
  napi_value out;
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "color", ca);
  napi_set_named_property(env, out, "albedo", aa);
  napi_set_named_property(env, out, "normal", na);
  napi_set_named_property(env, out, "history", ha);
  return out;
}

/* denoise(ctx, film, camera, {sigmaL, sigmaN, sigmaP, iterations}) -> {illumVar, radiance} (runDenoiser). */
static napi_value js_denoise(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  film_holder *fh = argc > 1 ? (film_holder *)get_external(env, argv[1]) : NULL;
  lucent_camera cam;
  if (!ctx || !fh || argc < 3 || !read_camera(env, argv[2], &cam)) {
    napi_throw_type_error(env, NULL, "denoise(ctx, film, camera, params)");
    return NULL;
  }
  lucent_denoise_params p = {4.0f, 64.0f, 1.5f, 0};
  if (argc > 3) {
    p.sigma_l = (float)num(env, prop(env, argv[3], "sigmaL"), 4);
    p.sigma_n = (float)num(env, prop(env, argv[3], "sigmaN"), 64);
    p.sigma_p = (float)num(env, prop(env, argv[3], "sigmaP"), 1.5);
    p.iterations = (uint32_t)num(env, prop(env, argv[3], "iterations"), 0);
  }
  CHECK_LUCENT(lucent_denoise(ctx, fh->film, &cam, &p, NULL));
  uint32_t w, h, n, s;
  CHECK_LUCENT(lucent_film_info(fh->film, &w, &h, &n, &s));
  size_t px = (size_t)w * h;
  void *iv, *rad;
  napi_value ia = new_typed(env, napi_float32_array, px * 4, 4, &iv), ra = new_typed(env, napi_float32_array, px * 4, 4, &rad);
  CHECK_LUCENT(lucent_film_read_denoised(ctx, fh->film, (float *)iv, (float *)rad));
  napi_value out;
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "illumVar", ia);
  napi_set_named_property(env, out, "radiance", ra);
  return out;
}

/* ---- asynchronous calls --------------------------------------------------
 * buildBVHAsync (bvh-async.ts:8-30), traceSamplesAsync and denoiseAsync run
 * the C-ABI call on a libuv worker thread (napi_async_work) and settle a
 * Promise; the arguments' JavaScript objects are referenced until then.
 * A context is not re-entrant: the JS facade queues calls per renderer. */
typedef enum { ASYNC_BVH, ASYNC_TRACE, ASYNC_DENOISE } async_kind;
typedef struct {
  napi_async_work work;
  napi_deferred deferred;
  async_kind kind;
  napi_ref refs[4];
  int nrefs;
  lucent_status status;
  char msg[512];
  /* buildBVH */
  const float *pos;
  uint32_t tris, max_leaf, bins, count;
  float *nodes;
  uint32_t *idx;
  double bounds[6];
  /* trace / denoise */
  lucent_ctx *ctx;
  lucent_scene *scene;
  lucent_film *film;
  lucent_camera cam;
  lucent_render_settings rs;
  uint32_t samples;
  lucent_denoise_params dp;
  float *iv, *rad;
  size_t px;
} async_call;

static void async_execute(napi_env env, void *data) {
  async_call *a = (async_call *)data;
  (void)env;
  switch (a->kind) {
    case ASYNC_BVH:
      a->status = lucent_build_bvh(a->pos, a->tris, a->max_leaf, a->bins, a->nodes, &a->count, a->idx, a->bounds);
      break;
    case ASYNC_TRACE:
      a->status = lucent_trace_samples(a->ctx, a->scene, a->film, &a->cam, &a->rs, a->samples, NULL);
      if (a->status == LUCENT_OK) a->status = lucent_synchronize(a->ctx);
      break;
    case ASYNC_DENOISE:
      a->status = lucent_denoise(a->ctx, a->film, &a->cam, &a->dp, NULL);
      if (a->status == LUCENT_OK) a->status = lucent_film_read_denoised(a->ctx, a->film, a->iv, a->rad);
      break;
  }
  /* The message is thread-local: read it on this thread. */
  if (a->status != LUCENT_OK) snprintf(a->msg, sizeof a->msg, "%s", lucent_last_error(a->ctx));
}

static void async_complete(napi_env env, napi_status status, void *data) {
  async_call *a = (async_call *)data;
  if (a->status != LUCENT_OK || status != napi_ok) {
    char code[32];
    snprintf(code, sizeof code, "LUCENT_%d", (int)a->status);
    napi_value c, m, err;
    napi_create_string_utf8(env, code, NAPI_AUTO_LENGTH, &c);
    napi_create_string_utf8(env, status != napi_ok ? "asynchronous call cancelled" : a->msg, NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, c, m, &err);
    napi_reject_deferred(env, a->deferred, err);
  } else {
    napi_value result = NULL;
    if (a->kind == ASYNC_BVH) {
      void *idx_data;
      napi_value idx = new_typed(env, napi_uint32_array, a->tris, 4, &idx_data);
      if (idx) memcpy(idx_data, a->idx, (size_t)a->tris * sizeof(uint32_t));
      result = idx ? bvh_result(env, a->nodes, a->count, idx, a->bounds) : NULL;
    } else if (a->kind == ASYNC_DENOISE) {
      void *iv, *rad;
      napi_value ia = new_typed(env, napi_float32_array, a->px * 4, 4, &iv);
      napi_value ra = new_typed(env, napi_float32_array, a->px * 4, 4, &rad);
      if (ia && ra) {
        memcpy(iv, a->iv, a->px * 16);
        memcpy(rad, a->rad, a->px * 16);
        napi_create_object(env, &result);
        napi_set_named_property(env, result, "illumVar", ia);
        napi_set_named_property(env, result, "radiance", ra);
      }
    } else {
      napi_get_undefined(env, &result);
    }
    if (result) {
      napi_resolve_deferred(env, a->deferred, result);
    } else {
      napi_value c, m, err;
      napi_create_string_utf8(env, "LUCENT_ASYNC", NAPI_AUTO_LENGTH, &c);
      napi_create_string_utf8(env, "could not build the result", NAPI_AUTO_LENGTH, &m);
      napi_create_error(env, c, m, &err);
      napi_reject_deferred(env, a->deferred, err);
    }
  }
  for (int i = 0; i < a->nrefs; i++) napi_delete_reference(env, a->refs[i]);
  free(a->nodes);
  free(a->idx);
  free(a->iv);
  free(a->rad);
  napi_delete_async_work(env, a->work);
  free(a);
}

static napi_value start_async(napi_env env, async_call *a, napi_value *args, int nargs, const char *name) {
  napi_value promise, res_name;
  for (int i = 0; i < nargs; i++) {
    if (napi_create_reference(env, args[i], 1, &a->refs[a->nrefs]) == napi_ok) a->nrefs++;
  }
  if (napi_create_promise(env, &a->deferred, &promise) != napi_ok ||
      napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &res_name) != napi_ok ||
      napi_create_async_work(env, NULL, res_name, async_execute, async_complete, a, &a->work) != napi_ok ||
      napi_queue_async_work(env, a->work) != napi_ok) {
    for (int i = 0; i < a->nrefs; i++) napi_delete_reference(env, a->refs[i]);
    free(a->nodes);
    free(a->idx);
    free(a->iv);
    free(a->rad);
    free(a);
    napi_throw_error(env, NULL, "could not start the asynchronous call");
    return NULL;
  }
  return promise;
}

/* buildBVHAsync(positions, options) -> Promise<BVH> (bvh-async.ts:8-30). */
static napi_value js_build_bvh_async(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  void *pos;
  size_t len;
  if (argc < 1 || !typed(env, argv[0], napi_float32_array, &pos, &len)) {
    napi_throw_type_error(env, NULL, "buildBVHAsync: positions must be a Float32Array");
    return NULL;
  }
  if (len % 9 != 0) {
    char msg[96];
    snprintf(msg, sizeof msg, "positions length %zu is not a multiple of 9", len);
    napi_throw_error(env, NULL, msg);
    return NULL;
  }
  async_call *a = (async_call *)calloc(1, sizeof(async_call));
  if (!a) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  a->kind = ASYNC_BVH;
  a->pos = (const float *)pos;
  a->tris = (uint32_t)(len / 9);
  a->max_leaf = argc > 1 ? (uint32_t)num(env, prop(env, argv[1], "maxLeafSize"), 4) : 4;
  a->bins = argc > 1 ? (uint32_t)num(env, prop(env, argv[1], "bins"), 12) : 12;
  a->nodes = (float *)malloc((a->tris > 0 ? 2 * (size_t)a->tris - 1 : 1) * 8 * sizeof(float));
  a->idx = (uint32_t *)malloc((a->tris > 0 ? a->tris : 1) * sizeof(uint32_t));
  if (!a->nodes || !a->idx) {
    free(a->nodes);
    free(a->idx);
    free(a);
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  return start_async(env, a, argv, 1, "lucent.buildBVHAsync");
}

/* traceSamplesAsync(ctx, scene, film, camera, settings, count) -> Promise<void>. */
static napi_value js_trace_samples_async(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  scene_holder *sh = argc > 1 ? (scene_holder *)get_external(env, argv[1]) : NULL;
  film_holder *fh = argc > 2 ? (film_holder *)get_external(env, argv[2]) : NULL;
  async_call *a = (async_call *)calloc(1, sizeof(async_call));
  if (!a) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  if (!ctx || !sh || !fh || argc < 6 || !read_camera(env, argv[3], &a->cam)) {
    free(a);
    napi_throw_type_error(env, NULL, "traceSamplesAsync(ctx, scene, film, camera, settings, count)");
    return NULL;
  }
  a->kind = ASYNC_TRACE;
  a->ctx = ctx;
  a->scene = sh->scene;
  a->film = fh->film;
  read_settings(env, argv[4], &a->rs);
  a->samples = (uint32_t)num(env, argv[5], 1);
  return start_async(env, a, argv, 3, "lucent.traceSamplesAsync");
}

/* denoiseAsync(ctx, film, camera, params) -> Promise<{illumVar, radiance}>. */
static napi_value js_denoise_async(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, NULL, NULL));
  lucent_ctx *ctx = argc > 0 ? get_ctx(env, argv[0]) : NULL;
  film_holder *fh = argc > 1 ? (film_holder *)get_external(env, argv[1]) : NULL;
  async_call *a = (async_call *)calloc(1, sizeof(async_call));
  if (!a) {
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  if (!ctx || !fh || argc < 3 || !read_camera(env, argv[2], &a->cam)) {
    free(a);
    napi_throw_type_error(env, NULL, "denoiseAsync(ctx, film, camera, params)");
    return NULL;
  }
  a->kind = ASYNC_DENOISE;
  a->ctx = ctx;
  a->film = fh->film;
  a->dp.sigma_l = 4.0f;
  a->dp.sigma_n = 64.0f;
  a->dp.sigma_p = 1.5f;
  a->dp.iterations = 0;
  if (argc > 3) {
    a->dp.sigma_l = (float)num(env, prop(env, argv[3], "sigmaL"), 4);
    a->dp.sigma_n = (float)num(env, prop(env, argv[3], "sigmaN"), 64);
    a->dp.sigma_p = (float)num(env, prop(env, argv[3], "sigmaP"), 1.5);
    a->dp.iterations = (uint32_t)num(env, prop(env, argv[3], "iterations"), 0);
  }
  uint32_t w, h, n, sd;
  lucent_film_info(fh->film, &w, &h, &n, &sd);
  a->px = (size_t)w * h;
  a->iv = (float *)malloc(a->px * 16);
  a->rad = (float *)malloc(a->px * 16);
  if (!a->iv || !a->rad) {
    free(a->iv);
    free(a->rad);
    free(a);
    napi_throw_error(env, NULL, "out of memory");
    return NULL;
  }
  return start_async(env, a, argv, 2, "lucent.denoiseAsync");
}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"version", 0, js_version, 0, 0, 0, napi_default, 0},
      {"buildBVH", 0, js_build_bvh, 0, 0, 0, napi_default, 0},
      {"buildBVHAsync", 0, js_build_bvh_async, 0, 0, 0, napi_default, 0},
      {"traceSamplesAsync", 0, js_trace_samples_async, 0, 0, 0, napi_default, 0},
      {"denoiseAsync", 0, js_denoise_async, 0, 0, 0, napi_default, 0},
      {"decodeHDR", 0, js_decode_hdr, 0, 0, 0, napi_default, 0},
      {"gltfParse", 0, js_gltf_parse, 0, 0, 0, napi_default, 0},
      {"gltfSetResource", 0, js_gltf_set_resource, 0, 0, 0, napi_default, 0},
      {"gltfBuild", 0, js_gltf_build, 0, 0, 0, napi_default, 0},
      {"createContext", 0, js_create_context, 0, 0, 0, napi_default, 0},
      {"sceneFromBVH", 0, js_scene_from_bvh, 0, 0, 0, napi_default, 0},
      {"uploadScene", 0, js_upload_scene, 0, 0, 0, napi_default, 0},
      {"setLighting", 0, js_set_lighting, 0, 0, 0, napi_default, 0},
      {"intersect", 0, js_intersect, 0, 0, 0, napi_default, 0},
      {"createFilm", 0, js_create_film, 0, 0, 0, napi_default, 0},
      {"resetFilm", 0, js_reset_film, 0, 0, 0, napi_default, 0},
      {"filmInfo", 0, js_film_info, 0, 0, 0, napi_default, 0},
      {"traceSamples", 0, js_trace_samples, 0, 0, 0, napi_default, 0},
      {"traceReprojected", 0, js_trace_reprojected, 0, 0, 0, napi_default, 0},
      {"screenshot", 0, js_screenshot, 0, 0, 0, napi_default, 0},
      {"readFilm", 0, js_read_film, 0, 0, 0, napi_default, 0},
      {"denoise", 0, js_denoise, 0, 0, 0, napi_default, 0},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
