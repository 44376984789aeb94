/*
 * TEST INFRASTRUCTURE — the parity oracle, not the product.
 *
 * FP64 restatement of Lucent's per-sample shader path and its host-side
 * table producers:
 *   pcg/rand/basis/samplers/offsetRay   src/shaders/common.glsl:10-108
 *   traceScene/occluded/alphaOpaque     src/shaders/scene.glsl:136-377
 *   environment and light sampling      src/shaders/lights.glsl:25-158
 *   layered BSDF                        src/shaders/bsdf.glsl:28-248
 *   cameraRay/surfaceAt/tracePath/main  src/shaders/trace.frag:53-272, 324-355
 *   buildAliasTable                     src/math/alias.ts:17-64
 *   buildEnvSampler                     src/scene/environment.ts:52-85
 *   sheenAlbedo(LUT)                    src/scene/sheen.ts:24-61
 *
 * Everything is double precision except what the reference fixes in float32
 * storage: the scene texels, the uniforms (passed in as float32 values),
 * the random numbers (float(pcg) * 2^-32, so both sides draw identical
 * values) and the half-precision sheen table.  Closest hits are found by an
 * exhaustive-order FP64 traversal (lowest t wins), so the oracle is
 * independent of the product's node layout and traversal order.
 *
 * Only tests/ and __graft_entry__.smoke() may load this library.
 *
 * Parity pins: the RNG against the published PCG hash constants and the
 * reference seeding formula; the BSDF against closed forms (GGX
 * normalisation, the Lambert value at normal incidence, the dielectric
 * sampler's reflection frequency = Schlick) and the sampler's histogram
 * against the evaluated pdf; alias tables against the reference's known
 * answers (test/environment.test.ts:16-44, 112-136); the sheen table against
 * its Monte-Carlo test (test/environment.test.ts:160-183); the integrator
 * (or_trace_pixels) against what the mathematics fixes
 * (tests/test_oracle_integrator.py): an empty scene and a directly viewed
 * emitter exactly, a lone convex sphere's directional albedo under a
 * constant environment (NEE, MIS and BSDF sampling together) against an
 * independent quadrature, a Lambert plane under a quad light against the
 * rectangle form factor, the environment pdf integrated over the sphere,
 * and alpha-masked hits against brute force.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  const float *geo;  /* 12 floats per leaf slot */
  const float *attr; /* 16 floats per slot or NULL */
  uint32_t tri_count;
  const float *nodes; /* reference layout, 8 floats per node */
  uint32_t node_count;
  const float *materials; /* 40 floats per material */
  uint32_t material_count;
  const float *spheres;
  const int32_t *sphere_mat;
  const int32_t *sphere_light;
  uint32_t sphere_count;
  const float *lights; /* 16 floats per light */
  uint32_t light_count;
  int32_t ground_material;
  float ground_height;
  const float *env_rgb; /* rows top to bottom */
  int32_t env_w, env_h;
  const float *env_table;
  int32_t grid_w, grid_h;
  float env_cos, env_sin, env_intensity;
  int32_t env_sampling, env_background;
  float backdrop[3];
  int32_t sun_enabled;
  float sun_dir[3], sun_radiance[3], sun_cos, sun_pdf;
  const float *sheen_lut; /* 16x16 half-rounded */
  const uint32_t *color_maps;
  const uint32_t *data_maps;
  int32_t tex_size, color_layers, data_layers;
} or_scene;

typedef struct {
  uint32_t width, height, max_bounces;
  float clamp;
  float cam_pos[3], cam_right[3], cam_up[3], cam_fwd[3];
  float tan_half_fov, aperture, focus_dist;
} or_frame;

/* ------------------------------------------------------------------------ */
typedef struct {
  double x, y, z;
} d3;
static d3 V(double x, double y, double z) {
  d3 r = {x, y, z};
  return r;
}
static d3 Vf(const float *p) { return V(p[0], p[1], p[2]); }
static d3 add(d3 a, d3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub(d3 a, d3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 mul(d3 a, d3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static d3 scl(d3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }
static d3 neg(d3 a) { return V(-a.x, -a.y, -a.z); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross(d3 a, d3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static double len(d3 a) { return sqrt(dot(a, a)); }
static d3 nrm(d3 a) { return scl(a, 1.0 / len(a)); }
static double maxd(double a, double b) { return a > b ? a : b; }
static double mind(double a, double b) { return a < b ? a : b; }
static double clampd(double x, double a, double b) { return mind(maxd(x, a), b); }
static double max3(d3 v) { return maxd(v.x, maxd(v.y, v.z)); }
static double mixd(double a, double b, double t) { return a * (1 - t) + b * t; }
static d3 mix3(d3 a, d3 b, double t) { return add(scl(a, 1 - t), scl(b, t)); }
static double lum(d3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
static d3 reflect3(d3 i, d3 n) { return sub(i, scl(n, 2 * dot(n, i))); }
static d3 refract3(d3 i, d3 n, double eta) {
  double d = dot(n, i);
  double k = 1 - eta * eta * (1 - d * d);
  if (k < 0) return V(0, 0, 0);
  return sub(scl(i, eta), scl(n, eta * d + sqrt(k)));
}

static const double PI = 3.14159265358979323846;
static const double TWO_PI = 6.28318530717958647692;
static const double INF_D = 1e30;
static const double EPS_D = 1e-6;

/* common.glsl:13-22 */
typedef struct {
  uint32_t s;
} rng_t;
static uint32_t pcg(rng_t *r) {
  uint32_t st = r->s;
  r->s = st * 747796405u + 2891336453u;
  uint32_t word = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
  return (word >> 22u) ^ word;
}
static double rnd(rng_t *r) { return (double)(float)pcg(r) * (1.0 / 4294967296.0); }

uint32_t or_pcg_stream(uint32_t pixel_index, uint32_t seed, uint32_t n, uint32_t *out) {
  rng_t r = {pixel_index * 1973u + seed * 9277u + 26699u};
  pcg(&r);
  for (uint32_t i = 0; i < n; i++) out[i] = pcg(&r);
  return r.s;
}

static void basis(d3 n, d3 *t, d3 *b) {
  double s = n.z >= 0 ? 1 : -1;
  double a = -1 / (s + n.z);
  double c = n.x * n.y * a;
  *t = V(1 + s * n.x * n.x * a, s * c, -s * n.x);
  *b = V(c, s + n.y * n.y * a, -n.y);
}
static d3 to_world(d3 v, d3 t, d3 b, d3 n) { return add(add(scl(t, v.x), scl(b, v.y)), scl(n, v.z)); }
static d3 to_local(d3 v, d3 t, d3 b, d3 n) { return V(dot(v, t), dot(v, b), dot(v, n)); }
static d3 cos_hemi(double u0, double u1) {
  double r = sqrt(u0), phi = TWO_PI * u1;
  return V(r * cos(phi), r * sin(phi), sqrt(maxd(0, 1 - u0)));
}
static d3 uni_cone(double u0, double u1, double cm) {
  double ct = 1 - u0 * (1 - cm);
  double st = sqrt(maxd(0, 1 - ct * ct));
  double phi = TWO_PI * u1;
  return V(st * cos(phi), st * sin(phi), ct);
}
static void concentric(double u0, double u1, double *x, double *y) {
  double ox = 2 * u0 - 1, oy = 2 * u1 - 1, r, th;
  if (ox == 0 && oy == 0) {
    *x = *y = 0;
    return;
  }
  if (fabs(ox) > fabs(oy)) {
    r = ox;
    th = (PI / 4) * (oy / ox);
  } else {
    r = oy;
    th = (PI / 2) - (PI / 4) * (ox / oy);
  }
  *x = r * cos(th);
  *y = r * sin(th);
}
static double mis(double a, double b) {
  double a2 = a * a, b2 = b * b;
  return a2 / maxd(a2 + b2, 1e-10);
}
/* offsetRay (common.glsl:95-108): integer ulp steps on the float32 position. */
static double off_axis(double p, double n) {
  float pf = (float)p;
  int32_t o = (int32_t)(256.0f * (float)n);
  int32_t bits;
  memcpy(&bits, &pf, 4);
  bits += pf < 0.0f ? -o : o;
  float pi;
  memcpy(&pi, &bits, 4);
  return fabs(pf) < 1.0 / 32.0 ? (double)pf + (1.0 / 65536.0) * n : (double)pi;
}
static d3 offset_ray(d3 p, d3 n) { return V(off_axis(p.x, n.x), off_axis(p.y, n.y), off_axis(p.z, n.z)); }

/* ---- textures ----------------------------------------------------------- */
static double srgb(double c) { return c <= 0.04045 ? c / 12.92 : pow((c + 0.055) / 1.055, 2.4); }
static void texel8(const uint32_t *maps, int size, int layer, int x, int y, int is_srgb, double *o) {
  x = ((x % size) + size) % size;
  y = ((y % size) + size) % size;
  uint32_t p = maps[((size_t)layer * size + y) * size + x];
  o[0] = (p & 255u) / 255.0;
  o[1] = ((p >> 8) & 255u) / 255.0;
  o[2] = ((p >> 16) & 255u) / 255.0;
  o[3] = (p >> 24) / 255.0;
  if (is_srgb)
    for (int k = 0; k < 3; k++) o[k] = srgb(o[k]);
}
static void sample_layer(const uint32_t *maps, int size, int layer, double u, double v, int is_srgb, double *o) {
  double x = u * size - 0.5, y = v * size - 0.5, fx = floor(x), fy = floor(y), tx = x - fx, ty = y - fy;
  double a[4], b[4], c[4], d[4];
  texel8(maps, size, layer, (int)fx, (int)fy, is_srgb, a);
  texel8(maps, size, layer, (int)fx + 1, (int)fy, is_srgb, b);
  texel8(maps, size, layer, (int)fx, (int)fy + 1, is_srgb, c);
  texel8(maps, size, layer, (int)fx + 1, (int)fy + 1, is_srgb, d);
  for (int k = 0; k < 4; k++) o[k] = mixd(mixd(a[k], b[k], tx), mixd(c[k], d[k], tx), ty);
}

/* ---- scene.glsl --------------------------------------------------------- */
static const float *mat_ptr(const or_scene *s, int i) { return s->materials + (size_t)i * 40; }

static int hit_tri(d3 ro, d3 rd, d3 v0, d3 v1, d3 v2, double tmax, double *t, double *bu, double *bv) {
  d3 e1 = sub(v1, v0), e2 = sub(v2, v0), p = cross(rd, e2);
  double det = dot(e1, p);
  if (fabs(det) < 1e-12) return 0;
  double inv = 1 / det;
  d3 sv = sub(ro, v0);
  double u = dot(sv, p) * inv;
  if (u < 0 || u > 1) return 0;
  d3 q = cross(sv, e1);
  double v = dot(rd, q) * inv;
  if (v < 0 || u + v > 1) return 0;
  double tt = dot(e2, q) * inv;
  if (tt <= EPS_D || tt >= tmax) return 0;
  *t = tt;
  *bu = u;
  *bv = v;
  return 1;
}
static double hit_sphere(d3 ro, d3 rd, const float *s4, double tmax) {
  d3 oc = sub(ro, Vf(s4));
  double b = dot(oc, rd), c = dot(oc, oc) - (double)s4[3] * s4[3], disc = b * b - c;
  if (disc < 0) return INF_D;
  double sq = sqrt(disc), t = -b - sq;
  if (t > EPS_D && t < tmax) return t;
  t = -b + sq;
  if (t > EPS_D && t < tmax) return t;
  return INF_D;
}
static double hit_ground(const or_scene *s, d3 ro, d3 rd, double tmax) {
  if (s->ground_material < 0 || fabs(rd.y) < 1e-9) return INF_D;
  double t = (s->ground_height - ro.y) / rd.y;
  return t > EPS_D && t < tmax ? t : INF_D;
}
static void tri_uvs(const or_scene *s, uint32_t slot, double *uv) {
  if (!s->attr) {
    for (int k = 0; k < 6; k++) uv[k] = 0;
    return;
  }
  const float *a = s->attr + (size_t)slot * 16;
  uv[0] = a[3];
  uv[1] = a[12];
  uv[2] = a[7];
  uv[3] = a[13];
  uv[4] = a[11];
  uv[5] = a[14];
}
static int alpha_opaque(const or_scene *s, uint32_t slot, double bu, double bv, int mat, rng_t *r) {
  double uv[6];
  tri_uvs(s, slot, uv);
  double w = 1 - bu - bv;
  double uu = uv[0] * w + uv[2] * bu + uv[4] * bv, vv = uv[1] * w + uv[3] * bu + uv[5] * bv;
  const float *m = mat_ptr(s, mat);
  double alpha = m[11];
  int mode = (int)m[15];
  int map = (int)m[16];
  if (map >= 0 && s->color_maps && map < s->color_layers) {
    double tu = m[32] * uu + m[33] * vv + m[34], tv = m[36] * uu + m[37] * vv + m[38], o[4];
    sample_layer(s->color_maps, s->tex_size, map, tu, tv, 1, o);
    alpha *= o[3];
  }
  if (mode == 1) return alpha >= m[23];
  return rnd(r) < alpha;
}
static double slab(const float *n, d3 o, d3 inv, double tmax) {
  double t0x = (n[0] - o.x) * inv.x, t1x = (n[4] - o.x) * inv.x;
  double t0y = (n[1] - o.y) * inv.y, t1y = (n[5] - o.y) * inv.y;
  double t0z = (n[2] - o.z) * inv.z, t1z = (n[6] - o.z) * inv.z;
  double a = 0, b = tmax, lo, hi;
  lo = fmin(t0x, t1x);
  hi = fmax(t0x, t1x);
  if (lo == lo) a = fmax(a, lo);
  if (hi == hi) b = fmin(b, hi);
  lo = fmin(t0y, t1y);
  hi = fmax(t0y, t1y);
  if (lo == lo) a = fmax(a, lo);
  if (hi == hi) b = fmin(b, hi);
  lo = fmin(t0z, t1z);
  hi = fmax(t0z, t1z);
  if (lo == lo) a = fmax(a, lo);
  if (hi == hi) b = fmin(b, hi);
  /* Slightly inflated: a rounding in the slab must never hide a hit. */
  return b * (1 + 1e-9) + 1e-12 >= a ? a : INF_D;
}
/* Fetch counters for the algorithmic-bytes model (per calling thread):
 * 0 paths, 1 closest rays, 2 closest interior visits (64-byte child pairs),
 * 3 closest triangle tests, 4 shadow rays, 5 shadow node visits (32 bytes),
 * 6 shadow triangle tests, 7 surface vertices, 8 environment NEE samples,
 * 9 escaped rays, 10 light NEE samples. */
static _Thread_local double *g_counts = NULL;
#define COUNT(i, v) \
  do {                \
    if (g_counts) g_counts[i] += (v); \
  } while (0)

/* Closest accepted triangle (t in (EPS, closest)) in the reference shader's
 * visiting order (scene.glsl:321-376: root box untested, near child first);
 * the FP64 slab is inflated so rounding never hides a hit, and the result is
 * the lowest t found, whatever the order. */
static int bvh_closest(const or_scene *s, d3 ro, d3 rd, double *closest, double *bu, double *bv, rng_t *r) {
  if (s->tri_count == 0) return -1;
  COUNT(1, 1);
  d3 inv = V(1 / rd.x, 1 / rd.y, 1 / rd.z);
  uint32_t stack[128];
  int sp = 0, best = -1;
  uint32_t node = 0;
  for (;;) {
    const float *nd = s->nodes + (size_t)node * 8;
    uint32_t first = (uint32_t)nd[3], count = (uint32_t)nd[7];
    if (count > 0 || s->node_count == 1) {
      COUNT(3, count);
      for (uint32_t k = first; k < first + count; k++) {
        const float *g = s->geo + (size_t)k * 12;
        double t, u, v;
        if (hit_tri(ro, rd, Vf(g), Vf(g + 4), Vf(g + 8), *closest, &t, &u, &v)) {
          if (g[11] >= 0.5f && !alpha_opaque(s, k, u, v, (int)g[3], r)) continue;
          *closest = t;
          best = (int)k;
          *bu = u;
          *bv = v;
        }
      }
      if (sp == 0) break;
      node = stack[--sp];
      continue;
    }
    COUNT(2, 1);
    uint32_t left = first, right = first + 1;
    double tl = slab(s->nodes + (size_t)left * 8, ro, inv, *closest);
    double tr = slab(s->nodes + (size_t)right * 8, ro, inv, *closest);
    if (tl > tr) {
      double tt = tl;
      tl = tr;
      tr = tt;
      uint32_t ti = left;
      left = right;
      right = ti;
    }
    if (tl == INF_D) {
      if (sp == 0) break;
      node = stack[--sp];
    } else {
      node = left;
      if (tr != INF_D) {
        if (sp >= 128) return -2;
        stack[sp++] = right;
      }
    }
  }
  return best;
}
static int occluded(const or_scene *s, d3 ro, d3 rd, double tmax, rng_t *r) {
  for (uint32_t i = 0; i < s->sphere_count; i++)
    if (hit_sphere(ro, rd, s->spheres + 4 * i, tmax) < tmax) return 1;
  if (hit_ground(s, ro, rd, tmax) < tmax) return 1;
  if (s->tri_count == 0) return 0;
  COUNT(4, 1);
  d3 inv = V(1 / rd.x, 1 / rd.y, 1 / rd.z);
  uint32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    uint32_t n = stack[--sp];
    const float *nd = s->nodes + (size_t)n * 8;
    COUNT(5, 1);
    if (slab(nd, ro, inv, tmax) == INF_D) continue;
    uint32_t first = (uint32_t)nd[3], count = (uint32_t)nd[7];
    if (count > 0 || s->node_count == 1) {
      for (uint32_t k = first; k < first + count; k++) {
        COUNT(6, 1);
        const float *g = s->geo + (size_t)k * 12;
        double t, u, v;
        if (hit_tri(ro, rd, Vf(g), Vf(g + 4), Vf(g + 8), tmax, &t, &u, &v) &&
            (g[11] < 0.5f || alpha_opaque(s, k, u, v, (int)g[3], r)))
          return 1;
      }
    } else {
      if (sp + 2 > 128) return 1;
      stack[sp++] = first + 1;
      stack[sp++] = first;
    }
  }
  return 0;
}

typedef struct {
  double t;
  d3 pos, n, ng, dpdu, dpdv;
  int front, material, light;
  double u, v;
} hit_t;

static int trace_scene(const or_scene *s, d3 ro, d3 rd, hit_t *h, rng_t *r) {
  double closest = INF_D, bu = 0, bv = 0;
  int best_sphere = -1, ground = 0;
  for (uint32_t i = 0; i < s->sphere_count; i++) {
    double t = hit_sphere(ro, rd, s->spheres + 4 * i, closest);
    if (t < closest) {
      closest = t;
      best_sphere = (int)i;
    }
  }
  double tg = hit_ground(s, ro, rd, closest);
  if (tg < closest) {
    closest = tg;
    best_sphere = -1;
    ground = 1;
  }
  int tri = bvh_closest(s, ro, rd, &closest, &bu, &bv, r);
  if (tri >= 0) {
    best_sphere = -1;
    ground = 0;
  }
  if (tri < 0 && best_sphere < 0 && !ground) return 0;
  h->t = closest;
  h->pos = add(ro, scl(rd, closest));
  h->u = h->v = 0;
  h->dpdu = h->dpdv = V(0, 0, 0);
  h->light = -1;
  if (ground) {
    h->front = rd.y < 0;
    h->n = V(0, h->front ? 1 : -1, 0);
    h->ng = h->n;
    h->material = s->ground_material;
    h->u = h->pos.x;
    h->v = h->pos.z;
    return 1;
  }
  if (best_sphere >= 0) {
    const float *sp = s->spheres + 4 * best_sphere;
    d3 n = scl(sub(h->pos, Vf(sp)), 1.0 / sp[3]);
    h->front = dot(rd, n) < 0;
    h->n = h->front ? n : neg(n);
    h->ng = h->n;
    h->material = s->sphere_mat[best_sphere];
    h->light = s->sphere_light[best_sphere];
    return 1;
  }
  const float *g = s->geo + (size_t)tri * 12;
  d3 p0 = Vf(g), p1 = Vf(g + 4), p2 = Vf(g + 8);
  d3 e1 = sub(p1, p0), e2 = sub(p2, p0);
  d3 ng = nrm(cross(e1, e2)), ns;
  double w = 1 - bu - bv;
  if (s->attr) {
    const float *a = s->attr + (size_t)tri * 16;
    double uv[6];
    tri_uvs(s, (uint32_t)tri, uv);
    h->u = uv[0] * w + uv[2] * bu + uv[4] * bv;
    h->v = uv[1] * w + uv[3] * bu + uv[5] * bv;
    double d1x = uv[2] - uv[0], d1y = uv[3] - uv[1], d2x = uv[4] - uv[0], d2y = uv[5] - uv[1];
    double det = d1x * d2y - d1y * d2x;
    if (fabs(det) > 1e-12) {
      double rr = 1 / det;
      h->dpdu = scl(sub(scl(e1, d2y), scl(e2, d1y)), rr);
      h->dpdv = scl(sub(scl(e2, d1x), scl(e1, d2x)), rr);
    }
    ns = nrm(add(add(scl(Vf(a), w), scl(Vf(a + 4), bu)), scl(Vf(a + 8), bv)));
  } else {
    ns = ng;
  }
  h->front = dot(rd, ng) < 0;
  if (!h->front) {
    ng = neg(ng);
    ns = neg(ns);
  }
  if (dot(ns, ng) < 0.05) ns = ng;
  h->n = ns;
  h->ng = ng;
  h->material = (int)g[3];
  h->light = (int)g[7];
  return 1;
}

/* ---- lights.glsl -------------------------------------------------------- */
typedef struct {
  d3 dir;
  double dist, pdf;
  d3 rad;
} lsample;

static d3 env_local(const or_scene *s, d3 d) {
  return V(s->env_cos * d.x - s->env_sin * d.z, d.y, s->env_sin * d.x + s->env_cos * d.z);
}
static d3 env_world(const or_scene *s, d3 l) {
  return V(s->env_cos * l.x + s->env_sin * l.z, l.y, -s->env_sin * l.x + s->env_cos * l.z);
}
static void eq_uv(d3 l, double *u, double *v) {
  *u = atan2(l.z, l.x) * (0.5 / PI) + 0.5;
  *v = acos(clampd(l.y, -1, 1)) / PI;
}
static d3 env_lookup(const or_scene *s, double u, double v) {
  double x = u * s->env_w - 0.5, y = v * s->env_h - 0.5, fx = floor(x), fy = floor(y), tx = x - fx, ty = y - fy;
  int x0 = ((int)fx % s->env_w + s->env_w) % s->env_w, x1 = (x0 + 1) % s->env_w;
  int y0 = (int)fy, y1 = (int)fy + 1;
  y0 = y0 < 0 ? 0 : (y0 >= s->env_h ? s->env_h - 1 : y0);
  y1 = y1 < 0 ? 0 : (y1 >= s->env_h ? s->env_h - 1 : y1);
  const float *m = s->env_rgb;
  d3 a = Vf(m + ((size_t)y0 * s->env_w + x0) * 3), b = Vf(m + ((size_t)y0 * s->env_w + x1) * 3);
  d3 c = Vf(m + ((size_t)y1 * s->env_w + x0) * 3), d = Vf(m + ((size_t)y1 * s->env_w + x1) * 3);
  return scl(mix3(mix3(a, b, tx), mix3(c, d, tx), ty), s->env_intensity);
}
static d3 env_radiance(const or_scene *s, d3 d) {
  double u, v;
  eq_uv(env_local(s, d), &u, &v);
  return env_lookup(s, u, v);
}
static double env_pdf(const or_scene *s, d3 d) {
  d3 l = env_local(s, d);
  double st = sqrt(maxd(0, 1 - l.y * l.y));
  if (st < 1e-6) return 0;
  double u, v;
  eq_uv(l, &u, &v);
  int cx = (int)(u * s->grid_w), cy = (int)(v * s->grid_h);
  if (cx > s->grid_w - 1) cx = s->grid_w - 1;
  if (cy > s->grid_h - 1) cy = s->grid_h - 1;
  return s->env_table[((size_t)cy * s->grid_w + cx) * 4 + 2] / (2 * PI * PI * st);
}
/* envPdf of a world direction (lights.glsl:49-56), exported for the
 * sphere-integral check. */
double or_env_pdf_dir(const or_scene *s, const double *d) { return env_pdf(s, V(d[0], d[1], d[2])); }
void or_env_pdf_batch(const or_scene *s, uint32_t n, const double *d, double *out) {
  for (uint32_t i = 0; i < n; i++) out[i] = env_pdf(s, V(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
}

/* bakeSky (environment.ts:106-132): the procedural gradient sky plus the
 * sun's glow (not its disc) in a width x width/2 equirect map, float32 rows
 * top to bottom.  Each stage stores into the float32 map as the reference's
 * Float32Array does: the gradient, then + colour * glow, then * intensity.
 * mix(a, b, t) = a + (b - a) * t as in the reference (not GLSL's form). */
void or_bake_sky(int enabled, const double *zenith, const double *horizon, const double *ground, double intensity,
                 double sun_elev_deg, double sun_azim_deg, const double *sun_color, double sun_intensity,
                 double sun_angular_radius_deg, int width, float *out) {
  int height = width / 2;
  for (size_t i = 0; i < (size_t)width * height * 3; i++) out[i] = 0.0f;
  if (!enabled || intensity <= 0) return;
  double el = sun_elev_deg * PI / 180, az = sun_azim_deg * PI / 180;
  d3 sd = V(cos(el) * sin(az), sin(el), cos(el) * cos(az));
  int sun_on = sun_intensity > 0 && sun_elev_deg > -sun_angular_radius_deg;
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) {
      double u = (x + 0.5) / width, v = (y + 0.5) / height;
      double phi = (u - 0.5) * 2 * PI, th = v * PI, st = sin(th);
      d3 d = V(st * cos(phi), cos(th), st * sin(phi));
      double t = d.y;
      float *o = out + ((size_t)y * width + x) * 3;
      for (int c = 0; c < 3; c++)
        o[c] = (float)(t >= 0 ? horizon[c] + (zenith[c] - horizon[c]) * pow(t, 0.55)
                              : horizon[c] + (ground[c] - horizon[c]) * mind(1, -t * 4));
      if (sun_on) {
        double sdot = maxd(0, d.x * sd.x + d.y * sd.y + d.z * sd.z);
        double glow = 0.12 * pow(sdot, 24) + 0.5 * pow(sdot, 400);
        for (int c = 0; c < 3; c++) o[c] = (float)((double)o[c] + sun_color[c] * glow);
      }
      for (int c = 0; c < 3; c++) o[c] = (float)((double)o[c] * intensity);
    }
}

static lsample sample_env(const or_scene *s, rng_t *r) {
  lsample ls;
  int n = s->grid_w * s->grid_h;
  /* int(rand() * float(n)) of the shader: a float32 product */
  int i = (int)((float)rnd(r) * (float)n);
  if (i > n - 1) i = n - 1;
  const float *e = s->env_table + (size_t)i * 4;
  int cell = i;
  double pdf_uv = e[2];
  if (rnd(r) >= e[0]) {
    cell = (int)e[1];
    pdf_uv = e[3];
  }
  double r0 = rnd(r), r1 = rnd(r);
  double u = ((cell % s->grid_w) + r0) / s->grid_w, v = ((cell / s->grid_w) + r1) / s->grid_h;
  double phi = (u - 0.5) * TWO_PI, th = v * PI, st = sin(th);
  ls.dir = env_world(s, V(st * cos(phi), cos(th), st * sin(phi)));
  ls.dist = INF_D;
  ls.pdf = st > 1e-6 ? pdf_uv / (2 * PI * PI * st) : 0;
  ls.rad = env_lookup(s, u, v);
  return ls;
}
static d3 environment(const or_scene *s, d3 d, int spec, double bpdf) {
  d3 L = env_radiance(s, d);
  if (!spec && s->env_sampling) L = scl(L, mis(bpdf, env_pdf(s, d)));
  if (s->sun_enabled && dot(d, Vf(s->sun_dir)) >= s->sun_cos) L = add(L, scl(Vf(s->sun_radiance), spec ? 1 : mis(bpdf, s->sun_pdf)));
  return L;
}
static lsample sample_light(const or_scene *s, int idx, d3 p, double u0, double u1) {
  lsample ls;
  const float *L = s->lights + idx * 16;
  ls.pdf = 0;
  ls.dir = V(0, 0, 0);
  ls.dist = 0;
  ls.rad = Vf(L + 12);
  if (L[3] < 0.5f) {
    d3 tc = sub(Vf(L), p);
    double d2 = dot(tc, tc), rr = L[7];
    if (d2 <= rr * rr) return ls;
    double cm = sqrt(maxd(0, 1 - rr * rr / d2));
    d3 axis = scl(tc, 1 / sqrt(d2)), t, b;
    basis(axis, &t, &b);
    ls.dir = to_world(uni_cone(u0, u1, cm), t, b, axis);
    ls.pdf = 1 / (TWO_PI * (1 - cm));
    float sp[4] = {L[0], L[1], L[2], L[7]};
    double t0 = hit_sphere(p, ls.dir, sp, INF_D);
    ls.dist = t0 < INF_D ? t0 : sqrt(d2);
  } else {
    d3 e1 = Vf(L + 4), e2 = Vf(L + 8);
    d3 q = add(add(Vf(L), scl(e1, u0)), scl(e2, u1));
    d3 tq = sub(q, p);
    double d2 = dot(tq, tq), d = sqrt(d2);
    ls.dir = scl(tq, 1 / d);
    ls.dist = d;
    d3 nc = cross(e1, e2);
    double area = len(nc);
    double cl = fabs(dot(scl(nc, 1 / area), ls.dir));
    if (cl < 1e-4) return ls;
    ls.pdf = d2 / (cl * area);
  }
  return ls;
}
static double light_pdf(const or_scene *s, int idx, d3 p, d3 dir, double dist) {
  const float *L = s->lights + idx * 16;
  if (L[3] < 0.5f) {
    d3 tc = sub(Vf(L), p);
    double d2 = dot(tc, tc), rr = L[7];
    if (d2 <= rr * rr) return 0;
    return 1 / (TWO_PI * (1 - sqrt(maxd(0, 1 - rr * rr / d2))));
  }
  d3 nc = cross(Vf(L + 4), Vf(L + 8));
  double area = len(nc);
  double cl = fabs(dot(scl(nc, 1 / area), dir));
  if (cl < 1e-4) return 0;
  return dist * dist / (cl * area);
}

/* ---- bsdf.glsl ---------------------------------------------------------- */
typedef struct {
  d3 albedo;
  double metallic, roughness, ior, transmission;
  d3 emission;
  double clearcoat, cc_rough;
  d3 sheen;
  double sheen_rough;
  int thin;
  d3 absorption, n, nc;
} surf_t;

double or_d_ggx(double ndh, double a) {
  double a2 = a * a, d = ndh * ndh * (a2 - 1) + 1;
  return a2 / (PI * d * d);
}
double or_g1_ggx(double ndx, double a) {
  double a2 = a * a;
  return 2 * ndx / (ndx + sqrt(a2 + (1 - a2) * ndx * ndx));
}
static d3 schlick3(double c, d3 f0) {
  double f = pow(1 - clampd(c, 0, 1), 5);
  return add(f0, scl(sub(V(1, 1, 1), f0), f));
}
static double schlick1(double c, double f0) { return f0 + (1 - f0) * pow(1 - clampd(c, 0, 1), 5); }
static double diel_f0(double ior) {
  double r = (ior - 1) / (ior + 1);
  return r * r;
}
static d3 vndf(d3 ve, double a, double u0, double u1) {
  d3 vh = nrm(V(a * ve.x, a * ve.y, ve.z));
  double lensq = vh.x * vh.x + vh.y * vh.y;
  d3 t1v = lensq > 0 ? scl(V(-vh.y, vh.x, 0), 1 / sqrt(lensq)) : V(1, 0, 0);
  d3 t2v = cross(vh, t1v);
  double r = sqrt(u0), phi = TWO_PI * u1, t1 = r * cos(phi), t2 = r * sin(phi), sv = 0.5 * (1 + vh.z);
  t2 = (1 - sv) * sqrt(maxd(0, 1 - t1 * t1)) + sv * t2;
  d3 nh = add(add(scl(t1v, t1), scl(t2v, t2)), scl(vh, sqrt(maxd(0, 1 - t1 * t1 - t2 * t2))));
  return nrm(V(a * nh.x, a * nh.y, maxd(0, nh.z)));
}
/* Visible-normal sample exposed for closed-form tests. */
void or_sample_vndf(const double *ve, double a, double u0, double u1, double *out) {
  d3 r = vndf(V(ve[0], ve[1], ve[2]), a, u0, u1);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
static double a_rough(double r) { return maxd(r * r, 1e-3); }
static double d_charlie(double alpha, double ndh) {
  double inv = 1 / alpha, s2 = maxd(1 - ndh * ndh, 0);
  return (2 + inv) * pow(s2, inv * 0.5) / TWO_PI;
}
static double v_neubelt(double l, double v) { return 1 / (4 * (l + v - l * v)); }
static double sheen_lookup(const or_scene *s, double ndv, double r) {
  double x = ndv * 15, y = r * 15, fx = floor(x), fy = floor(y), tx = x - fx, ty = y - fy;
  int x0 = (int)fx, y0 = (int)fy, x1 = x0 + 1, y1 = y0 + 1;
  x0 = x0 < 0 ? 0 : (x0 > 15 ? 15 : x0);
  x1 = x1 < 0 ? 0 : (x1 > 15 ? 15 : x1);
  y0 = y0 < 0 ? 0 : (y0 > 15 ? 15 : y0);
  y1 = y1 < 0 ? 0 : (y1 > 15 ? 15 : y1);
  const float *L = s->sheen_lut;
  return mixd(mixd(L[y0 * 16 + x0], L[y0 * 16 + x1], tx), mixd(L[y1 * 16 + x0], L[y1 * 16 + x1], tx), ty);
}
static double spec_prob(d3 f0, double metallic, double ndv) {
  return mixd(clampd(lum(schlick3(ndv, f0)), 0.1, 0.9), 1, metallic);
}
static double coat_prob(const surf_t *s, d3 wo) {
  if (s->clearcoat <= 0) return 0;
  double fc = s->clearcoat * schlick1(maxd(dot(s->nc, wo), 0), 0.04);
  double base = (1 - fc) * maxd(lum(s->albedo), 0.04);
  return clampd(fc / (fc + base), 0.15, 0.85);
}
static d3 base_f0(const surf_t *s) {
  double d = diel_f0(s->ior);
  return mix3(V(d, d, d), s->albedo, s->metallic);
}
static d3 eval_opaque(const or_scene *sc, const surf_t *s, d3 wo, d3 wi, double *pdf) {
  d3 n = s->n;
  double ndl = dot(n, wi), ndv = dot(n, wo);
  *pdf = 0;
  if (ndl <= 0 || ndv <= 0) return V(0, 0, 0);
  double a = a_rough(s->roughness);
  d3 H = nrm(add(wo, wi));
  double ndh = maxd(dot(n, H), 0), vdh = maxd(dot(wo, H), 0);
  d3 f0 = base_f0(s), F = schlick3(vdh, f0);
  double D = or_d_ggx(ndh, a), g1v = or_g1_ggx(ndv, a), G = g1v * or_g1_ggx(ndl, a);
  d3 spec = scl(F, D * G / maxd(4 * ndv * ndl, 1e-6));
  d3 diff = scl(mul(sub(V(1, 1, 1), F), s->albedo), (1 - s->metallic) / PI);
  d3 f = add(diff, spec);
  *pdf = mixd(ndl / PI, g1v * D / maxd(4 * ndv, 1e-6), spec_prob(f0, s->metallic, ndv));
  double smax = max3(s->sheen);
  if (smax > 0) {
    double sr = maxd(s->sheen_rough, 0.07);
    d3 sh = scl(s->sheen, d_charlie(sr * sr, ndh) * v_neubelt(ndl, ndv));
    double k = 1 - smax * maxd(sheen_lookup(sc, ndv, sr), sheen_lookup(sc, ndl, sr));
    f = add(sh, scl(f, k));
  }
  double pc = coat_prob(s, wo);
  if (pc > 0) {
    double cv = dot(s->nc, wo), cl = dot(s->nc, wi), coat = 0, pdfc = 0;
    if (cv > 0 && cl > 0) {
      double ac = a_rough(s->cc_rough), ch = maxd(dot(s->nc, H), 0), dc = or_d_ggx(ch, ac), g1c = or_g1_ggx(cv, ac);
      coat = dc * g1c * or_g1_ggx(cl, ac) * schlick1(vdh, 0.04) / (4 * cv * cl);
      pdfc = g1c * dc / (4 * cv);
    }
    double fcv = schlick1(maxd(cv, 0), 0.04);
    f = add(scl(f, 1 - s->clearcoat * fcv), scl(V(coat, coat, coat), s->clearcoat));
    *pdf = mixd(*pdf, pdfc, pc);
  }
  return f;
}
typedef struct {
  d3 dir, w;
  double pdf;
} bsample;
static bsample sample_opaque(const or_scene *sc, const surf_t *s, d3 wo, rng_t *r) {
  bsample bs;
  bs.pdf = 0;
  bs.w = V(0, 0, 0);
  bs.dir = V(0, 0, 0);
  d3 wi, t, b;
  if (rnd(r) < coat_prob(s, wo)) {
    basis(s->nc, &t, &b);
    d3 wol = to_local(wo, t, b, s->nc);
    if (wol.z <= 0) return bs;
    double u0 = rnd(r), u1 = rnd(r);
    wi = to_world(reflect3(neg(wol), vndf(wol, a_rough(s->cc_rough), u0, u1)), t, b, s->nc);
  } else {
    basis(s->n, &t, &b);
    d3 wol = to_local(wo, t, b, s->n), wil;
    if (wol.z <= 0) return bs;
    if (rnd(r) < spec_prob(base_f0(s), s->metallic, wol.z)) {
      double u0 = rnd(r), u1 = rnd(r);
      wil = reflect3(neg(wol), vndf(wol, a_rough(s->roughness), u0, u1));
    } else {
      double u0 = rnd(r), u1 = rnd(r);
      wil = cos_hemi(u0, u1);
    }
    wi = to_world(wil, t, b, s->n);
  }
  double ndl = dot(s->n, wi);
  if (ndl <= 0) return bs;
  double pdf;
  d3 f = eval_opaque(sc, s, wo, wi, &pdf);
  if (pdf <= 0) return bs;
  bs.dir = wi;
  bs.pdf = pdf;
  bs.w = scl(f, ndl / pdf);
  return bs;
}
static int sample_dielectric(const surf_t *s, int front, d3 ng, d3 wo, rng_t *r, d3 *wi, d3 *w, int *refl) {
  d3 n = s->n, t, b;
  basis(n, &t, &b);
  d3 wol = to_local(wo, t, b, n);
  if (wol.z <= 0) return 0;
  d3 H = n;
  if (s->roughness > 0) {
    double u0 = rnd(r), u1 = rnd(r);
    H = to_world(vndf(wol, a_rough(s->roughness), u0, u1), t, b, n);
  }
  double ci = dot(wo, H);
  if (ci <= 0) {
    H = n;
    ci = wol.z;
  }
  double eta = front ? 1 / s->ior : s->ior, s2t = eta * eta * (1 - ci * ci);
  double fr = schlick1(ci, diel_f0(s->ior));
  if (s->thin)
    fr = 2 * fr / (1 + fr);
  else if (s2t > 1)
    fr = 1;
  if (rnd(r) < fr) {
    *wi = reflect3(neg(wo), H);
    *w = V(1, 1, 1);
    *refl = 1;
    return dot(*wi, ng) > 0;
  }
  *wi = s->thin ? neg(wo) : refract3(neg(wo), H, eta);
  *w = s->albedo;
  *refl = 0;
  return dot(*wi, ng) < 0;
}

/* ---- trace.frag --------------------------------------------------------- */
static d3 clamp_c(d3 c, int bounce, double cl) {
  if (bounce == 0 || cl <= 0) return c;
  double l = lum(c);
  return l > cl ? scl(c, cl / l) : c;
}
static surf_t surface_at(const or_scene *sc, const hit_t *h, d3 wo, d3 *em) {
  const float *m = mat_ptr(sc, h->material);
  surf_t s;
  double tu = m[32] * h->u + m[33] * h->v + m[34], tv = m[36] * h->u + m[37] * h->v + m[38];
  s.albedo = Vf(m);
  if (m[7] > 0) {
    double qx = floor(h->pos.x * m[7] + 0.5), qy = floor(h->pos.y * m[7] + 0.5), qz = floor(h->pos.z * m[7] + 0.5);
    double sum = qx + qy + qz;
    if (sum - 2 * floor(sum / 2) >= 0.5) s.albedo = Vf(m + 12);
  }
  int tex = sc->tex_size > 0, mb = (int)m[16], mmr = (int)m[17], mn = (int)m[18], me = (int)m[19];
  double o[4];
  if (mb >= 0 && mb < sc->color_layers && tex) {
    sample_layer(sc->color_maps, sc->tex_size, mb, tu, tv, 1, o);
    s.albedo = mul(s.albedo, V(o[0], o[1], o[2]));
  }
  s.metallic = m[3];
  s.roughness = m[4];
  if (mmr >= 0 && mmr < sc->data_layers && tex) {
    sample_layer(sc->data_maps, sc->tex_size, mmr, tu, tv, 0, o);
    s.roughness *= o[1];
    s.metallic *= o[2];
  }
  *em = Vf(m + 8);
  if (me >= 0 && me < sc->color_layers && tex) {
    sample_layer(sc->color_maps, sc->tex_size, me, tu, tv, 1, o);
    *em = mul(*em, V(o[0], o[1], o[2]));
  }
  s.nc = h->n;
  s.n = h->n;
  if (mn >= 0 && mn < sc->data_layers && tex && dot(h->dpdu, h->dpdu) > 0) {
    d3 N = h->front ? h->n : neg(h->n);
    d3 T = nrm(sub(h->dpdu, scl(N, dot(N, h->dpdu))));
    d3 B = cross(N, T);
    if (dot(B, neg(h->dpdv)) < 0) B = neg(B);
    sample_layer(sc->data_maps, sc->tex_size, mn, tu, tv, 0, o);
    d3 tn = V((o[0] * 2 - 1) * m[20], (o[1] * 2 - 1) * m[20], o[2] * 2 - 1);
    d3 mapped = nrm(add(add(scl(T, tn.x), scl(B, tn.y)), scl(N, maxd(tn.z, 1e-3))));
    if (!h->front) mapped = neg(mapped);
    if (dot(mapped, wo) > 1e-3 && dot(mapped, h->ng) > 0) s.n = mapped;
  }
  s.ior = m[5];
  s.transmission = m[6] * (1 - s.metallic);
  s.emission = *em;
  s.clearcoat = m[21];
  s.cc_rough = m[22];
  s.sheen = Vf(m + 24);
  s.sheen_rough = m[27];
  s.thin = m[35] > 0.5f;
  if (m[31] > 0) {
    d3 c = V(maxd(m[28], 1e-4), maxd(m[29], 1e-4), maxd(m[30], 1e-4));
    s.absorption = scl(V(-log(c.x), -log(c.y), -log(c.z)), 1.0 / m[31]);
  } else {
    s.absorption = V(0, 0, 0);
  }
  return s;
}

typedef struct {
  d3 albedo, normal;
  double depth, coverage;
  int done;
  d3 position;  /* primary hit (trace.frag:158-159) */
  double reuse; /* trace.frag:187-193 */
} guide_t;

static double smoothstep_d(double e0, double e1, double x) {
  double t = clampd((x - e0) / (e1 - e0), 0, 1);
  return t * t * (3 - 2 * t);
}

static d3 trace_path(const or_scene *sc, const or_frame *f, d3 ro, d3 rd, rng_t *r, guide_t *g) {
  d3 rad = V(0, 0, 0), thr = V(1, 1, 1), tint = V(1, 1, 1);
  int spec = 1;
  double bpdf = 0;
  g->albedo = V(1, 1, 1);
  g->normal = neg(rd);
  g->depth = 1e4;
  g->coverage = 0;
  g->done = 0;
  g->position = V(0, 0, 0);
  g->reuse = 1;
  for (int bounce = 0; bounce <= (int)f->max_bounces; bounce++) {
    hit_t h;
    if (!trace_scene(sc, ro, rd, &h, r)) {
      COUNT(9, 1);
      d3 env = (bounce == 0 && sc->env_background == 0) ? Vf(sc->backdrop) : environment(sc, rd, spec, bpdf);
      rad = add(rad, clamp_c(mul(thr, env), bounce, f->clamp));
      if (!g->done) {
        g->albedo = tint;
        g->normal = neg(rd);
        g->done = 1;
      }
      break;
    }
    COUNT(7, 1);
    if (bounce == 0) {
      g->depth = h.t;
      g->coverage = 1;
      g->position = h.pos;
    }
    d3 wo = neg(rd), em;
    surf_t s = surface_at(sc, &h, wo, &em);
    d3 ng = h.ng;
    if (!h.front && !s.thin && s.transmission > 0) {
      d3 a = scl(s.absorption, -h.t);
      thr = mul(thr, V(exp(a.x), exp(a.y), exp(a.z)));
    }
    if (em.x > 0 || em.y > 0 || em.z > 0) {
      double w = 1;
      if (!spec && h.light >= 0) w = mis(bpdf, light_pdf(sc, h.light, ro, rd, h.t) / sc->light_count);
      rad = add(rad, clamp_c(scl(mul(thr, em), w), bounce, f->clamp));
      if (!g->done) {
        g->albedo = tint;
        g->normal = s.n;
        g->done = 1;
      }
      if (h.light >= 0) break;
    }
    int glass = s.transmission > 0 && rnd(r) < s.transmission;
    if (bounce == 0) {
      double sharp = 1 - smoothstep_d(0.1, 0.7, s.roughness);
      double glossy = maxd(sharp * (0.5 * (1 - s.metallic) + 1.0 * s.metallic), s.transmission);
      glossy = maxd(glossy, s.clearcoat * 0.5 * (1 - smoothstep_d(0.08, 0.5, s.cc_rough)));
      g->reuse = 1 - glossy;
    }
    int mirror = !glass && s.metallic > 0.9 && s.roughness < 0.1;
    if (!g->done && !((glass || mirror) && bounce < 3)) {
      g->albedo = mul(tint, s.albedo);
      g->normal = s.n;
      g->done = 1;
    }
    if (glass) {
      d3 wi, w;
      int refl;
      if (!sample_dielectric(&s, h.front, ng, wo, r, &wi, &w, &refl)) break;
      thr = mul(thr, w);
      if (!g->done) tint = mul(tint, w);
      ro = offset_ray(h.pos, refl ? ng : neg(ng));
      rd = wi;
      spec = 1;
    } else {
      if (!g->done) tint = mul(tint, s.albedo);
      d3 so = offset_ray(h.pos, ng), n = s.n;
      if (sc->sun_enabled) {
        double u0 = rnd(r), u1 = rnd(r);
        d3 t, b, sd = Vf(sc->sun_dir);
        basis(sd, &t, &b);
        d3 wi = to_world(uni_cone(u0, u1, sc->sun_cos), t, b, sd);
        double ndl = dot(n, wi);
        if (ndl > 0 && dot(ng, wi) > 0 && !occluded(sc, so, wi, INF_D, r)) {
          double pb;
          d3 fv = eval_opaque(sc, &s, wo, wi, &pb);
          double w = mis(sc->sun_pdf, pb);
          rad = add(rad, clamp_c(scl(mul(scl(mul(thr, fv), ndl), Vf(sc->sun_radiance)), w / sc->sun_pdf), bounce,
                                 f->clamp));
        }
      }
      if (sc->env_sampling) {
        COUNT(8, 1);
        lsample ls = sample_env(sc, r);
        double ndl = dot(n, ls.dir);
        if (ls.pdf > 0 && ndl > 0 && dot(ng, ls.dir) > 0 && !occluded(sc, so, ls.dir, INF_D, r)) {
          double pb;
          d3 fv = eval_opaque(sc, &s, wo, ls.dir, &pb);
          double w = mis(ls.pdf, pb);
          rad = add(rad, clamp_c(scl(mul(scl(mul(thr, fv), ndl), ls.rad), w / ls.pdf), bounce, f->clamp));
        }
      }
      if (sc->light_count > 0) {
        COUNT(10, 1);
        int li = (int)((float)rnd(r) * (float)sc->light_count);
        if (li > (int)sc->light_count - 1) li = (int)sc->light_count - 1;
        double u0 = rnd(r), u1 = rnd(r);
        lsample ls = sample_light(sc, li, h.pos, u0, u1);
        double ndl = dot(n, ls.dir);
        if (ls.pdf > 0 && ndl > 0 && dot(ng, ls.dir) > 0 && !occluded(sc, so, ls.dir, ls.dist * (1 - 1e-4), r)) {
          double pl = ls.pdf / sc->light_count, pb;
          d3 fv = eval_opaque(sc, &s, wo, ls.dir, &pb);
          double w = mis(pl, pb);
          rad = add(rad, clamp_c(scl(mul(scl(mul(thr, fv), ndl), ls.rad), w / pl), bounce, f->clamp));
        }
      }
      bsample bs = sample_opaque(sc, &s, wo, r);
      if (bs.pdf <= 0 || dot(bs.dir, ng) <= 0) break;
      thr = mul(thr, bs.w);
      bpdf = bs.pdf;
      spec = 0;
      ro = so;
      rd = bs.dir;
    }
    if (bounce >= 2) {
      double p = mind(0.95, max3(thr));
      if (rnd(r) >= p) break;
      thr = scl(thr, 1 / p);
    }
  }
  return rad;
}

/* main() of trace.frag up to tracePath (trace.frag:324-330): seed, camera
 * ray (cameraRay, trace.frag:53-62), path.  Returns the radiance and the
 * guide; *rd_out is the camera ray direction. */
static d3 sample_pixel(const or_scene *sc, const or_frame *f, uint32_t x, uint32_t y, uint32_t seed, guide_t *g,
                       d3 *rd_out) {
  rng_t r = {(x + y * f->width) * 1973u + seed * 9277u + 26699u};
  pcg(&r);
  double r0 = rnd(&r), r1 = rnd(&r);
  double ux = (x + 0.5 + r0) / f->width * 2 - 1, uy = (y + 0.5 + r1) / f->height * 2 - 1;
  double aspect = (double)f->width / f->height;
  d3 F = Vf(f->cam_fwd), R = Vf(f->cam_right), U = Vf(f->cam_up), P = Vf(f->cam_pos);
  d3 dir = nrm(add(add(F, scl(R, ux * aspect * f->tan_half_fov)), scl(U, uy * f->tan_half_fov)));
  d3 focus = add(P, scl(dir, f->focus_dist / dot(dir, F)));
  double l0 = rnd(&r), l1 = rnd(&r), lx, ly;
  concentric(l0, l1, &lx, &ly);
  d3 ro = add(add(P, scl(R, lx * f->aperture)), scl(U, ly * f->aperture));
  d3 rd = nrm(sub(focus, ro));
  *rd_out = rd;
  return trace_path(sc, f, ro, rd, &r, g);
}

/* Traces `samples` samples (seeds seed0, seed0+1, ...) at each listed pixel
 * (x, y with rows bottom-up) and writes per pixel: mean rgb, mean lum^2,
 * mean albedo rgb, mean coverage, mean normal xyz, mean depth (12 doubles),
 * plus the per-sample radiance (3 doubles per sample) when per_sample != NULL
 * and the fetch counters (16 doubles, accumulated) when counts16 != NULL. */
int or_trace_pixels(const or_scene *sc, const or_frame *f, uint32_t seed0, uint32_t samples, const uint32_t *xy,
                    uint32_t npix, double *out12, double *per_sample, double *counts16) {
  g_counts = counts16;
  if (counts16) counts16[0] += (double)npix * samples;
  for (uint32_t p = 0; p < npix; p++) {
    uint32_t x = xy[2 * p], y = xy[2 * p + 1];
    double acc[12] = {0};
    for (uint32_t k = 0; k < samples; k++) {
      guide_t g;
      d3 rd;
      d3 c = sample_pixel(sc, f, x, y, seed0 + k, &g, &rd);
      if (!isfinite(c.x) || !isfinite(c.y) || !isfinite(c.z)) c = V(0, 0, 0);
      double l = lum(c);
      double vals[12] = {c.x, c.y, c.z, l * l, g.albedo.x, g.albedo.y, g.albedo.z, g.coverage,
                         g.normal.x, g.normal.y, g.normal.z, g.depth};
      for (int i = 0; i < 12; i++) acc[i] += vals[i];
      if (per_sample) {
        double *ps = per_sample + ((size_t)p * samples + k) * 3;
        ps[0] = c.x;
        ps[1] = c.y;
        ps[2] = c.z;
      }
    }
    for (int i = 0; i < 12; i++) out12[(size_t)p * 12 + i] = acc[i] / samples;
  }
  g_counts = NULL;
  return 0;
}


/* ---- temporal reprojection (trace.frag:274-355 with uReproject = 1) ---- */
typedef struct {
  uint32_t width, height;
  float cam_pos[3], cam_right[3], cam_up[3], cam_fwd[3];
  float tan_half_fov, history_cap;
  const float *color4, *albedo4, *normal4, *history; /* previous running means, rows bottom-up */
} or_prev;

/* projectToPrevious (trace.frag:274-282). */
static int project_prev(const or_prev *pv, d3 d, double *px, double *py) {
  double z = dot(d, Vf(pv->cam_fwd));
  if (z <= 1e-6) return 0;
  double aspect = (double)pv->width / pv->height;
  double nx = dot(d, Vf(pv->cam_right)) / (z * pv->tan_half_fov * aspect);
  double ny = dot(d, Vf(pv->cam_up)) / (z * pv->tan_half_fov);
  *px = (nx * 0.5 + 0.5) * pv->width;
  *py = (ny * 0.5 + 0.5) * pv->height;
  return *px >= 0 && *py >= 0 && *px < pv->width && *py < pv->height;
}

/* reprojectHistory (trace.frag:284-322); prev13 = colour 4, albedo 4,
 * normal 4, history. */
static int reproject_history(const or_prev *pv, const guide_t *g, d3 rd, double prev13[13]) {
  for (int i = 0; i < 13; i++) prev13[i] = 0;
  int surface = g->coverage > 0.5;
  d3 d = surface ? sub(g->position, Vf(pv->cam_pos)) : rd;
  double px, py;
  if (!project_prev(pv, d, &px, &py)) return 0;
  double expected = len(d);
  double fx = px - 0.5, fy = py - 0.5;
  int bx = (int)floor(fx), by = (int)floor(fy);
  double tx = fx - floor(fx), ty = fy - floor(fy), wsum = 0;
  for (int i = 0; i < 4; i++) {
    int ox = i & 1, oy = i >> 1, qx = bx + ox, qy = by + oy;
    if (qx < 0 || qy < 0 || qx >= (int)pv->width || qy >= (int)pv->height) continue;
    size_t q = (size_t)qy * pv->width + qx;
    const float *a = pv->albedo4 + q * 4, *n = pv->normal4 + q * 4, *c = pv->color4 + q * 4;
    int same;
    if (surface) {
      d3 qn = nrm(V(n[0] + 1e-6, n[1] + 1e-6, n[2] + 1e-6));
      same = a[3] > 0.5 && fabs(n[3] - expected) < 0.02 * expected && dot(qn, g->normal) > 0.85;
    } else {
      same = a[3] < 0.5;
    }
    if (!same) continue;
    double w = (ox == 1 ? tx : 1 - tx) * (oy == 1 ? ty : 1 - ty);
    for (int k = 0; k < 4; k++) {
      prev13[k] += c[k] * w;
      prev13[4 + k] += a[k] * w;
      prev13[8 + k] += n[k] * w;
    }
    prev13[12] += pv->history[q] * w;
    wsum += w;
  }
  if (wsum < 1e-3) return 0;
  for (int k = 0; k < 12; k++) prev13[k] /= wsum;
  prev13[12] = mind(prev13[12] / wsum, pv->history_cap * g->reuse);
  return 1;
}

/* One reprojecting sample (seed `seed`) at each listed pixel of the frame:
 * out13 per pixel = colour rgb + lum^2, albedo rgb + coverage, normal xyz +
 * depth, history after the sample (FP64, no FP16 rounding); flags[p] = 1 if
 * the history was reprojected, 0 if the pixel started over. */
int or_trace_reprojected(const or_scene *sc, const or_frame *f, const or_prev *pv, uint32_t seed, const uint32_t *xy,
                         uint32_t npix, double *out13, uint8_t *flags) {
  for (uint32_t p = 0; p < npix; p++) {
    uint32_t x = xy[2 * p], y = xy[2 * p + 1];
    guide_t g;
    d3 rd;
    d3 c = sample_pixel(sc, f, x, y, seed, &g, &rd);
    if (!isfinite(c.x) || !isfinite(c.y) || !isfinite(c.z)) c = V(0, 0, 0);
    double l = lum(c), prev[13];
    int ok = reproject_history(pv, &g, rd, prev);
    if (!ok) prev[12] = 0;
    double k = 1.0 / (prev[12] + 1.0), j = 1 - k;
    double now[12] = {c.x, c.y, c.z, l * l, g.albedo.x, g.albedo.y, g.albedo.z, g.coverage,
                      g.normal.x, g.normal.y, g.normal.z, g.depth};
    double *o = out13 + (size_t)p * 13;
    for (int i = 0; i < 12; i++) o[i] = prev[i] * j + now[i] * k;
    o[12] = prev[12] + 1;
    if (flags) flags[p] = (uint8_t)ok;
  }
  return 0;
}

/* ---- host tables -------------------------------------------------------- */
/* alias.ts:17-64 — Vose with LIFO worklists.  Returns -1 on invalid input. */
int or_alias_table(const double *w, uint32_t n, float *prob, uint32_t *alias, double *pmf, double *total_out) {
  if (n == 0) return -1;
  double total = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (!(w[i] >= 0) || !isfinite(w[i])) return -1;
    total += w[i];
  }
  for (uint32_t i = 0; i < n; i++) pmf[i] = total > 0 ? w[i] / total : 1.0 / n;
  double *scaled = malloc(n * sizeof(double));
  uint32_t *small = malloc(n * sizeof(uint32_t)), *large = malloc(n * sizeof(uint32_t));
  uint32_t ns = 0, nl = 0;
  for (uint32_t i = 0; i < n; i++) {
    scaled[i] = pmf[i] * n;
    if (scaled[i] < 1)
      small[ns++] = i;
    else
      large[nl++] = i;
  }
  while (ns > 0 && nl > 0) {
    uint32_t s = small[--ns], l = large[--nl];
    prob[s] = (float)scaled[s];
    alias[s] = l;
    scaled[l] = scaled[l] + scaled[s] - 1;
    if (scaled[l] < 1)
      small[ns++] = l;
    else
      large[nl++] = l;
  }
  while (nl > 0) {
    uint32_t l = large[--nl];
    prob[l] = 1;
    alias[l] = l;
  }
  while (ns > 0) {
    uint32_t s = small[--ns];
    prob[s] = 1;
    alias[s] = s;
  }
  *total_out = total;
  free(scaled);
  free(small);
  free(large);
  return 0;
}

/* environment.ts:52-85 — returns grid size through gw/gh; table 4 floats/cell. */
int or_env_sampler(const float *rgb, uint32_t width, uint32_t height, uint32_t max_width, float *table,
                   uint32_t *gw_out, uint32_t *gh_out, double *total) {
  uint32_t gw = width < max_width ? width : max_width;
  if (gw < 1) gw = 1;
  uint32_t gh = (uint32_t)floor((double)height * gw / width + 0.5);
  if (gh > height) gh = height;
  if (gh < 1) gh = 1;
  size_t n = (size_t)gw * gh;
  double *w = malloc(n * sizeof(double));
  for (uint32_t cy = 0; cy < gh; cy++) {
    uint32_t y0 = (uint32_t)floor((double)cy * height / gh);
    uint32_t y1 = (uint32_t)floor((double)(cy + 1) * height / gh);
    if (y1 < y0 + 1) y1 = y0 + 1;
    double st = sin((PI * (cy + 0.5)) / gh);
    for (uint32_t cx = 0; cx < gw; cx++) {
      uint32_t x0 = (uint32_t)floor((double)cx * width / gw);
      uint32_t x1 = (uint32_t)floor((double)(cx + 1) * width / gw);
      if (x1 < x0 + 1) x1 = x0 + 1;
      double sum = 0;
      for (uint32_t y = y0; y < y1; y++)
        for (uint32_t x = x0; x < x1; x++) {
          const float *p = rgb + ((size_t)y * width + x) * 3;
          sum += 0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2];
        }
      w[(size_t)cy * gw + cx] = (sum / ((double)(y1 - y0) * (x1 - x0))) * st;
    }
  }
  float *prob = malloc(n * sizeof(float));
  uint32_t *alias = malloc(n * sizeof(uint32_t));
  double *pmf = malloc(n * sizeof(double));
  int rc = or_alias_table(w, (uint32_t)n, prob, alias, pmf, total);
  if (rc == 0)
    for (size_t i = 0; i < n; i++) {
      table[i * 4] = prob[i];
      table[i * 4 + 1] = (float)alias[i];
      table[i * 4 + 2] = (float)(pmf[i] * n);
      table[i * 4 + 3] = (float)(pmf[alias[i]] * n);
    }
  *gw_out = gw;
  *gh_out = gh;
  free(w);
  free(prob);
  free(alias);
  free(pmf);
  return rc;
}

/* sheen.ts:24-50 */
double or_sheen_albedo(double cos_v, double roughness, uint32_t steps) {
  double r = maxd(0.07, roughness), alpha = r * r;
  double v0 = sqrt(maxd(0, 1 - cos_v * cos_v)), v2 = cos_v;
  uint32_t phi_steps = steps / 2;
  double dth = PI / 2 / steps, dph = PI / phi_steps, e = 0;
  for (uint32_t i = 0; i < steps; i++) {
    double th = (i + 0.5) * dth, ch = cos(th), sh = sin(th), d = d_charlie(alpha, ch);
    for (uint32_t j = 0; j < phi_steps; j++) {
      double ph = (j + 0.5) * dph;
      double h0 = sh * cos(ph);
      double vh = v0 * h0 + v2 * ch;
      if (vh <= 0) continue;
      double cl = 2 * vh * ch - v2;
      if (cl <= 0) continue;
      e += d * v_neubelt(cl, cos_v) * cl * 4 * vh * sh * dth * dph;
    }
  }
  return mind(1, 2 * e);
}

/* Round to the nearest IEEE half (ties to even); the R16F upload. */
float or_round_half(float f) {
  if (!isfinite(f)) return f;
  double a = fabs((double)f);
  if (a >= 65520.0) return f > 0 ? INFINITY : -INFINITY;
  int e;
  frexp(a, &e);
  int eh = e - 1 < -14 ? -14 : e - 1;
  double q = ldexp(1.0, eh - 10);
  double r = nearbyint(a / q) * q;
  return (float)(f < 0 ? -r : r);
}

void or_sheen_lut_half(float *out) {
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) {
      double cv = maxd(1e-3, x / 15.0);
      out[y * 16 + x] = or_round_half((float)or_sheen_albedo(cv, y / 15.0, 96));
    }
}

/* BSDF evaluation for closed-form tests: surface from a material's 40 floats,
 * shading normal (0,0,1). */
void or_eval_opaque(const float *mat40, const float *sheen_lut, const double *wo, const double *wi, double *f3,
                    double *pdf) {
  or_scene sc;
  memset(&sc, 0, sizeof(sc));
  sc.sheen_lut = sheen_lut;
  surf_t s;
  s.albedo = Vf(mat40);
  s.metallic = mat40[3];
  s.roughness = mat40[4];
  s.ior = mat40[5];
  s.transmission = mat40[6];
  s.clearcoat = mat40[21];
  s.cc_rough = mat40[22];
  s.sheen = Vf(mat40 + 24);
  s.sheen_rough = mat40[27];
  s.thin = mat40[35] > 0.5f;
  s.n = V(0, 0, 1);
  s.nc = s.n;
  d3 r = eval_opaque(&sc, &s, V(wo[0], wo[1], wo[2]), V(wi[0], wi[1], wi[2]), pdf);
  f3[0] = r.x;
  f3[1] = r.y;
  f3[2] = r.z;
}

/* One opaque BSDF sample (weight f*cos/pdf), drawing from a PCG stream. */
/* or_eval_opaque over n direction pairs (local frame, normal +z), for
 * quadratures of the BSDF. */
void or_eval_opaque_batch(const float *mat40, const float *sheen_lut, uint32_t n, const double *wo,
                          const double *wi, double *f3, double *pdf) {
  for (uint32_t i = 0; i < n; i++) or_eval_opaque(mat40, sheen_lut, wo + 3 * i, wi + 3 * i, f3 + 3 * i, pdf + i);
}

void or_sample_opaque(const float *mat40, const float *sheen_lut, const double *wo, uint32_t rng_state,
                      double *dir, double *w3, double *pdf) {
  or_scene sc;
  memset(&sc, 0, sizeof(sc));
  sc.sheen_lut = sheen_lut;
  surf_t s;
  s.albedo = Vf(mat40);
  s.metallic = mat40[3];
  s.roughness = mat40[4];
  s.ior = mat40[5];
  s.transmission = mat40[6];
  s.clearcoat = mat40[21];
  s.cc_rough = mat40[22];
  s.sheen = Vf(mat40 + 24);
  s.sheen_rough = mat40[27];
  s.thin = mat40[35] > 0.5f;
  s.n = V(0, 0, 1);
  s.nc = s.n;
  rng_t r = {rng_state};
  bsample bs = sample_opaque(&sc, &s, V(wo[0], wo[1], wo[2]), &r);
  dir[0] = bs.dir.x;
  dir[1] = bs.dir.y;
  dir[2] = bs.dir.z;
  w3[0] = bs.w.x;
  w3[1] = bs.w.y;
  w3[2] = bs.w.z;
  *pdf = bs.pdf;
}

/* One dielectric sample: returns 1 if the path continues, reflected flag. */
int or_sample_dielectric(const float *mat40, int front, const double *wo, uint32_t rng_state, double *wi, int *refl) {
  surf_t s;
  memset(&s, 0, sizeof(s));
  s.albedo = Vf(mat40);
  s.roughness = mat40[4];
  s.ior = mat40[5];
  s.thin = mat40[35] > 0.5f;
  s.n = V(0, 0, 1);
  s.nc = s.n;
  rng_t r = {rng_state};
  d3 w, o;
  int ok = sample_dielectric(&s, front, V(0, 0, 1), V(wo[0], wo[1], wo[2]), &r, &o, &w, refl);
  wi[0] = o.x;
  wi[1] = o.y;
  wi[2] = o.z;
  return ok;
}

/* ---- FP64 path-replay adjoint ------------------------------------------ */
/* Derivatives of the opaque BSDF value f with respect to (albedo r, g, b,
 * roughness, metallic) by central differences in double precision: an
 * evaluation independent of the product's forward-mode dual numbers. */
static int eval_grad_fd(const or_scene *sc, const surf_t *s, d3 wo, d3 wi, double f[3], double df[3][5]) {
  double pdf;
  d3 f0 = eval_opaque(sc, s, wo, wi, &pdf);
  if (pdf <= 0 && f0.x == 0 && f0.y == 0 && f0.z == 0) {
    if (dot(s->n, wi) <= 0 || dot(s->n, wo) <= 0) return 0;
  }
  f[0] = f0.x;
  f[1] = f0.y;
  f[2] = f0.z;
  for (int i = 0; i < 5; i++) {
    surf_t p = *s, m = *s;
    double *vp, *vm, base;
    switch (i) {
      case 0: vp = &p.albedo.x; vm = &m.albedo.x; base = s->albedo.x; break;
      case 1: vp = &p.albedo.y; vm = &m.albedo.y; base = s->albedo.y; break;
      case 2: vp = &p.albedo.z; vm = &m.albedo.z; base = s->albedo.z; break;
      case 3: vp = &p.roughness; vm = &m.roughness; base = s->roughness; break;
      default: vp = &p.metallic; vm = &m.metallic; base = s->metallic; break;
    }
    double h = 1e-6;
    *vp = base + h;
    *vm = base - h;
    double pp, pm;
    d3 fp = eval_opaque(sc, &p, wo, wi, &pp), fm = eval_opaque(sc, &m, wo, wi, &pm);
    df[0][i] = (fp.x - fm.x) / (2 * h);
    df[1][i] = (fp.y - fm.y) / (2 * h);
    df[2][i] = (fp.z - fm.z) / (2 * h);
  }
  return 1;
}

typedef struct {
  double *mat;  /* material_count * 5 */
  double *env;  /* env_w * env_h * 3 or NULL */
} grads_t;

static void env_taps_d(const or_scene *s, double u, double v, int idx[4], double w[4]) {
  double x = u * s->env_w - 0.5, y = v * s->env_h - 0.5, fx = floor(x), fy = floor(y), tx = x - fx, ty = y - fy;
  int x0 = ((int)fx % s->env_w + s->env_w) % s->env_w, x1 = (x0 + 1) % s->env_w;
  int y0 = (int)fy, y1 = (int)fy + 1;
  y0 = y0 < 0 ? 0 : (y0 >= s->env_h ? s->env_h - 1 : y0);
  y1 = y1 < 0 ? 0 : (y1 >= s->env_h ? s->env_h - 1 : y1);
  idx[0] = y0 * s->env_w + x0;
  idx[1] = y0 * s->env_w + x1;
  idx[2] = y1 * s->env_w + x0;
  idx[3] = y1 * s->env_w + x1;
  w[0] = (1 - tx) * (1 - ty);
  w[1] = tx * (1 - ty);
  w[2] = (1 - tx) * ty;
  w[3] = tx * ty;
}
static void add_env_grad(const or_scene *s, grads_t *g, double u, double v, d3 wgt) {
  if (!g->env) return;
  int idx[4];
  double w[4];
  env_taps_d(s, u, v, idx, w);
  for (int k = 0; k < 4; k++) {
    double sc = w[k] * s->env_intensity;
    g->env[(size_t)idx[k] * 3] += wgt.x * sc;
    g->env[(size_t)idx[k] * 3 + 1] += wgt.y * sc;
    g->env[(size_t)idx[k] * 3 + 2] += wgt.z * sc;
  }
}
static void add_f_grad(grads_t *g, int mat, d3 coeff, double df[3][5]) {
  double c[3] = {coeff.x, coeff.y, coeff.z};
  for (int i = 0; i < 5; i++) g->mat[mat * 5 + i] += c[0] * df[0][i] + c[1] * df[1][i] + c[2] * df[2][i];
}

/* Replay of trace_path accumulating d(dl · L)/dθ (material parameters through
 * f, environment texels through every lookup); sampling decisions, pdfs, MIS
 * weights and Russian roulette held constant. */
static void replay_path(const or_scene *sc, const or_frame *f, d3 ro, d3 rd, rng_t *r, d3 L, d3 dl, grads_t *g) {
  d3 lr = L, beta = V(1, 1, 1);
  int spec = 1;
  double bpdf = 0;
  for (int bounce = 0; bounce <= (int)f->max_bounces; bounce++) {
    hit_t h;
    if (!trace_scene(sc, ro, rd, &h, r)) {
      if (!(bounce == 0 && sc->env_background == 0)) {
        double w = (!spec && sc->env_sampling) ? mis(bpdf, env_pdf(sc, rd)) : 1;
        double u, v;
        eq_uv(env_local(sc, rd), &u, &v);
        add_env_grad(sc, g, u, v, scl(mul(dl, beta), w));
      }
      return;
    }
    d3 wo = neg(rd), em;
    surf_t s = surface_at(sc, &h, wo, &em);
    d3 ng = h.ng;
    if (!h.front && !s.thin && s.transmission > 0) {
      d3 a = scl(s.absorption, -h.t);
      beta = mul(beta, V(exp(a.x), exp(a.y), exp(a.z)));
    }
    if (em.x > 0 || em.y > 0 || em.z > 0) {
      double w = 1;
      if (!spec && h.light >= 0) w = mis(bpdf, light_pdf(sc, h.light, ro, rd, h.t) / sc->light_count);
      lr = sub(lr, scl(mul(beta, em), w));
      if (h.light >= 0) return;
    }
    int glass = s.transmission > 0 && rnd(r) < s.transmission;
    if (glass) {
      d3 wi, w;
      int refl;
      if (!sample_dielectric(&s, h.front, ng, wo, r, &wi, &w, &refl)) return;
      if (!refl) {
        if (s.albedo.x > 0) g->mat[h.material * 5 + 0] += dl.x * lr.x / s.albedo.x;
        if (s.albedo.y > 0) g->mat[h.material * 5 + 1] += dl.y * lr.y / s.albedo.y;
        if (s.albedo.z > 0) g->mat[h.material * 5 + 2] += dl.z * lr.z / s.albedo.z;
      }
      beta = mul(beta, w);
      ro = offset_ray(h.pos, refl ? ng : neg(ng));
      rd = wi;
      spec = 1;
    } else {
      d3 so = offset_ray(h.pos, ng), n = s.n;
      double fv[3], dfv[3][5];
      if (sc->sun_enabled) {
        double u0 = rnd(r), u1 = rnd(r);
        d3 t, b, sd = Vf(sc->sun_dir);
        basis(sd, &t, &b);
        d3 wi = to_world(uni_cone(u0, u1, sc->sun_cos), t, b, sd);
        double ndl = dot(n, wi);
        if (ndl > 0 && dot(ng, wi) > 0 && !occluded(sc, so, wi, INF_D, r)) {
          double pb;
          eval_opaque(sc, &s, wo, wi, &pb);
          double k = ndl * mis(sc->sun_pdf, pb) / sc->sun_pdf;
          if (eval_grad_fd(sc, &s, wo, wi, fv, dfv)) {
            d3 li = scl(Vf(sc->sun_radiance), k);
            lr = sub(lr, mul(mul(beta, V(fv[0], fv[1], fv[2])), li));
            add_f_grad(g, h.material, mul(mul(dl, beta), li), dfv);
          }
        }
      }
      if (sc->env_sampling) {
        lsample ls = sample_env(sc, r);
        double ndl = dot(n, ls.dir);
        if (ls.pdf > 0 && ndl > 0 && dot(ng, ls.dir) > 0 && !occluded(sc, so, ls.dir, INF_D, r)) {
          double pb;
          eval_opaque(sc, &s, wo, ls.dir, &pb);
          double k = ndl * mis(ls.pdf, pb) / ls.pdf;
          if (eval_grad_fd(sc, &s, wo, ls.dir, fv, dfv)) {
            d3 fvec = V(fv[0], fv[1], fv[2]);
            lr = sub(lr, scl(mul(mul(beta, fvec), ls.rad), k));
            add_f_grad(g, h.material, scl(mul(mul(dl, beta), ls.rad), k), dfv);
            /* The sampled map coordinates, recovered from the direction. */
            double u, v;
            eq_uv(env_local(sc, ls.dir), &u, &v);
            add_env_grad(sc, g, u, v, scl(mul(mul(dl, beta), fvec), k));
          }
        }
      }
      if (sc->light_count > 0) {
        int li = (int)((float)rnd(r) * (float)sc->light_count);
        if (li > (int)sc->light_count - 1) li = (int)sc->light_count - 1;
        double u0 = rnd(r), u1 = rnd(r);
        lsample ls = sample_light(sc, li, h.pos, u0, u1);
        double ndl = dot(n, ls.dir);
        if (ls.pdf > 0 && ndl > 0 && dot(ng, ls.dir) > 0 && !occluded(sc, so, ls.dir, ls.dist * (1 - 1e-4), r)) {
          double pl = ls.pdf / sc->light_count, pb;
          eval_opaque(sc, &s, wo, ls.dir, &pb);
          double k = ndl * mis(pl, pb) / pl;
          if (eval_grad_fd(sc, &s, wo, ls.dir, fv, dfv)) {
            lr = sub(lr, scl(mul(mul(beta, V(fv[0], fv[1], fv[2])), ls.rad), k));
            add_f_grad(g, h.material, scl(mul(mul(dl, beta), ls.rad), k), dfv);
          }
        }
      }
      bsample bs = sample_opaque(sc, &s, wo, r);
      if (bs.pdf <= 0 || dot(bs.dir, ng) <= 0) return;
      if (eval_grad_fd(sc, &s, wo, bs.dir, fv, dfv)) {
        d3 coeff = V(fv[0] > 1e-300 ? dl.x * lr.x / fv[0] : 0, fv[1] > 1e-300 ? dl.y * lr.y / fv[1] : 0,
                     fv[2] > 1e-300 ? dl.z * lr.z / fv[2] : 0);
        add_f_grad(g, h.material, coeff, dfv);
      }
      beta = mul(beta, bs.w);
      bpdf = bs.pdf;
      spec = 0;
      ro = so;
      rd = bs.dir;
    }
    if (bounce >= 2) {
      double p = mind(0.95, max3(beta));
      if (rnd(r) >= p) return;
      beta = scl(beta, 1 / p);
    }
  }
}

/* Gradient of sum_p dl[p] · mean_radiance[p] over the listed pixels (rows
 * bottom-up), `samples` samples from seed0; accumulates into grad_mat
 * (material_count * 5) and grad_env (env texels * 3, may be NULL). */
int or_adjoint_pixels(const or_scene *sc, const or_frame *f, uint32_t seed0, uint32_t samples, const uint32_t *xy,
                      const double *dl3, uint32_t npix, double *grad_mat, double *grad_env) {
  grads_t g = {grad_mat, grad_env};
  for (uint32_t p = 0; p < npix; p++) {
    uint32_t x = xy[2 * p], y = xy[2 * p + 1];
    d3 dl = scl(V(dl3[p * 3], dl3[p * 3 + 1], dl3[p * 3 + 2]), 1.0 / samples);
    for (uint32_t k = 0; k < samples; k++) {
      rng_t r = {(x + y * f->width) * 1973u + (seed0 + k) * 9277u + 26699u};
      pcg(&r);
      double r0 = rnd(&r), r1 = rnd(&r);
      double ux = (x + 0.5 + r0) / f->width * 2 - 1, uy = (y + 0.5 + r1) / f->height * 2 - 1;
      double aspect = (double)f->width / f->height;
      d3 F = Vf(f->cam_fwd), R = Vf(f->cam_right), U = Vf(f->cam_up), P = Vf(f->cam_pos);
      d3 dir = nrm(add(add(F, scl(R, ux * aspect * f->tan_half_fov)), scl(U, uy * f->tan_half_fov)));
      d3 focus = add(P, scl(dir, f->focus_dist / dot(dir, F)));
      double l0 = rnd(&r), l1 = rnd(&r), lx, ly;
      concentric(l0, l1, &lx, &ly);
      d3 ro = add(add(P, scl(R, lx * f->aperture)), scl(U, ly * f->aperture));
      d3 rd = nrm(sub(focus, ro));
      rng_t replay_rng = r;
      guide_t gd;
      d3 L = trace_path(sc, f, ro, rd, &r, &gd);
      if (!isfinite(L.x) || !isfinite(L.y) || !isfinite(L.z)) continue;
      replay_path(sc, f, ro, rd, &replay_rng, L, dl, &g);
    }
  }
  return 0;
}
