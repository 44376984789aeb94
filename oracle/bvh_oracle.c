/*
 * TEST INFRASTRUCTURE — the parity oracle, not the product.
 *
 * A plain-C, FP64 restatement of the CPU geometry code on Lucent's
 * path-tracing hot path: the binned-SAH BVH builder, the Möller–Trumbore and
 * slab tests, the CPU closest-hit traversal, the brute-force reference, the
 * reference shader's own traversal order (to count the node and triangle
 * fetches that define the algorithmic bytes of a ray), the test-suite LCG and
 * its random triangle soup, and the torus-knot generator.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library.  It is compiled with -ffp-contract=off so that every
 * double operation rounds exactly as a JavaScript number operation does.
 *
 * Parity pins (what fixes each function besides this text):
 *   or_intersect_triangle / or_aabb_intersect : the reference's known-answer
 *       tests (test/ray.test.ts:9-37, test/aabb.test.ts:21-33) and closed forms.
 *   or_intersect_bvh : brute force (or_intersect_all) on the reference's own
 *       regenerated vectors (test/bvh.test.ts:90-132) plus random cases.
 *   or_build_bvh : invariants (test/bvh.test.ts:31-86) and a bit-exact
 *       cross-check against the product's independent C++ builder.  Its exact
 *       node arrays are otherwise "parity unpinned" (the TypeScript cannot run
 *       in this image).
 *   or_torus_knot : vertex counts, closedness and unit normals
 *       (test/mesh.test.ts:47-55).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------------------
 * JavaScript Math.min / Math.max on numbers: NaN wins, -0 < +0.
 * ------------------------------------------------------------------------- */
static double js_min2(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return signbit(a) ? a : b;
  return a < b ? a : b;
}
static double js_max2(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}

/* test/bvh.test.ts:7-13 — s = (s*1664525 + 1013904223) >>> 0; return s / 2^32.
 * The product s*1664525 stays below 2^53, so JS doubles equal uint32 wrap. */
double or_lcg_next(uint32_t *state) {
  *state = *state * 1664525u + 1013904223u;
  return (double)*state / 4294967296.0;
}

/* test/bvh.test.ts:15-28 — count triangles scattered in [-5,5]^3, each vertex
 * within ±0.5 of its centre; values rounded to float32 on store. */
void or_random_soup(uint32_t count, uint32_t *state, float *out) {
  for (uint32_t i = 0; i < count; i++) {
    double cx = or_lcg_next(state) * 10 - 5;
    double cy = or_lcg_next(state) * 10 - 5;
    double cz = or_lcg_next(state) * 10 - 5;
    for (int k = 0; k < 3; k++) {
      double x = cx + or_lcg_next(state) - 0.5;
      out[i * 9 + k * 3] = (float)x;
      double y = cy + or_lcg_next(state) - 0.5;
      out[i * 9 + k * 3 + 1] = (float)y;
      double z = cz + or_lcg_next(state) - 0.5;
      out[i * 9 + k * 3 + 2] = (float)z;
    }
  }
}

/* ---------------------------------------------------------------------------
 * vec3.ts:15-33 — left-to-right, no fused multiply-add.
 * ------------------------------------------------------------------------- */
static void v_sub(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void v_cross(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static void v_normalize(const double *a, double *o) {
  double len = sqrt(v_dot(a, a));
  if (len == 0) {
    o[0] = o[1] = o[2] = 0;
    return;
  }
  double s = 1 / len;
  o[0] = a[0] * s;
  o[1] = a[1] * s;
  o[2] = a[2] * s;
}

/* ray.ts:22-45 — two-sided Möller–Trumbore in FP64. */
int or_intersect_triangle(const double *ro, const double *rd, const double *v0, const double *v1,
                          const double *v2, double tmin, double tmax, double *t_out, double *u_out,
                          double *v_out) {
  double e1[3], e2[3], p[3], s[3], q[3];
  v_sub(v1, v0, e1);
  v_sub(v2, v0, e2);
  v_cross(rd, e2, p);
  double det = v_dot(e1, p);
  if (fabs(det) < 1e-12) return 0;
  double inv_det = 1 / det;
  v_sub(ro, v0, s);
  double u = v_dot(s, p) * inv_det;
  if (u < 0 || u > 1) return 0;
  v_cross(s, e1, q);
  double v = v_dot(rd, q) * inv_det;
  if (v < 0 || u + v > 1) return 0;
  double t = v_dot(e2, q) * inv_det;
  if (t <= tmin || t >= tmax) return 0;
  *t_out = t;
  *u_out = u;
  *v_out = v;
  return 1;
}

/* ray.ts:48-67 — nearest root in (tmin, tmax), or NAN for none. */
double or_intersect_sphere(const double *ro, const double *rd, const double *c, double radius,
                           double tmin, double tmax) {
  double oc[3];
  v_sub(ro, c, oc);
  double b = v_dot(oc, rd);
  double cc = v_dot(oc, oc) - radius * radius;
  double disc = b * b - cc;
  if (disc < 0) return NAN;
  double sq = sqrt(disc);
  double t = -b - sq;
  if (t <= tmin || t >= tmax) {
    t = -b + sq;
    if (t <= tmin || t >= tmax) return NAN;
  }
  return t;
}

/* aabb.ts:84-98 — FP64 slab test from tmin = 0; INFINITY on a miss. */
double or_aabb_intersect(const double *bmin, const double *bmax, const double *o, const double *inv,
                         double tmax_in) {
  double tmin = 0, tmax = tmax_in;
  for (int i = 0; i < 3; i++) {
    double t0 = (bmin[i] - o[i]) * inv[i];
    double t1 = (bmax[i] - o[i]) * inv[i];
    if (t0 > t1) {
      double tt = t0;
      t0 = t1;
      t1 = tt;
    }
    if (t0 > tmin) tmin = t0;
    if (t1 < tmax) tmax = t1;
    if (tmax < tmin) return INFINITY;
  }
  return tmin;
}

/* ---------------------------------------------------------------------------
 * bvh.ts:35-40 boxArea.
 * ------------------------------------------------------------------------- */
static double box_area(double x0, double y0, double z0, double x1, double y1, double z1) {
  double dx = x1 - x0, dy = y1 - y0, dz = z1 - z0;
  return dx < 0 ? 0 : 2 * (dx * dy + dy * dz + dz * dx);
}

typedef struct {
  const float *tb, *cen;
  uint32_t *tri;
  uint32_t bins;
  double box[6], cbox[6], bin_scale[3];
} build_state;

static void measure(build_state *s, uint32_t start, uint32_t end) {
  for (int a = 0; a < 3; a++) {
    s->box[a] = INFINITY;
    s->box[a + 3] = -INFINITY;
    s->cbox[a] = INFINITY;
    s->cbox[a + 3] = -INFINITY;
  }
  for (uint32_t i = start; i < end; i++) {
    uint32_t t = s->tri[i];
    for (int a = 0; a < 3; a++) {
      double lo = s->tb[t * 6 + a], hi = s->tb[t * 6 + 3 + a], c = s->cen[t * 3 + a];
      if (lo < s->box[a]) s->box[a] = lo;
      if (hi > s->box[a + 3]) s->box[a + 3] = hi;
      if (c < s->cbox[a]) s->cbox[a] = c;
      if (c > s->cbox[a + 3]) s->cbox[a + 3] = c;
    }
  }
}

static double bin_of(const build_state *s, uint32_t t, int axis) {
  double f = floor(((double)s->cen[t * 3 + axis] - s->cbox[axis]) * s->bin_scale[axis]);
  return js_min2((double)(s->bins - 1), f);
}

static void write_node(float *nodes, uint32_t idx, const double *b, double left_first, double count) {
  float *o = nodes + (size_t)idx * 8;
  o[0] = (float)b[0];
  o[1] = (float)b[1];
  o[2] = (float)b[2];
  o[3] = (float)left_first;
  o[4] = (float)b[3];
  o[5] = (float)b[4];
  o[6] = (float)b[5];
  o[7] = (float)count;
}

/* bvh.ts:52-252 — binned SAH build.  Returns 0, or -1 on a bad argument. */
int or_build_bvh(const float *pos, uint32_t tri_count, uint32_t max_leaf, uint32_t bins, float *nodes,
                 uint32_t *node_count_out, uint32_t *tri_idx, double *bounds6) {
  if (bins < 2 || bins > 256) return -1;
  if (tri_count == 0) {
    double zero[6] = {0, 0, 0, 0, 0, 0};
    write_node(nodes, 0, zero, 0, 0);
    *node_count_out = 1;
    memset(bounds6, 0, 6 * sizeof(double));
    return 0;
  }
  float *tb = (float *)malloc((size_t)tri_count * 6 * sizeof(float));
  float *cen = (float *)malloc((size_t)tri_count * 3 * sizeof(float));
  for (uint32_t i = 0; i < tri_count; i++) {
    for (int a = 0; a < 3; a++) {
      double p0 = pos[(size_t)i * 9 + a], p1 = pos[(size_t)i * 9 + 3 + a], p2 = pos[(size_t)i * 9 + 6 + a];
      double lo = js_min2(js_min2(p0, p1), p2);
      double hi = js_max2(js_max2(p0, p1), p2);
      tb[(size_t)i * 6 + a] = (float)lo;
      tb[(size_t)i * 6 + 3 + a] = (float)hi;
      cen[(size_t)i * 3 + a] = (float)((lo + hi) * 0.5);
    }
  }
  for (uint32_t i = 0; i < tri_count; i++) tri_idx[i] = i;

  build_state s;
  s.tb = tb;
  s.cen = cen;
  s.tri = tri_idx;
  s.bins = bins;
  double *bin_box = (double *)malloc(3 * bins * 6 * sizeof(double));
  uint32_t *bin_count = (uint32_t *)calloc(3 * bins, sizeof(uint32_t));
  double *right_area = (double *)calloc(bins, sizeof(double));
  uint32_t *right_count = (uint32_t *)calloc(bins, sizeof(uint32_t));
  int splittable[3];

  /* Work stack of (node, start, end) triples; the left child is popped first. */
  size_t cap = 64, sp = 0;
  uint32_t *stack = (uint32_t *)malloc(cap * 3 * sizeof(uint32_t));
  uint32_t node_count = 1;
  stack[0] = 0;
  stack[1] = 0;
  stack[2] = tri_count;
  sp = 1;
  while (sp > 0) {
    sp--;
    uint32_t node_idx = stack[sp * 3], start = stack[sp * 3 + 1], end = stack[sp * 3 + 2];
    uint32_t count = end - start;
    measure(&s, start, end);
    if (count <= max_leaf) {
      write_node(nodes, node_idx, s.box, start, count);
      continue;
    }
    double *cb = s.cbox;
    int longest = (cb[3] - cb[0] >= cb[4] - cb[1] && cb[3] - cb[0] >= cb[5] - cb[2]) ? 0
                  : (cb[4] - cb[1] >= cb[5] - cb[2])                                 ? 1
                                                                                      : 2;
    for (int a = 0; a < 3; a++) {
      double extent = cb[a + 3] - cb[a];
      splittable[a] = extent > 1e-9 && (count >= 512 || a == longest);
      s.bin_scale[a] = (double)bins / extent;
      if (!splittable[a]) continue;
      for (uint32_t k = a * bins; k < (a + 1) * bins; k++) {
        bin_box[k * 6] = bin_box[k * 6 + 1] = bin_box[k * 6 + 2] = INFINITY;
        bin_box[k * 6 + 3] = bin_box[k * 6 + 4] = bin_box[k * 6 + 5] = -INFINITY;
        bin_count[k] = 0;
      }
    }
    for (uint32_t i = start; i < end; i++) {
      uint32_t t = tri_idx[i];
      for (int a = 0; a < 3; a++) {
        if (!splittable[a]) continue;
        uint32_t k = a * bins + (uint32_t)bin_of(&s, t, a);
        bin_count[k]++;
        for (int c = 0; c < 3; c++) {
          double lo = tb[(size_t)t * 6 + c], hi = tb[(size_t)t * 6 + 3 + c];
          if (lo < bin_box[k * 6 + c]) bin_box[k * 6 + c] = lo;
          if (hi > bin_box[k * 6 + 3 + c]) bin_box[k * 6 + 3 + c] = hi;
        }
      }
    }
    double best_cost = INFINITY;
    int best_axis = -1, best_bin = -1;
    for (int a = 0; a < 3; a++) {
      if (!splittable[a]) continue;
      double x0 = INFINITY, y0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, z1 = -INFINITY;
      uint32_t n = 0;
      for (uint32_t b = bins - 1; b > 0; b--) {
        const double *o = bin_box + (a * bins + b) * 6;
        x0 = js_min2(x0, o[0]);
        y0 = js_min2(y0, o[1]);
        z0 = js_min2(z0, o[2]);
        x1 = js_max2(x1, o[3]);
        y1 = js_max2(y1, o[4]);
        z1 = js_max2(z1, o[5]);
        n += bin_count[a * bins + b];
        right_area[b] = box_area(x0, y0, z0, x1, y1, z1);
        right_count[b] = n;
      }
      x0 = INFINITY, y0 = INFINITY, z0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, z1 = -INFINITY;
      uint32_t left_n = 0;
      for (uint32_t b = 0; b < bins - 1; b++) {
        const double *o = bin_box + (a * bins + b) * 6;
        x0 = js_min2(x0, o[0]);
        y0 = js_min2(y0, o[1]);
        z0 = js_min2(z0, o[2]);
        x1 = js_max2(x1, o[3]);
        y1 = js_max2(y1, o[4]);
        z1 = js_max2(z1, o[5]);
        left_n += bin_count[a * bins + b];
        uint32_t rn = right_count[b + 1];
        if (left_n == 0 || rn == 0) continue;
        double cost = box_area(x0, y0, z0, x1, y1, z1) * left_n + right_area[b + 1] * rn;
        if (cost < best_cost) {
          best_cost = cost;
          best_axis = a;
          best_bin = (int)b;
        }
      }
    }
    double leaf_cost = count * box_area(s.box[0], s.box[1], s.box[2], s.box[3], s.box[4], s.box[5]);
    int unprofitable = best_axis < 0 || best_cost >= leaf_cost;
    if (unprofitable && count <= max_leaf * 4) {
      write_node(nodes, node_idx, s.box, start, count);
      continue;
    }
    uint32_t mid;
    if (unprofitable) {
      mid = (start + end) >> 1;
    } else {
      /* bvh.ts:126-140 — swap-from-the-end partition. */
      int64_t i = start, j = (int64_t)end - 1;
      while (i <= j) {
        if (bin_of(&s, tri_idx[i], best_axis) <= best_bin) {
          i++;
        } else {
          uint32_t tmp = tri_idx[i];
          tri_idx[i] = tri_idx[j];
          tri_idx[j] = tmp;
          j--;
        }
      }
      mid = (uint32_t)i;
    }
    if (mid == start || mid == end) mid = (start + end) >> 1;
    uint32_t left = node_count++;
    uint32_t right = node_count++;
    write_node(nodes, node_idx, s.box, left, 0);
    if (sp + 2 > cap) {
      cap *= 2;
      stack = (uint32_t *)realloc(stack, cap * 3 * sizeof(uint32_t));
    }
    stack[sp * 3] = right;
    stack[sp * 3 + 1] = mid;
    stack[sp * 3 + 2] = end;
    sp++;
    stack[sp * 3] = left;
    stack[sp * 3 + 1] = start;
    stack[sp * 3 + 2] = mid;
    sp++;
  }
  measure(&s, 0, tri_count);
  memcpy(bounds6, s.box, 6 * sizeof(double));
  *node_count_out = node_count;
  free(stack);
  free(right_count);
  free(right_area);
  free(bin_count);
  free(bin_box);
  free(cen);
  free(tb);
  return 0;
}

static void load_tri(const float *pos, uint32_t tri, double *a, double *b, double *c) {
  const float *p = pos + (size_t)tri * 9;
  for (int k = 0; k < 3; k++) {
    a[k] = p[k];
    b[k] = p[3 + k];
    c[k] = p[6 + k];
  }
}

/* bvh.ts:265-320 — closest hit, left child first, 64-entry stack.  Returns 1 on
 * a hit, 0 on a miss, -1 if the stack would overflow (the reference's typed
 * array drops such writes; no tree used here is that deep). */
int or_intersect_bvh(const float *nodes, const uint32_t *tri_idx, uint32_t tri_count, const float *pos,
                     const double *ro, const double *rd, double tmin, double tmax, or_hit *hit) {
  if (tri_count == 0) return 0;
  double inv[3] = {1 / rd[0], 1 / rd[1], 1 / rd[2]};
  int32_t stack[64];
  int sp = 0;
  stack[sp++] = 0;
  int found = 0;
  double closest = tmax;
  while (sp > 0) {
    int32_t n = stack[--sp];
    const float *o = nodes + (size_t)n * 8;
    double bmin[3] = {o[0], o[1], o[2]}, bmax[3] = {o[4], o[5], o[6]};
    if (or_aabb_intersect(bmin, bmax, ro, inv, closest) == INFINITY) continue;
    uint32_t left_first = (uint32_t)o[3];
    uint32_t count = (uint32_t)o[7];
    if (count > 0) {
      for (uint32_t i = 0; i < count; i++) {
        uint32_t tri = tri_idx[left_first + i];
        double a[3], b[3], c[3], t, u, v;
        load_tri(pos, tri, a, b, c);
        if (or_intersect_triangle(ro, rd, a, b, c, tmin, closest, &t, &u, &v)) {
          closest = t;
          found = 1;
          hit->t = t;
          hit->u = u;
          hit->v = v;
          hit->tri = tri;
          hit->slot = left_first + i;
        }
      }
    } else {
      if (sp + 2 > 64) return -1;
      stack[sp++] = (int32_t)left_first + 1;
      stack[sp++] = (int32_t)left_first;
    }
  }
  return found;
}

/* bvh.ts:323-348 — brute force over every triangle in index order. */
int or_intersect_all(const float *pos, uint32_t tri_count, const double *ro, const double *rd, double tmin,
                     double tmax, or_hit *hit) {
  int found = 0;
  double closest = tmax;
  for (uint32_t tri = 0; tri < tri_count; tri++) {
    double a[3], b[3], c[3], t, u, v;
    load_tri(pos, tri, a, b, c);
    if (or_intersect_triangle(ro, rd, a, b, c, tmin, closest, &t, &u, &v)) {
      closest = t;
      found = 1;
      hit->t = t;
      hit->u = u;
      hit->v = v;
      hit->tri = tri;
      hit->slot = tri;
    }
  }
  return found;
}

/* Batched form of or_intersect_bvh over float32 rays (used for the large
 * configurations; the caller splits the batch across threads).  Misses write
 * id = UINT32_MAX and t = +inf. Returns the number of stack overflows. */
int or_intersect_bvh_batch(const float *nodes, const uint32_t *tri_idx, uint32_t tri_count, const float *pos,
                           const float *org3, const float *dir3, uint32_t n, double tmin, double tmax,
                           double *t_out, uint32_t *id_out, double *u_out, double *v_out) {
  int overflow = 0;
  for (uint32_t i = 0; i < n; i++) {
    double ro[3] = {org3[i * 3], org3[i * 3 + 1], org3[i * 3 + 2]};
    double rd[3] = {dir3[i * 3], dir3[i * 3 + 1], dir3[i * 3 + 2]};
    or_hit h;
    int r = or_intersect_bvh(nodes, tri_idx, tri_count, pos, ro, rd, tmin, tmax, &h);
    if (r < 0) overflow++;
    if (r == 1) {
      t_out[i] = h.t;
      id_out[i] = h.tri;
      u_out[i] = h.u;
      v_out[i] = h.v;
    } else {
      t_out[i] = INFINITY;
      id_out[i] = UINT32_MAX;
      u_out[i] = 0;
      v_out[i] = 0;
    }
  }
  return overflow;
}

/* Same, double-precision rays (the reference's JS-number rays). */
int or_intersect_bvh_batch_f64(const float *nodes, const uint32_t *tri_idx, uint32_t tri_count,
                               const float *pos, const double *org3, const double *dir3, uint32_t n,
                               double tmin, double tmax, double *t_out, uint32_t *id_out, double *u_out,
                               double *v_out) {
  int overflow = 0;
  for (uint32_t i = 0; i < n; i++) {
    or_hit h;
    int r = or_intersect_bvh(nodes, tri_idx, tri_count, pos, org3 + i * 3, dir3 + i * 3, tmin, tmax, &h);
    if (r < 0) overflow++;
    if (r == 1) {
      t_out[i] = h.t;
      id_out[i] = h.tri;
      u_out[i] = h.u;
      v_out[i] = h.v;
    } else {
      t_out[i] = INFINITY;
      id_out[i] = UINT32_MAX;
      u_out[i] = 0;
      v_out[i] = 0;
    }
  }
  return overflow;
}

int or_intersect_all_batch_f64(const float *pos, uint32_t tri_count, const double *org3, const double *dir3,
                               uint32_t n, double tmin, double tmax, double *t_out, uint32_t *id_out,
                               double *u_out, double *v_out) {
  for (uint32_t i = 0; i < n; i++) {
    or_hit h;
    if (or_intersect_all(pos, tri_count, org3 + i * 3, dir3 + i * 3, tmin, tmax, &h)) {
      t_out[i] = h.t;
      id_out[i] = h.tri;
      u_out[i] = h.u;
      v_out[i] = h.v;
    } else {
      t_out[i] = INFINITY;
      id_out[i] = UINT32_MAX;
      u_out[i] = 0;
      v_out[i] = 0;
    }
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * The reference shader's closest-hit traversal order (scene.glsl:321-376) in
 * float32: root box untested, near child first, 40-entry stack that drops on
 * overflow, at most 8192 iterations.  Used only to count, per ray, the
 * interior visits P (each fetches a 64-byte child pair) and the triangle
 * tests T (48 bytes each) that define the algorithmic bytes of a ray.
 * ------------------------------------------------------------------------- */
static float hit_aabb_f32(const float *ro, const float *inv, const float *bmin, const float *bmax, float tmax) {
  float tn = 0.0f, tf = tmax;
  float lo[3], hi[3];
  for (int i = 0; i < 3; i++) {
    float t0 = (bmin[i] - ro[i]) * inv[i];
    float t1 = (bmax[i] - ro[i]) * inv[i];
    lo[i] = fminf(t0, t1);
    hi[i] = fmaxf(t0, t1);
  }
  tn = fmaxf(fmaxf(lo[0], lo[1]), fmaxf(lo[2], 0.0f));
  tf = fminf(fminf(hi[0], hi[1]), fminf(hi[2], tmax));
  return tf >= tn ? tn : 1e30f;
}

static int hit_tri_f32(const float *ro, const float *rd, const float *v0, const float *v1, const float *v2,
                       float tmax, float *t_out) {
  float e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  float e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  float p[3] = {rd[1] * e2[2] - rd[2] * e2[1], rd[2] * e2[0] - rd[0] * e2[2], rd[0] * e2[1] - rd[1] * e2[0]};
  float det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
  if (fabsf(det) < 1e-12f) return 0;
  float inv_det = 1.0f / det;
  float s[3] = {ro[0] - v0[0], ro[1] - v0[1], ro[2] - v0[2]};
  float u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv_det;
  if (u < 0.0f || u > 1.0f) return 0;
  float q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
  float v = (rd[0] * q[0] + rd[1] * q[1] + rd[2] * q[2]) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return 0;
  float t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv_det;
  if (t <= 1e-6f || t >= tmax) return 0;
  *t_out = t;
  return 1;
}

void or_shader_order_counts(const float *nodes, const uint32_t *tri_idx, uint32_t tri_count, const float *pos,
                            const float *org3, const float *dir3, uint32_t n, uint64_t *interior_visits,
                            uint64_t *tri_tests, uint64_t *leaf_visits) {
  uint64_t P = 0, T = 0, L = 0;
  for (uint32_t r = 0; r < n; r++) {
    const float *ro = org3 + (size_t)r * 3, *rd = dir3 + (size_t)r * 3;
    float inv[3] = {1.0f / rd[0], 1.0f / rd[1], 1.0f / rd[2]};
    float closest = 1e30f;
    int32_t stack[40];
    int sp = 0;
    int32_t node = 0;
    if (tri_count == 0) continue;
    for (int iter = 0; iter < 8192; iter++) {
      const float *a = nodes + (size_t)node * 8;
      int32_t left_first = (int32_t)a[3];
      int32_t count = (int32_t)a[7];
      if (count > 0) {
        L++;
        for (int k = 0; k < count; k++) {
          uint32_t tri = tri_idx[left_first + k];
          const float *v = pos + (size_t)tri * 9;
          float t;
          T++;
          if (hit_tri_f32(ro, rd, v, v + 3, v + 6, closest, &t)) closest = t;
        }
        if (sp == 0) break;
        node = stack[--sp];
        continue;
      }
      P++;
      int32_t left = left_first, right = left_first + 1;
      const float *la = nodes + (size_t)left * 8, *ra = nodes + (size_t)right * 8;
      float tl = hit_aabb_f32(ro, inv, la, la + 4, closest);
      float tr = hit_aabb_f32(ro, inv, ra, ra + 4, closest);
      if (tl > tr) {
        float tt = tl;
        tl = tr;
        tr = tt;
        int32_t ti = left;
        left = right;
        right = ti;
      }
      if (tl == 1e30f) {
        if (sp == 0) break;
        node = stack[--sp];
      } else {
        node = left;
        if (tr != 1e30f && sp < 40) stack[sp++] = right;
      }
    }
  }
  *interior_visits = P;
  *tri_tests = T;
  *leaf_visits = L;
}

/* Depth of the deepest leaf (root = 0). */
uint32_t or_bvh_depth(const float *nodes, uint32_t node_count) {
  if (node_count == 0) return 0;
  uint32_t *depth = (uint32_t *)calloc(node_count, sizeof(uint32_t));
  uint32_t best = 0;
  /* Children are always numbered after their parent. */
  for (uint32_t n = 0; n < node_count; n++) {
    const float *o = nodes + (size_t)n * 8;
    if (o[7] == 0 && node_count > 1) {
      uint32_t l = (uint32_t)o[3];
      if (l + 1 < node_count) {
        depth[l] = depth[n] + 1;
        depth[l + 1] = depth[n] + 1;
      }
    }
    if (depth[n] > best) best = depth[n];
  }
  free(depth);
  return best;
}

/* ---------------------------------------------------------------------------
 * mesh.ts:95-156 — (p, q) torus knot with ring normals; f64 math stored f32.
 * Writes tubular*radial*2 triangles (9 floats each) into pos and nrm.
 * ------------------------------------------------------------------------- */
static void knot_curve(double t, double p, double q, double radius, double *o) {
  double cu = cos(t), su = sin(t);
  double qu_over_p = (q / p) * t;
  double cs = cos(qu_over_p);
  o[0] = radius * (2 + cs) * 0.5 * cu;
  o[1] = radius * (2 + cs) * su * 0.5;
  o[2] = radius * sin(qu_over_p) * 0.5;
}

void or_torus_knot(double p, double q, double radius, double tube, uint32_t tubular, uint32_t radial,
                   float *pos, float *nrm) {
  double *ring = (double *)malloc((size_t)tubular * radial * 3 * sizeof(double));
  double *ringn = (double *)malloc((size_t)tubular * radial * 3 * sizeof(double));
  for (uint32_t i = 0; i < tubular; i++) {
    double u = ((double)i / tubular) * p * M_PI * 2;
    double P1[3], P2[3], T[3], N0[3], B[3], N[3], tmp[3];
    knot_curve(u, p, q, radius, P1);
    knot_curve(u + 0.01, p, q, radius, P2);
    v_sub(P2, P1, T);
    N0[0] = P2[0] + P1[0];
    N0[1] = P2[1] + P1[1];
    N0[2] = P2[2] + P1[2];
    v_cross(T, N0, tmp);
    v_normalize(tmp, B);
    v_cross(B, T, tmp);
    v_normalize(tmp, N);
    for (uint32_t j = 0; j < radial; j++) {
      double v = ((double)j / radial) * M_PI * 2;
      double cx = -tube * cos(v);
      double cy = tube * sin(v);
      double off[3] = {N[0] * cx + B[0] * cy, N[1] * cx + B[1] * cy, N[2] * cx + B[2] * cy};
      double *r = ring + ((size_t)i * radial + j) * 3;
      r[0] = P1[0] + off[0];
      r[1] = P1[1] + off[1];
      r[2] = P1[2] + off[2];
      v_normalize(off, ringn + ((size_t)i * radial + j) * 3);
    }
  }
  size_t o = 0;
  for (uint32_t i = 0; i < tubular; i++) {
    uint32_t i1 = (i + 1) % tubular;
    for (uint32_t j = 0; j < radial; j++) {
      uint32_t j1 = (j + 1) % radial;
      uint32_t tri[2][6] = {{i, j, i1, j, i1, j1}, {i, j, i1, j1, i, j1}};
      for (int t = 0; t < 2; t++) {
        for (int k = 0; k < 3; k++) {
          size_t src = ((size_t)tri[t][k * 2] * radial + tri[t][k * 2 + 1]) * 3;
          for (int c = 0; c < 3; c++) {
            pos[o + k * 3 + c] = (float)ring[src + c];
            nrm[o + k * 3 + c] = (float)ringn[src + c];
          }
        }
        o += 9;
      }
    }
  }
  free(ring);
  free(ringn);
}
