/* TEST INFRASTRUCTURE — declarations of the FP64 parity oracle (see bvh_oracle.c
 * and pt_oracle.c).  Nothing in the product includes this header. */
#ifndef LUCENT_ORACLE_H
#define LUCENT_ORACLE_H
#include <stdint.h>

typedef struct {
  double t, u, v;
  uint32_t tri;  /* original triangle id */
  uint32_t slot; /* leaf-order position */
} or_hit;

double or_lcg_next(uint32_t *state);
void or_random_soup(uint32_t count, uint32_t *state, float *out);
int or_intersect_triangle(const double *ro, const double *rd, const double *v0, const double *v1,
                          const double *v2, double tmin, double tmax, double *t, double *u, double *v);
double or_intersect_sphere(const double *ro, const double *rd, const double *c, double radius, double tmin,
                           double tmax);
double or_aabb_intersect(const double *bmin, const double *bmax, const double *o, const double *inv,
                         double tmax);
int or_build_bvh(const float *pos, uint32_t tri_count, uint32_t max_leaf, uint32_t bins, float *nodes,
                 uint32_t *node_count, uint32_t *tri_idx, double *bounds6);
int or_intersect_bvh(const float *nodes, const uint32_t *tri_idx, uint32_t tri_count, const float *pos,
                     const double *ro, const double *rd, double tmin, double tmax, or_hit *hit);
int or_intersect_all(const float *pos, uint32_t tri_count, const double *ro, const double *rd, double tmin,
                     double tmax, or_hit *hit);
uint32_t or_bvh_depth(const float *nodes, uint32_t node_count);
void or_torus_knot(double p, double q, double radius, double tube, uint32_t tubular, uint32_t radial,
                   float *pos, float *nrm);
#endif
