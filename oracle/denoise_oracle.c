/*
 * TEST INFRASTRUCTURE — the parity oracle, not the product.
 *
 * FP64 restatement of Lucent's denoiser passes: variance.frag (demodulation
 * and variance estimate, src/shaders/variance.frag:15-55 with
 * denoise.glsl:10-25) and atrous.frag (src/shaders/atrous.frag:22-82), run
 * in sequence as Renderer.runDenoiser does (src/renderer.ts:788-830).  The
 * inputs are the film's stored values (float32 colour and history, FP16
 * guides widened to float), so the oracle differs from the product only in
 * the precision of its arithmetic.
 *
 * Only tests/ and __graft_entry__.smoke() may load this library.
 * Parity pins: with sigma_n = 0 and sigma_p, sigma_l very large every edge
 * weight is 1 and the filter reduces to a plain B3-spline convolution; a
 * constant input is a fixed point; these are checked in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double lum3(double r, double g, double b) { return 0.2126 * r + 0.7152 * g + 0.0722 * b; }
static void unit_normal(const float *n, double *o) {
  double l = sqrt((double)n[0] * n[0] + (double)n[1] * n[1] + (double)n[2] * n[2]);
  if (l > 1e-6) {
    o[0] = n[0] / l;
    o[1] = n[1] / l;
    o[2] = n[2] / l;
  } else {
    o[0] = 0;
    o[1] = 0;
    o[2] = 1;
  }
}
static double maxd(double a, double b) { return a > b ? a : b; }

static void variance_pass(const float *color, const float *albedo, const float *normal, const float *history, int w,
                          int h, double *out) {
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t i = (size_t)y * w + x;
      const float *c = color + i * 4, *a = albedo + i * 4;
      double *o = out + i * 4;
      if (a[3] < 0.5f) {
        o[0] = c[0];
        o[1] = c[1];
        o[2] = c[2];
        o[3] = 0;
        continue;
      }
      double samples = maxd(history[i], 1.0), la = maxd(lum3(a[0], a[1], a[2]), 1e-3), var;
      if (samples >= 4.0) {
        double l = lum3(c[0], c[1], c[2]);
        var = maxd(0.0, c[3] - l * l) / (la * la) / samples;
      } else {
        const float *np = normal + i * 4;
        double n[3], s1 = 0, s2 = 0, ws = 0;
        unit_normal(np, n);
        for (int dy = -3; dy <= 3; dy++)
          for (int dx = -3; dx <= 3; dx++) {
            int qx = x + dx, qy = y + dy;
            if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
            size_t q = (size_t)qy * w + qx;
            const float *aq = albedo + q * 4;
            if (aq[3] < 0.5f) continue;
            double u[3];
            unit_normal(normal + q * 4, u);
            double wn = pow(maxd(0, n[0] * u[0] + n[1] * u[1] + n[2] * u[2]), 32.0);
            double wz = fabs((double)np[3] - normal[q * 4 + 3]) <= 0.1 * np[3] ? 1 : 0;
            const float *cq = color + q * 4;
            double l = lum3(cq[0] / maxd(aq[0], 1e-3), cq[1] / maxd(aq[1], 1e-3), cq[2] / maxd(aq[2], 1e-3));
            s1 += wn * wz * l;
            s2 += wn * wz * l * l;
            ws += wn * wz;
          }
        double m1 = s1 / maxd(ws, 1e-6);
        var = maxd(0, s2 / maxd(ws, 1e-6) - m1 * m1) / samples;
      }
      o[0] = c[0] / maxd(a[0], 1e-3);
      o[1] = c[1] / maxd(a[1], 1e-3);
      o[2] = c[2] / maxd(a[2], 1e-3);
      o[3] = var;
    }
}

/* Frame geometry of the pixels being filtered: a window at (x0, y0) of a
 * full_w x full_h frame (the whole frame when x0 = y0 = 0). */
static int g_x0, g_y0, g_fw, g_fh;

static void primary_dir(const float *cam, int px, int py, int w, int h, double *d) {
  px += g_x0;
  py += g_y0;
  w = g_fw;
  h = g_fh;
  /* cam: pos[3] right[3] up[3] fwd[3] tan */
  double ux = (px + 0.5) / w * 2 - 1, uy = (py + 0.5) / h * 2 - 1, aspect = (double)w / h, t = cam[12];
  for (int k = 0; k < 3; k++) d[k] = cam[9 + k] + cam[3 + k] * (ux * aspect * t) + cam[6 + k] * (uy * t);
  double l = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; k++) d[k] /= l;
}

static void atrous_pass(const double *in, const float *albedo, const float *normal, int w, int h, int step,
                        const float *cam, double sl, double sn, double sp, double *out) {
  static const double hk[3] = {0.375, 0.25, 0.0625};
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      size_t i = (size_t)y * w + x;
      const double *cp = in + i * 4;
      double *o = out + i * 4;
      if (albedo[i * 4 + 3] < 0.5f) {
        memcpy(o, cp, 4 * sizeof(double));
        continue;
      }
      const float *np = normal + i * 4;
      double n[3], d[3], P[3];
      unit_normal(np, n);
      primary_dir(cam, x, y, w, h, d);
      for (int k = 0; k < 3; k++) P[k] = cam[k] + d[k] * np[3];
      double footprint = 2 * cam[12] * np[3] / g_fh;
      double var = 0, wv = 0;
      for (int oy = -1; oy <= 1; oy++)
        for (int ox = -1; ox <= 1; ox++) {
          int qx = x + ox, qy = y + oy;
          if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
          double k = (ox == 0 ? 0.5 : 0.25) * (oy == 0 ? 0.5 : 0.25);
          var += in[((size_t)qy * w + qx) * 4 + 3] * k;
          wv += k;
        }
      double denom = sl * sqrt(maxd(var / wv, 0)) + 1e-6, lp = lum3(cp[0], cp[1], cp[2]);
      double s[3] = {cp[0] * hk[0] * hk[0], cp[1] * hk[0] * hk[0], cp[2] * hk[0] * hk[0]};
      double svar = cp[3] * hk[0] * hk[0] * hk[0] * hk[0], wsum = hk[0] * hk[0];
      for (int oy = -2; oy <= 2; oy++)
        for (int ox = -2; ox <= 2; ox++) {
          if (ox == 0 && oy == 0) continue;
          int qx = x + ox * step, qy = y + oy * step;
          if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
          size_t q = (size_t)qy * w + qx;
          if (albedo[q * 4 + 3] < 0.5f) continue;
          const double *cq = in + q * 4;
          double u[3], qd[3], Q[3];
          unit_normal(normal + q * 4, u);
          double k = hk[abs(ox)] * hk[abs(oy)];
          double wn = pow(maxd(0, n[0] * u[0] + n[1] * u[1] + n[2] * u[2]), sn);
          primary_dir(cam, qx, qy, w, h, qd);
          for (int c = 0; c < 3; c++) Q[c] = cam[c] + qd[c] * normal[q * 4 + 3];
          double reach = sp * footprint * step * sqrt((double)(ox * ox + oy * oy));
          double wp = exp(-fabs(n[0] * (Q[0] - P[0]) + n[1] * (Q[1] - P[1]) + n[2] * (Q[2] - P[2])) / (reach + 1e-6));
          double wl = exp(-fabs(lp - lum3(cq[0], cq[1], cq[2])) / denom);
          double wgt = k * wn * wp * wl;
          for (int c = 0; c < 3; c++) s[c] += cq[c] * wgt;
          svar += cq[3] * wgt * wgt;
          wsum += wgt;
        }
      o[0] = s[0] / wsum;
      o[1] = s[1] / wsum;
      o[2] = s[2] / wsum;
      o[3] = svar / (wsum * wsum);
    }
}

/* cam13: pos, right, up, fwd (float32 uniforms), tanHalfFov.  out: illum+var
 * (4 doubles per pixel); remod: illumination * max(albedo, 1e-3) (3 doubles). */
int or_denoise_window(const float *color, const float *albedo, const float *normal, const float *history, int w,
                      int h, int x0, int y0, int full_w, int full_h, const float *cam13, double sigma_l,
                      double sigma_n, double sigma_p, int iterations, double *out, double *remod) {
  g_x0 = x0;
  g_y0 = y0;
  g_fw = full_w;
  g_fh = full_h;
  size_t n = (size_t)w * h;
  double *a = malloc(n * 4 * sizeof(double)), *b = malloc(n * 4 * sizeof(double));
  variance_pass(color, albedo, normal, history, w, h, a);
  for (int it = 0; it < iterations; it++) {
    atrous_pass(a, albedo, normal, w, h, 1 << it, cam13, sigma_l, sigma_n, sigma_p, b);
    double *t = a;
    a = b;
    b = t;
  }
  memcpy(out, a, n * 4 * sizeof(double));
  if (remod)
    for (size_t i = 0; i < n; i++)
      for (int c = 0; c < 3; c++) remod[i * 3 + c] = a[i * 4 + c] * maxd(albedo[i * 4 + c], 1e-3);
  free(a);
  free(b);
  return 0;
}

int or_denoise(const float *color, const float *albedo, const float *normal, const float *history, int w, int h,
               const float *cam13, double sigma_l, double sigma_n, double sigma_p, int iterations, double *out,
               double *remod) {
  return or_denoise_window(color, albedo, normal, history, w, h, 0, 0, w, h, cam13, sigma_l, sigma_n, sigma_p,
                           iterations, out, remod);
}
