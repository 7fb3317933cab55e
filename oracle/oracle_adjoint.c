/*
 * oracle_adjoint.c -- chi^2 of a spatially constant, shifted Gauss-Hermite
 * kernel and its exact gradient.  TEST INFRASTRUCTURE ONLY (see
 * oracle_core.h).
 *
 * Model (the reference's subtraction model, src/subtract.cpp:466-639, with
 * one spatial term, i.e. a constant kernel, and with the basis sampled at a
 * shifted position):
 *   K(u)   = sum_n a_n g_nx(ux - dx) g_ny(uy - dy),   u in [-r, r]^2
 *   D(x)   = S(x) - b - sum_u K(u) R(x - u)            (R zero outside)
 *   chi^2  = sum_x w(x) D(x)^2                         (w <= 0: excluded)
 * with g_n the sampled Gauss-Hermite function of src/basis.cpp:14-59,
 *   g_n(x) = [2^n n! sqrt(pi) beta]^(-1/2) H_n(x/beta) exp(-x^2 / 2 beta^2),
 * which the reference evaluates at x = i - r (dx = dy = 0).  The shift and
 * the beta derivative have no counterpart in the reference.
 *
 * Gradient (adjoint):  with C(u) = sum_x w D R(x - u),
 *   d chi^2 / d a_n  = -2 sum_u phi_n(u) C(u)
 *   d chi^2 / d b    = -2 sum_x w D
 *   d chi^2 / d p    = -2 sum_u C(u) dK(u)/dp        for p = beta, dx, dy
 * and the Hermite ladder identities
 *   dg_n/dx    = (1/beta) (sqrt(n/2) g_{n-1} - sqrt((n+1)/2) g_{n+1})
 *   dg_n/dbeta = -(1/beta) (g_n / 2 + x dg_n/dx).
 *
 * Pinning (tests/test_oracle_adjoint.py): at zero shift the samples equal
 * gh_init (itself bit-exact against the reference's basis.cpp); the ladder
 * derivatives equal central differences of the direct evaluation; the model
 * equals or_subtract with one knot (the constant-kernel mode) at zero shift;
 * an impulse reference reproduces K; the analytic gradient equals central
 * finite differences of chi^2 (step 1e-3).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_core.h"

/* g_n(x_i), dg_n/dx and dg_n/dbeta at x_i = i - radius - shift, n = 0..n_max
 * (each (n_max + 1) x ks, row n). */
OR_EXPORT int or_gh_shifted(double beta, int n_max, int radius, double shift, double *g,
                            double *dg_dx, double *dg_dbeta) {
  if (!(beta > 0.0) || n_max < 0 || radius < 1) return -1;
  const int ks = 2 * radius + 1;
  const int nn = n_max + 2; /* one order beyond n_max for the ladder */
  double *gg = (double *)malloc(sizeof(double) * nn * ks);
  double *lnorm = (double *)malloc(sizeof(double) * nn);
  const double lnpi_half = 0.5 * log(acos(-1.0));
  for (int n = 0; n < nn; ++n)
    lnorm[n] = -0.5 * (n * log(2.0) + lgamma(n + 1.0) + lnpi_half + log(beta));
  for (int i = 0; i < ks; ++i) {
    const double x = (double)(i - radius) - shift;
    const double t = x / beta;
    const double e = exp(-0.5 * t * t);
    double hm = 1.0, h = 2.0 * t;
    gg[i] = exp(lnorm[0]) * hm * e;
    gg[ks + i] = exp(lnorm[1]) * h * e;
    for (int n = 2; n < nn; ++n) {
      const double hn = 2.0 * t * h - 2.0 * (n - 1) * hm;
      gg[(size_t)n * ks + i] = exp(lnorm[n]) * hn * e;
      hm = h;
      h = hn;
    }
  }
  for (int n = 0; n <= n_max; ++n)
    for (int i = 0; i < ks; ++i) {
      const double x = (double)(i - radius) - shift;
      const double gn = gg[(size_t)n * ks + i];
      const double lo = n > 0 ? sqrt(0.5 * n) * gg[(size_t)(n - 1) * ks + i] : 0.0;
      const double hi = sqrt(0.5 * (n + 1)) * gg[(size_t)(n + 1) * ks + i];
      const double d = (lo - hi) / beta;
      if (g) g[(size_t)n * ks + i] = gn;
      if (dg_dx) dg_dx[(size_t)n * ks + i] = d;
      if (dg_dbeta) dg_dbeta[(size_t)n * ks + i] = -(0.5 * gn + x * d) / beta;
    }
  free(gg);
  free(lnorm);
  return 0;
}

/* Dense kernel K (ks x ks, row = uy + r) and, when dk != NULL, the three
 * derivative kernels dK/dbeta, dK/ddx, dK/ddy (3 x ks x ks).
 * params = [a_0 .. a_{nc-1}, b, beta, dx, dy]. */
OR_EXPORT int or_shifted_kernel(const double *params, int n_max, int radius, double *k2,
                                double *dk) {
  const int nc = (n_max + 1) * (n_max + 2) / 2, ks = 2 * radius + 1;
  const double beta = params[nc + 1], dx = params[nc + 2], dy = params[nc + 3];
  const size_t tab = (size_t)(n_max + 1) * ks;
  double *gx = (double *)malloc(sizeof(double) * tab * 6);
  double *dgx = gx + tab, *bgx = gx + 2 * tab, *gy = gx + 3 * tab, *dgy = gx + 4 * tab,
         *bgy = gx + 5 * tab;
  if (or_gh_shifted(beta, n_max, radius, dx, gx, dgx, bgx) ||
      or_gh_shifted(beta, n_max, radius, dy, gy, dgy, bgy)) {
    free(gx);
    return -1;
  }
  memset(k2, 0, sizeof(double) * ks * ks);
  if (dk) memset(dk, 0, sizeof(double) * 3 * ks * ks);
  int c = 0;
  for (int tot = 0; tot <= n_max; ++tot)
    for (int ny = 0; ny <= tot; ++ny, ++c) {
      const int nx = tot - ny;
      const double a = params[c];
      for (int iy = 0; iy < ks; ++iy)
        for (int ix = 0; ix < ks; ++ix) {
          const size_t u = (size_t)iy * ks + ix;
          const double vx = gx[(size_t)nx * ks + ix], vy = gy[(size_t)ny * ks + iy];
          k2[u] += a * vy * vx;
          if (dk) {
            dk[u] += a * (bgy[(size_t)ny * ks + iy] * vx + vy * bgx[(size_t)nx * ks + ix]);
            /* d/d(dx) of g(i - r - dx) is -g'(x) */
            dk[(size_t)ks * ks + u] -= a * vy * dgx[(size_t)nx * ks + ix];
            dk[(size_t)2 * ks * ks + u] -= a * dgy[(size_t)ny * ks + iy] * vx;
          }
        }
    }
  free(gx);
  return 0;
}

/* chi^2 and (grad != NULL) its gradient with respect to params (length
 * nc + 4); diff (optional) receives D.  Image arithmetic in `real` (the FP64
 * build is the accuracy reference); all sums over pixels in double. */
OR_EXPORT int or_kernel_chi2(int w, int h, const float *sci, const float *ref,
                             const float *weight, const double *params, int n_max, int radius,
                             double *chi2, double *grad, real *diff) {
  const int nc = (n_max + 1) * (n_max + 2) / 2, ks = 2 * radius + 1, r = radius;
  if (w < 1 || h < 1 || radius < 1 || n_max < 0) return -1;
  double *k2 = (double *)malloc(sizeof(double) * ks * ks * 4);
  double *dk = k2 + (size_t)ks * ks;
  if (or_shifted_kernel(params, n_max, radius, k2, grad ? dk : NULL)) {
    free(k2);
    return -1;
  }
  real *kr = (real *)malloc(sizeof(real) * ks * ks);
  for (int i = 0; i < ks * ks; ++i) kr[i] = (real)k2[i];
  const real b = (real)params[nc];
  double *cu = (double *)calloc((size_t)ks * ks, sizeof(double));
  double total = 0.0, sum_wd = 0.0;
#pragma omp parallel
  {
    double *cl = (double *)calloc((size_t)ks * ks, sizeof(double));
    double tl = 0.0, sl = 0.0;
#pragma omp for schedule(static)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        real m = 0;
        for (int iy = 0; iy < ks; ++iy) {
          const int yy = y - (iy - r);
          if (yy < 0 || yy >= h) continue;
          for (int ix = 0; ix < ks; ++ix) {
            const int xx = x - (ix - r);
            if (xx < 0 || xx >= w) continue;
            m = RFMA(kr[iy * ks + ix], (real)ref[(size_t)yy * w + xx], m);
          }
        }
        const real d = ((real)sci[(size_t)y * w + x] - b) - m;
        if (diff) diff[(size_t)y * w + x] = d;
        const double wt = weight ? (double)weight[(size_t)y * w + x] : 1.0;
        if (!(wt > 0.0)) continue;
        const double wd = wt * (double)d;
        tl += wd * (double)d;
        if (!grad) continue;
        sl += wd;
        for (int iy = 0; iy < ks; ++iy) {
          const int yy = y - (iy - r);
          if (yy < 0 || yy >= h) continue;
          for (int ix = 0; ix < ks; ++ix) {
            const int xx = x - (ix - r);
            if (xx < 0 || xx >= w) continue;
            cl[iy * ks + ix] += wd * (double)ref[(size_t)yy * w + xx];
          }
        }
      }
#pragma omp critical
    {
      total += tl;
      sum_wd += sl;
      for (int i = 0; i < ks * ks; ++i) cu[i] += cl[i];
    }
    free(cl);
  }
  *chi2 = total;
  if (grad) {
    const size_t tab = (size_t)(n_max + 1) * ks;
    double *gx = (double *)malloc(sizeof(double) * tab * 2);
    double *gy = gx + tab;
    or_gh_shifted(params[nc + 1], n_max, radius, params[nc + 2], gx, NULL, NULL);
    or_gh_shifted(params[nc + 1], n_max, radius, params[nc + 3], gy, NULL, NULL);
    int c = 0;
    for (int tot = 0; tot <= n_max; ++tot)
      for (int ny = 0; ny <= tot; ++ny, ++c) {
        const int nx = tot - ny;
        double s = 0.0;
        for (int iy = 0; iy < ks; ++iy)
          for (int ix = 0; ix < ks; ++ix)
            s += gy[(size_t)ny * ks + iy] * gx[(size_t)nx * ks + ix] * cu[iy * ks + ix];
        grad[c] = -2.0 * s;
      }
    grad[nc] = -2.0 * sum_wd;
    for (int p = 0; p < 3; ++p) {
      double s = 0.0;
      for (int i = 0; i < ks * ks; ++i) s += dk[(size_t)p * ks * ks + i] * cu[i];
      grad[nc + 1 + p] = -2.0 * s;
    }
    free(gx);
  }
  free(cu);
  free(kr);
  free(k2);
  return 0;
}
