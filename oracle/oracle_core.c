/*
 * oracle_core.c -- basis, thin-plate spline, coarse field lattice and order
 * statistics of the CPU oracle.  TEST INFRASTRUCTURE ONLY (see oracle_core.h).
 *
 * Pinning (how each function is checked, tests/test_oracle_*.py):
 *   gh_init            == the reference's own basis.cpp compiled into
 *                         oracle/_ref (bit-exact), and the closed form
 *                         g_n = [2^n n! sqrt(pi) beta]^-1/2 H_n(x/b) e^{-x^2/2b^2}
 *                         evaluated with numpy's hermval.
 *   tps_init/design    == numpy: Z against LAPACK's complete QR (numpy.linalg.qr,
 *                         same Householder sign convention as Eigen), the
 *                         penalty against Z^T E Z, the design against
 *                         [U(r) Z | 1 u v] built independently.
 *   cf_build/cf_point  exact mode == numpy design @ C; coarse mode bilinear
 *                         bound of test_subtract.py:45-70.
 *   fill_value_f       == numpy sort of the stride-7 sample, element n/2.
 */
#include "oracle_core.h"

#include <float.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Gauss-Hermite basis: src/basis.cpp:9-59                                   */
/* ------------------------------------------------------------------------ */
int gh_init(gh_basis *b, double beta, int n_max, int radius) {
  memset(b, 0, sizeof(*b));
  if (!(beta > 0.0) || n_max < 0 || radius < 1) return -1;
  b->beta = beta;
  b->n_max = n_max;
  b->radius = radius;
  b->ks = 2 * radius + 1;
  b->nc = (n_max + 1) * (n_max + 2) / 2;
  const int ks = b->ks;
  b->b1d = (double *)calloc((size_t)(n_max + 1) * ks, sizeof(double));
  b->onx = (int *)malloc(sizeof(int) * b->nc);
  b->ony = (int *)malloc(sizeof(int) * b->nc);
  const double lnpi_half = 0.5 * log(acos(-1.0));
  double *lnorm = (double *)malloc(sizeof(double) * (n_max + 1));
  for (int n = 0; n <= n_max; ++n)
    lnorm[n] = -0.5 * (n * log(2.0) + lgamma(n + 1.0) + lnpi_half + log(beta));
  for (int i = 0; i < ks; ++i) {
    const double t = (double)(i - radius) / beta;
    const double g = exp(-0.5 * t * t);
    double hm = 1.0, h = 2.0 * t;
    b->b1d[i] = exp(lnorm[0]) * hm * g;
    if (n_max >= 1) b->b1d[ks + i] = exp(lnorm[1]) * h * g;
    for (int n = 2; n <= n_max; ++n) {
      const double hn = 2.0 * t * h - 2.0 * (n - 1) * hm;
      b->b1d[(size_t)n * ks + i] = exp(lnorm[n]) * hn * g;
      hm = h;
      h = hn;
    }
  }
  free(lnorm);
  int c = 0;
  for (int tot = 0; tot <= n_max; ++tot)
    for (int ny = 0; ny <= tot; ++ny) {
      b->onx[c] = tot - ny;
      b->ony[c] = ny;
      ++c;
    }
  return 0;
}

void gh_free(gh_basis *b) {
  free(b->b1d);
  free(b->onx);
  free(b->ony);
  memset(b, 0, sizeof(*b));
}

void gh_kernel2d(const gh_basis *b, int c, real *out) {
  const int ks = b->ks;
  const double *gx = b->b1d + (size_t)b->onx[c] * ks;
  const double *gy = b->b1d + (size_t)b->ony[c] * ks;
  for (int iy = 0; iy < ks; ++iy)
    for (int ix = 0; ix < ks; ++ix) out[iy * ks + ix] = (real)(gy[iy] * gx[ix]);
}

/* ------------------------------------------------------------------------ */
/* Thin-plate spline: src/spatial.cpp:13-111                                 */
/* ------------------------------------------------------------------------ */
static double tps_green(double r2) { return r2 > 0.0 ? 0.5 * r2 * log(r2) : 0.0; }

/* Householder QR of the k x 3 trend matrix [1 u v] and the complete Q,
 * Q = H0 H1 H2, with the LAPACK/Eigen reflector convention
 * beta = -sign(alpha) * ||x||, v = x / (alpha - beta), tau = (beta - alpha)/beta
 * (Eigen HouseholderQR as used at spatial.cpp:54-57). */
static void trend_nullspace(const double *kn, int k, double *Z) {
  double *a = (double *)malloc(sizeof(double) * k * 3);
  double *v = (double *)malloc(sizeof(double) * k * 3);
  double tau[3];
  for (int i = 0; i < k; ++i) {
    a[i * 3 + 0] = 1.0;
    a[i * 3 + 1] = kn[i * 2 + 0];
    a[i * 3 + 2] = kn[i * 2 + 1];
  }
  for (int j = 0; j < 3; ++j) {
    const double c0 = a[j * 3 + j];
    double tail = 0.0;
    for (int i = j + 1; i < k; ++i) tail += a[i * 3 + j] * a[i * 3 + j];
    for (int i = 0; i < k; ++i) v[i * 3 + j] = 0.0;
    v[j * 3 + j] = 1.0;
    if (tail <= DBL_MIN) {
      tau[j] = 0.0;
      continue;
    }
    double beta = sqrt(c0 * c0 + tail);
    if (c0 >= 0.0) beta = -beta;
    for (int i = j + 1; i < k; ++i) v[i * 3 + j] = a[i * 3 + j] / (c0 - beta);
    tau[j] = (beta - c0) / beta;
    a[j * 3 + j] = beta;
    for (int i = j + 1; i < k; ++i) a[i * 3 + j] = 0.0;
    for (int c = j + 1; c < 3; ++c) {
      double s = 0.0;
      for (int i = j; i < k; ++i) s += v[i * 3 + j] * a[i * 3 + c];
      s *= tau[j];
      for (int i = j; i < k; ++i) a[i * 3 + c] -= s * v[i * 3 + j];
    }
  }
  /* Q = H0 H1 H2 I: apply reflectors in reverse order to the identity. */
  double *q = (double *)calloc((size_t)k * k, sizeof(double));
  for (int i = 0; i < k; ++i) q[i * k + i] = 1.0;
  for (int j = 2; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double s = 0.0;
      for (int i = j; i < k; ++i) s += v[i * 3 + j] * q[i * k + c];
      s *= tau[j];
      for (int i = j; i < k; ++i) q[i * k + c] -= s * v[i * 3 + j];
    }
  }
  for (int i = 0; i < k; ++i)
    for (int m = 0; m < k - 3; ++m) Z[i * (k - 3) + m] = q[i * k + 3 + m];
  free(q);
  free(a);
  free(v);
}

int tps_init(tps_basis *t, const double *knots, int k) {
  memset(t, 0, sizeof(*t));
  if (k != 1 && k < 3) return -1;
  t->k = k;
  t->kn = (double *)calloc((size_t)k * 2, sizeof(double));
  t->pen = (double *)calloc((size_t)k * k, sizeof(double));
  if (k == 1) { /* constant field: design row = [1], no penalty */
    t->scale = 1.0;
    return 0;
  }
  double sx = 0.0, sy = 0.0;
  for (int i = 0; i < k; ++i) {
    sx += knots[i * 2];
    sy += knots[i * 2 + 1];
  }
  t->cx = sx / k;
  t->cy = sy / k;
  double ex = 0.0, ey = 0.0;
  for (int i = 0; i < k; ++i) {
    ex = fmax(ex, fabs(knots[i * 2] - t->cx));
    ey = fmax(ey, fabs(knots[i * 2 + 1] - t->cy));
  }
  t->scale = fmax(fmax(ex, ey), 1e-12);
  for (int i = 0; i < k; ++i) {
    t->kn[i * 2] = (knots[i * 2] - t->cx) / t->scale;
    t->kn[i * 2 + 1] = (knots[i * 2 + 1] - t->cy) / t->scale;
  }
  const int nr = k - 3;
  t->Z = (double *)calloc((size_t)k * (nr > 0 ? nr : 1), sizeof(double));
  trend_nullspace(t->kn, k, t->Z);
  if (nr > 0) {
    double *e = (double *)malloc(sizeof(double) * k * k);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        const double dx = t->kn[i * 2] - t->kn[j * 2];
        const double dy = t->kn[i * 2 + 1] - t->kn[j * 2 + 1];
        e[i * k + j] = tps_green(dx * dx + dy * dy);
      }
    double *ez = (double *)malloc(sizeof(double) * k * nr);
    for (int i = 0; i < k; ++i)
      for (int m = 0; m < nr; ++m) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += e[i * k + j] * t->Z[j * nr + m];
        ez[i * nr + m] = s;
      }
    for (int a = 0; a < nr; ++a)
      for (int b = 0; b < nr; ++b) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += t->Z[i * nr + a] * ez[i * nr + b];
        t->pen[a * k + b] = s;
      }
    for (int a = 0; a < k; ++a) /* symmetrise */
      for (int b = a + 1; b < k; ++b) {
        const double m = 0.5 * (t->pen[a * k + b] + t->pen[b * k + a]);
        t->pen[a * k + b] = m;
        t->pen[b * k + a] = m;
      }
    free(e);
    free(ez);
  }
  return 0;
}

void tps_free(tps_basis *t) {
  free(t->kn);
  free(t->Z);
  free(t->pen);
  memset(t, 0, sizeof(*t));
}

void tps_design_row(const tps_basis *t, double px, double py, double *row) {
  const int k = t->k;
  if (k == 1) {
    row[0] = 1.0;
    return;
  }
  const int nr = k - 3;
  const double ux = (px - t->cx) / t->scale, uy = (py - t->cy) / t->scale;
  double sig[256];
  for (int j = 0; j < k; ++j) {
    const double dx = ux - t->kn[j * 2], dy = uy - t->kn[j * 2 + 1];
    sig[j] = tps_green(dx * dx + dy * dy);
  }
  for (int m = 0; m < nr; ++m) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += sig[j] * t->Z[j * nr + m];
    row[m] = s;
  }
  row[nr] = 1.0;
  row[nr + 1] = ux;
  row[nr + 2] = uy;
}

double tps_min_spacing(const tps_basis *t) {
  if (t->k < 2) return 0.0;
  double best = INFINITY;
  for (int i = 0; i < t->k; ++i)
    for (int j = i + 1; j < t->k; ++j) {
      const double dx = t->kn[i * 2] - t->kn[j * 2];
      const double dy = t->kn[i * 2 + 1] - t->kn[j * 2 + 1];
      const double d = dx * dx + dy * dy;
      if (d < best) best = d;
    }
  return isfinite(best) ? sqrt(best) * t->scale : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Coarse field lattice: src/subtract.cpp:113-316                            */
/* ------------------------------------------------------------------------ */
static int *axis_coords(int len, int stride, int *count) {
  int n = 0;
  for (int p = 0; p < len - 1; p += stride) ++n;
  int *g = (int *)malloc(sizeof(int) * (n + 1));
  n = 0;
  for (int p = 0; p < len - 1; p += stride) g[n++] = p;
  g[n++] = len - 1;
  *count = n;
  return g;
}

static void fields_at(const tps_basis *t, const double *theta, int nf, double px,
                      double py, double *out) {
  double row[256];
  tps_design_row(t, px, py, row);
  for (int f = 0; f < nf; ++f) {
    double s = 0.0;
    for (int m = 0; m < t->k; ++m) s += row[m] * theta[(size_t)f * t->k + m];
    out[f] = s;
  }
}

int cf_build(coarse_fields *cf, const tps_basis *t, const double *theta, int nc,
             int w, int h) {
  memset(cf, 0, sizeof(*cf));
  cf->nf = nc + 1;
  cf->w = w;
  cf->h = h;
  const int nf = cf->nf;
  int stride = 1;
  const double sp = tps_min_spacing(t);
  if (sp > 0.0) {
    stride = (int)(sp / 16.0);
    if (stride < 1) stride = 1;
    if (stride > 64) stride = 64;
  }
  cf->stride = stride;
  if (t->k == 1) {
    /* Constant fields: a 2x2 lattice with equal corners reproduces the
     * constant exactly through the ramp (fma(0, w, a) == a). */
    cf->exact = 0;
    cf->ncx = cf->ncy = 2;
    stride = (w > h ? w : h) + 1;
  } else if (stride < 4 || w <= 2 * stride || h <= 2 * stride) {
    cf->exact = 1;
    cf->full = (real *)malloc(sizeof(real) * (size_t)nf * w * h);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
      double v[512];
      for (int x = 0; x < w; ++x) {
        fields_at(t, theta, nf, x, y, v);
        for (int f = 0; f < nf; ++f)
          cf->full[((size_t)f * h + y) * w + x] = (real)v[f];
      }
    }
    return 0;
  }
  if (t->k == 1) {
    cf->gx = (int *)malloc(sizeof(int) * 2);
    cf->gy = (int *)malloc(sizeof(int) * 2);
    cf->gx[0] = 0;
    cf->gx[1] = w - 1;
    cf->gy[0] = 0;
    cf->gy[1] = h - 1;
  } else {
    cf->gx = axis_coords(w, stride, &cf->ncx);
    cf->gy = axis_coords(h, stride, &cf->ncy);
  }
  const int ncx = cf->ncx, ncy = cf->ncy;
  cf->coarse = (double *)malloc(sizeof(double) * (size_t)nf * ncx * ncy);
  for (int iy = 0; iy < ncy; ++iy)
    for (int ix = 0; ix < ncx; ++ix) {
      double v[512];
      fields_at(t, theta, nf, cf->gx[ix], cf->gy[iy], v);
      for (int f = 0; f < nf; ++f) cf->coarse[((size_t)f * ncy + iy) * ncx + ix] = v[f];
    }
  cf->cellx = (int *)malloc(sizeof(int) * w);
  cf->wx = (real *)malloc(sizeof(real) * w);
  for (int x = 0; x < w; ++x) {
    int c0 = x / stride;
    if (c0 > ncx - 2) c0 = ncx - 2;
    const int x0 = cf->gx[c0], x1 = cf->gx[c0 + 1];
    cf->cellx[x] = c0;
    cf->wx[x] = x1 > x0 ? (real)(x - x0) / (real)(x1 - x0) : (real)0;
  }
  cf->celly = (int *)malloc(sizeof(int) * h);
  cf->wy = (real *)malloc(sizeof(real) * h);
  for (int y = 0; y < h; ++y) {
    int c0 = y / stride;
    if (c0 > ncy - 2) c0 = ncy - 2;
    const int y0 = cf->gy[c0], y1 = cf->gy[c0 + 1];
    cf->celly[y] = c0;
    cf->wy[y] = y1 > y0 ? (real)(y - y0) / (real)(y1 - y0) : (real)0;
  }
  return 0;
}

void cf_free(coarse_fields *cf) {
  free(cf->coarse);
  free(cf->gx);
  free(cf->gy);
  free(cf->cellx);
  free(cf->celly);
  free(cf->wx);
  free(cf->wy);
  free(cf->full);
  memset(cf, 0, sizeof(*cf));
}

/* Bilinear cell ramp (subtract.cpp:187-205): value = base + slope * wx with
 * base = a00 + (a10-a00) fy and slope = (a01-a00) + ((a11-a10)-(a01-a00)) fy,
 * each product-sum contracted to a fused multiply-add. */
real cf_point(const coarse_fields *cf, int f, int x, int y) {
  if (cf->exact) return cf->full[((size_t)f * cf->h + y) * cf->w + x];
  const int jx = cf->cellx[x], jy = cf->celly[y];
  const double *c0 = cf->coarse + ((size_t)f * cf->ncy + jy) * cf->ncx;
  const double *c1 = c0 + cf->ncx;
  const real a00 = (real)c0[jx], a01 = (real)c0[jx + 1];
  const real a10 = (real)c1[jx], a11 = (real)c1[jx + 1];
  const real fy = cf->wy[y];
  const real base = RFMA(a10 - a00, fy, a00);
  const real top = a01 - a00;
  const real slope = RFMA((a11 - a10) - top, fy, top);
  return RFMA(slope, cf->wx[x], base);
}

/* ------------------------------------------------------------------------ */
/* Order statistics (std::nth_element semantics: the k-th smallest value)    */
/* ------------------------------------------------------------------------ */
#define DEFINE_SELECT(NAME, T)                                            \
  T NAME(T *v, size_t n, size_t k) {                                      \
    size_t lo = 0, hi = n - 1;                                            \
    while (hi > lo) {                                                     \
      const size_t mid = lo + (hi - lo) / 2;                              \
      T a = v[lo], b = v[mid], c = v[hi], piv;                            \
      if ((a <= b && b <= c) || (c <= b && b <= a)) piv = b;              \
      else if ((b <= a && a <= c) || (c <= a && a <= b)) piv = a;         \
      else piv = c;                                                       \
      size_t i = lo, j = hi;                                              \
      while (i <= j) {                                                    \
        while (v[i] < piv) ++i;                                           \
        while (v[j] > piv) --j;                                           \
        if (i <= j) {                                                     \
          T tmp = v[i]; v[i] = v[j]; v[j] = tmp;                          \
          ++i;                                                            \
          if (j == 0) break;                                              \
          --j;                                                            \
        }                                                                 \
      }                                                                   \
      if (k <= j) hi = j;                                                 \
      else if (k >= i) lo = i;                                            \
      else return v[k];                                                   \
    }                                                                     \
    return v[k];                                                          \
  }
DEFINE_SELECT(select_kth_f, float)
DEFINE_SELECT(select_kth_d, double)

/* Median fill of the reference (subtract.cpp:20-51): finite, unmasked samples
 * at indices 7j; element size/2 after nth_element; 0 when none. */
float fill_value_f(const float *ref, const uint8_t *mask, size_t n) {
  const size_t ns = (n + 6) / 7;
  float *s = (float *)malloc(sizeof(float) * (ns ? ns : 1));
  size_t m = 0;
  for (size_t j = 0; j < ns; ++j) {
    const float v = ref[7 * j];
    if (isfinite(v) && !(mask && mask[7 * j] != MASK_GOOD)) s[m++] = v;
  }
  float out = 0.0f;
  if (m) out = select_kth_f(s, m, m / 2);
  free(s);
  return out;
}

/* Tile interpolation (subtract.cpp:211-247): per coarse cell the bilinear value
 * is the ramp base + slope * wx[x]; identical values to cf_point. */
void cf_fill_tile(const coarse_fields *cf, int f, int bx, int by, int bw, int bh,
                  real *dst) {
  if (cf->exact) {
    for (int ly = 0; ly < bh; ++ly)
      memcpy(dst + (size_t)ly * bw, cf->full + ((size_t)f * cf->h + by + ly) * cf->w + bx,
             sizeof(real) * bw);
    return;
  }
  const int ncx = cf->ncx;
  for (int ly = 0; ly < bh; ++ly) {
    const int y = by + ly, jy = cf->celly[y];
    const real fy = cf->wy[y];
    const double *c0 = cf->coarse + ((size_t)f * cf->ncy + jy) * ncx;
    const double *c1 = c0 + ncx;
    real *d = dst + (size_t)ly * bw;
    const int xend = bx + bw;
    for (int ix = cf->cellx[bx]; ix <= ncx - 2; ++ix) {
      const int xa = cf->gx[ix] > bx ? cf->gx[ix] : bx;
      if (xa >= xend) break;
      const int xbc = (ix == ncx - 2) ? cf->w : cf->gx[ix + 1];
      const int xb = xbc < xend ? xbc : xend;
      const real a00 = (real)c0[ix], a01 = (real)c0[ix + 1];
      const real a10 = (real)c1[ix], a11 = (real)c1[ix + 1];
      const real base = RFMA(a10 - a00, fy, a00);
      const real top = a01 - a00;
      const real slope = RFMA((a11 - a10) - top, fy, top);
      for (int x = xa; x < xb; ++x) d[x - bx] = RFMA(slope, cf->wx[x], base);
      if (xb >= xend) break;
    }
  }
}

/* OpenMP team size of the oracle's parallel regions (benchmark control). */
#include <omp.h>
OR_EXPORT int or_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
