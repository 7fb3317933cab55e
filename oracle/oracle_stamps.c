/* Oracle restatement of delta's stamp selection -- TEST INFRASTRUCTURE ONLY.
 *
 * Follows src/detect.cpp:84-314 (estimate_background, estimate_fwhm,
 * detect_stamps, select_stamps, select_stamps_from_catalog) with one
 * deliberate refinement: candidates of equal flux are ordered by pixel value,
 * then raster index (the reference's std::sort is unstable there, :224-225,
 * so any order is one it may produce; the value matters only where
 * v - median rounds two values to one flux).  Rounding follows the
 * reference's default build (g++ -O3 -march=x86-64-v3: AVX2/FMA, and g++
 * contracts in C++ even in ISO mode): the FWHM moment sums are written with
 * the fused multiply-adds GCC forms from them, and the detection threshold is
 * kept unfused, as GCC leaves it.  Both are pinned by discriminating tests.
 *
 * Pinned by tests/test_oracle_stamps.py against the reference's own
 * detect.cpp compiled into oracle/_ref (bit-exact stamps, background and
 * FWHM on random star fields with masks, NaNs, saturation and plateaus),
 * against the cases of the reference's tests/test_detect.py, and against the
 * closed forms (Gaussian FWHM = 2.3548 sigma; median/MAD of known samples).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int stamp_radius;
  double threshold_sigma;
  int max_stamps;
  double saturation;
  int isolation_radius;
  int border;
  int fwhm_radius;
  int match_radius;
} or_detect_params;

static const double kFwhmPerSigma = 2.354820045030949;

static int cmp_float(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}
static int cmp_double(const void* a, const void* b) {
  const double x = *(const double*)a, y = *(const double*)b;
  return (x > y) - (x < y);
}

/* Element size/2 of the finite samples (the nth_element statistic), or
 * `empty` when none is finite (src/detect.cpp:14-40).  Reorders v. */
static double median_f(float* v, size_t n, double empty) {
  size_t m = 0;
  for (size_t i = 0; i < n; ++i)
    if (isfinite(v[i])) v[m++] = v[i];
  if (m == 0) return empty;
  qsort(v, m, sizeof(float), cmp_float);
  return (double)v[m / 2];
}
static double median_d(double* v, size_t n) {
  size_t m = 0;
  for (size_t i = 0; i < n; ++i)
    if (isfinite(v[i])) v[m++] = v[i];
  if (m == 0) return NAN;
  qsort(v, m, sizeof(double), cmp_double);
  return v[m / 2];
}

/* src/detect.cpp:84-113: median and 1.4826*MAD of every step-th flat pixel
 * that is finite and unmasked, step = n / 2^18 for frames above 2^18 pixels. */
void or_estimate_background(const float* img, const uint8_t* mask, int w, int h,
                            double* median, double* sigma) {
  const size_t n = (size_t)w * h;
  const size_t target = (size_t)1 << 18;
  const size_t step = n > target ? n / target : 1;
  float* good = (float*)malloc(sizeof(float) * (n / step + 1));
  size_t m = 0;
  for (size_t i = 0; i < n; i += step) {
    const float v = img[i];
    if (!isfinite(v)) continue;
    if (mask && mask[i] != 0) continue;
    good[m++] = v;
  }
  if (m == 0) {
    free(good);
    *median = 0.0;
    *sigma = 0.0;
    return;
  }
  const double med = median_f(good, m, 0.0);
  for (size_t i = 0; i < m; ++i) good[i] = fabsf(good[i] - (float)med);
  const double mad = median_f(good, m, 0.0);
  free(good);
  *median = med;
  *sigma = 1.4826 * mad;
}

/* src/detect.cpp:115-157: flux-weighted second moments of max(v - bg, 0)
 * over the on-frame part of a (2r+1)^2 window; NaN without positive flux. */
double or_estimate_fwhm(const float* img, int w, int h, int x, int y, int radius,
                        double background) {
  double sum = 0.0, sx = 0.0, sy = 0.0;
  for (int dy = -radius; dy <= radius; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -radius; dx <= radius; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= w) continue;
      const double f = (double)img[(size_t)yy * w + xx] - background;
      const double wt = f > 0.0 ? f : 0.0;
      sum += wt;
      sx = fma(wt, (double)dx, sx);
      sy = fma(wt, (double)dy, sy);
    }
  }
  if (sum <= 0.0) return NAN;
  const double mx = sx / sum, my = sy / sum;
  double sxx = 0.0, syy = 0.0;
  for (int dy = -radius; dy <= radius; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -radius; dx <= radius; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= w) continue;
      const double f = (double)img[(size_t)yy * w + xx] - background;
      const double wt = f > 0.0 ? f : 0.0;
      sxx = fma(wt * (dx - mx), dx - mx, sxx);
      syy = fma(wt * (dy - my), dy - my, syy);
    }
  }
  const double var = 0.5 * (sxx / sum + syy / sum);
  if (var <= 0.0) return NAN;
  return kFwhmPerSigma * sqrt(var);
}

/* src/detect.cpp:55-75 */
static int footprint_ok(const float* img, const uint8_t* mask, int w, int h, int x, int y,
                        int radius, double saturation) {
  if (x - radius < 0 || x + radius >= w || y - radius < 0 || y + radius >= h) return 0;
  for (int dy = -radius; dy <= radius; ++dy)
    for (int dx = -radius; dx <= radius; ++dx) {
      const size_t i = (size_t)(y + dy) * w + (x + dx);
      const float v = img[i];
      if (!isfinite(v)) return 0;
      if (mask && mask[i] != 0) return 0;
      if (saturation > 0.0 && (double)v >= saturation) return 0;
    }
  return 1;
}

static int resolved_fwhm_radius(const or_detect_params* p) {
  if (p->fwhm_radius > 0) return p->fwhm_radius;
  return p->stamp_radius / 2 > 5 ? p->stamp_radius / 2 : 5;
}

typedef struct {
  int x, y;
  long idx;
  float v;
  double flux;
} cand_t;

/* flux descending, then value descending, then raster index ascending */
static int cmp_cand(const void* a, const void* b) {
  const cand_t* p = (const cand_t*)a;
  const cand_t* q = (const cand_t*)b;
  if (p->flux > q->flux) return -1;
  if (p->flux < q->flux) return 1;
  if (p->v > q->v) return -1;
  if (p->v < q->v) return 1;
  return (p->idx > q->idx) - (p->idx < q->idx);
}

/* src/detect.cpp:159-241.  Writes at most max(max_stamps, 0) stamps in
 * selection order; returns the count, or -1 when the resolved border is
 * below one pixel (the reference then reads outside the frame). */
int or_detect_stamps(const float* img, const uint8_t* mask, int w, int h,
                     const or_detect_params* p, int* xs, int* ys, double* flux, double* snr,
                     double* fwhm) {
  double med, sig;
  or_estimate_background(img, mask, w, h, &med, &sig);
  /* unfused, as the reference's compiler leaves it (the conditional operand
   * stops contraction there) */
  volatile double prod = p->threshold_sigma * (sig > 0.0 ? sig : 1.0);
  const double thr = med + prod;
  const int border = p->border > 0 ? p->border : p->stamp_radius;
  if (border < 1) return -1;
  const int frad = resolved_fwhm_radius(p);

  size_t cap = 1024, nc = 0;
  cand_t* c = (cand_t*)malloc(sizeof(cand_t) * cap);
  for (int y = border; y < h - border; ++y)
    for (int x = border; x < w - border; ++x) {
      const double v = (double)img[(size_t)y * w + x];
      if (v <= thr) continue;
      int peak = 1;
      for (int dy = -1; dy <= 1 && peak; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0) continue;
          if ((double)img[(size_t)(y + dy) * w + (x + dx)] > v) {
            peak = 0;
            break;
          }
        }
      if (!peak) continue;
      if (!footprint_ok(img, mask, w, h, x, y, p->stamp_radius, p->saturation)) continue;
      if (nc == cap) {
        cap *= 2;
        c = (cand_t*)realloc(c, sizeof(cand_t) * cap);
      }
      c[nc].x = x;
      c[nc].y = y;
      c[nc].idx = (long)y * w + x;
      c[nc].v = (float)v;
      c[nc].flux = v - med;
      ++nc;
    }
  qsort(c, nc, sizeof(cand_t), cmp_cand);

  const int iso = p->isolation_radius > 0 ? p->isolation_radius : 2 * p->stamp_radius;
  const long iso2 = (long)iso * iso;
  int kept = 0;
  for (size_t i = 0; i < nc; ++i) {
    if (kept >= p->max_stamps) break;
    int isolated = 1;
    for (int k = 0; k < kept; ++k) {
      const long ddx = c[i].x - xs[k], ddy = c[i].y - ys[k];
      if (ddx * ddx + ddy * ddy < iso2) {
        isolated = 0;
        break;
      }
    }
    if (!isolated) continue;
    xs[kept] = c[i].x;
    ys[kept] = c[i].y;
    flux[kept] = c[i].flux;
    snr[kept] = sig > 0.0 ? c[i].flux / sig : c[i].flux;
    fwhm[kept] = or_estimate_fwhm(img, w, h, c[i].x, c[i].y, frad, med);
    ++kept;
  }
  free(c);
  return kept;
}

/* src/detect.cpp:243-287 (detection) and :289-314 (catalog, when cat_x and
 * cat_y are given).  Output capacity: max(max_stamps, 0) for detection, n_cat
 * for a catalog.  direction: 0 = "reference", 1 = "science".  Returns the
 * stamp count, or -1 as or_detect_stamps. */
int or_select_stamps(const float* sci, const uint8_t* sci_mask, int ws, int hs,
                     const float* ref, const uint8_t* ref_mask, int wr, int hr,
                     const int* cat_x, const int* cat_y, int n_cat, const or_detect_params* p,
                     int* xs, int* ys, double* fwhm_sci, double* fwhm_ref,
                     double* med_sci, double* med_ref, int* direction) {
  int n = 0;
  if (cat_x && cat_y) {
    double ms, ss, mr, sr;
    or_estimate_background(sci, sci_mask, ws, hs, &ms, &ss);
    or_estimate_background(ref, ref_mask, wr, hr, &mr, &sr);
    const int frad = resolved_fwhm_radius(p);
    for (int i = 0; i < n_cat; ++i) {
      const int x = cat_x[i], y = cat_y[i];
      if (!footprint_ok(sci, sci_mask, ws, hs, x, y, p->stamp_radius, p->saturation)) continue;
      if (!footprint_ok(ref, ref_mask, wr, hr, x, y, p->stamp_radius, p->saturation)) continue;
      xs[n] = x;
      ys[n] = y;
      fwhm_sci[n] = or_estimate_fwhm(sci, ws, hs, x, y, frad, ms);
      fwhm_ref[n] = or_estimate_fwhm(ref, wr, hr, x, y, frad, mr);
      ++n;
    }
  } else {
    const int cap = p->max_stamps > 0 ? p->max_stamps : 0;
    int *sx = (int*)malloc(sizeof(int) * (cap + 1)), *sy = (int*)malloc(sizeof(int) * (cap + 1));
    int *rx = (int*)malloc(sizeof(int) * (cap + 1)), *ry = (int*)malloc(sizeof(int) * (cap + 1));
    double* buf = (double*)malloc(sizeof(double) * 6 * (cap + 1));
    const int ns = or_detect_stamps(sci, sci_mask, ws, hs, p, sx, sy, buf, buf + (cap + 1),
                                    buf + 2 * (cap + 1));
    const int nr = or_detect_stamps(ref, ref_mask, wr, hr, p, rx, ry, buf + 3 * (cap + 1),
                                    buf + 4 * (cap + 1), buf + 5 * (cap + 1));
    if (ns < 0 || nr < 0) {
      free(sx), free(sy), free(rx), free(ry), free(buf);
      return -1;
    }
    const double* sfw = buf + 2 * (cap + 1);
    const double* rfw = buf + 5 * (cap + 1);
    const long tol2 = (long)p->match_radius * p->match_radius;
    for (int r = 0; r < nr; ++r) {
      int best = -1;
      long best_d2 = tol2 + 1;
      for (int s = 0; s < ns; ++s) {
        const long ddx = rx[r] - sx[s], ddy = ry[r] - sy[s];
        const long d2 = ddx * ddx + ddy * ddy;
        if (d2 <= tol2 && d2 < best_d2) {
          best_d2 = d2;
          best = s;
        }
      }
      if (best < 0) continue;
      xs[n] = rx[r];
      ys[n] = ry[r];
      fwhm_sci[n] = sfw[best];
      fwhm_ref[n] = rfw[r];
      ++n;
    }
    free(sx), free(sy), free(rx), free(ry), free(buf);
  }
  double* tmp = (double*)malloc(sizeof(double) * (n + 1));
  memcpy(tmp, fwhm_sci, sizeof(double) * n);
  *med_sci = median_d(tmp, n);
  memcpy(tmp, fwhm_ref, sizeof(double) * n);
  *med_ref = median_d(tmp, n);
  free(tmp);
  *direction = (*med_ref <= *med_sci) ? 0 : 1;
  return n;
}
