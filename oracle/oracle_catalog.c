/* Oracle restatement of delta's catalog summaries -- TEST INFRASTRUCTURE ONLY.
 *
 * Follows src/catalog.cpp:262-697: counting-sort buckets of each component's
 * pixels (in raster order, as the reference's scatter leaves them when it runs
 * on one thread), summarize_component, flag_dipole_positives, fit_psf_shape,
 * nearest_bright_star_dist, enrich_entries and build_catalog.  The multiply-
 * adds are written out where GCC fuses them in the reference's default build
 * (g++ -O3 -march=x86-64-v3), and pinned by tests/test_oracle_catalog.py
 * against the reference's own catalog.cpp compiled into oracle/_ref and run on
 * one thread.  Components above kSequentialMax pixels sum their centroid in a
 * fixed order instead (the reference's order for them depends on its thread
 * schedule): kGiantSlots strided partial sums, a butterfly within each run of
 * 64 slots, then the runs' totals in slot order; that order is pinned against
 * numpy.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void or_seed_mask(const float *score, const uint8_t *mask, long n, double threshold,
                  int polarity, uint8_t *out);
int or_label_seed(const uint8_t *seed, int w, int h, int32_t *labels);

typedef struct {
  double threshold_sigma, threshold_sigma_dipole, expected_fwhm;
  double fwhm_tolerance_lo, fwhm_tolerance_hi;
  int aperture_radius;
  int exclude_bad_pixels;
  int fit_psf_shape;
  double shape_chi2_tolerance;
  double bright_star_radius;
  int max_shape_fit;
} or_catalog_params;

/* Field order: doubles, then int32, then bytes (no interior padding). */
typedef struct {
  double x, y, peak_snr, flux, expected_n_pix, fwhm_ratio;
  double psf_chi2, psf_amplitude, psf_amplitude_err, bright_star_dist;
  int32_t peak_x, peak_y, n_pix, psf_chi2_dof;
  uint8_t fwhm_consistent, mask_flags, is_dipole, shape_consistent, near_bright_star, quality;
  uint8_t pad[2];
} or_catalog_entry;

static const double kFwhmPerSigma = 2.354820045030949;
static const double kPi = 3.14159265358979323846;
static const long kSequentialMax = 8192;
enum { kGiantSlots = 1024 };

/* src/catalog.cpp:303-325 */
static double aperture_flux(const float *d, int w, int h, double cx, double cy, int radius) {
  const int x0 = (int)lround(cx), y0 = (int)lround(cy);
  const double r2 = (double)radius * radius;
  double sum = 0.0;
  for (int dy = -radius; dy <= radius; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -radius; dx <= radius; ++dx) {
      if (dx * dx + dy * dy > r2) continue;
      const int xx = x0 + dx;
      if (xx < 0 || xx >= w) continue;
      const float v = d[(size_t)yy * w + xx];
      if (isfinite(v)) sum += (double)v;
    }
  }
  return sum;
}

/* src/catalog.cpp:335-418 over the component's pixels in raster order. */
static void summarize(const long *pix, long count, const float *score, const float *diff,
                      const uint8_t *mask, const int32_t *labels, int32_t id, int w, int h,
                      int polarity, const or_catalog_params *p, or_catalog_entry *e) {
  memset(e, 0, sizeof(*e));
  e->shape_consistent = 1;
  e->bright_star_dist = INFINITY;
  e->n_pix = (int32_t)count;
  double sum = 0.0, sx = 0.0, sy = 0.0, best = -1.0, peak = 0.0;
  long best_idx = 0;
  uint8_t flags = 0;
  /* Above kSequentialMax pixels: kGiantSlots strided partial sums (pixel k
   * into slot k mod kGiantSlots), a butterfly within each run of 64 slots
   * (l + (l ^ 1), then + (l ^ 2), ... + (l ^ 32)), then the runs' totals
   * added in order -- a fixed order for giant components, where the
   * reference's order follows its thread schedule. */
  const int giant = count > kSequentialMax;
  double ps[kGiantSlots], px[kGiantSlots], py[kGiantSlots];
  if (giant)
    for (int l = 0; l < kGiantSlots; ++l) ps[l] = px[l] = py[l] = 0.0;
  for (long k = 0; k < count; ++k) {
    const long idx = pix[k];
    const int x = (int)(idx % w), y = (int)(idx / w);
    const double v = (double)score[idx];
    const double wt = fabs(v);
    if (giant) {
      const int l = (int)(k % kGiantSlots);
      ps[l] += wt;
      px[l] = fma(wt, (double)x, px[l]);
      py[l] = fma(wt, (double)y, py[l]);
    } else {
      sum += wt;
      sx = fma(wt, (double)x, sx);
      sy = fma(wt, (double)y, sy);
    }
    if (wt > best || (wt == best && idx < best_idx)) {
      best = wt;
      best_idx = idx;
      peak = v;
      e->peak_x = x;
      e->peak_y = y;
    }
    if (mask) {
      flags |= mask[idx];
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= h) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0) continue;
          const int xx = x + dx;
          if (xx < 0 || xx >= w) continue;
          const long n = (long)yy * w + xx;
          if (labels[n] != id) flags |= mask[n];
        }
      }
    }
  }
  if (giant) {
    for (int r0 = 0; r0 < kGiantSlots; r0 += 64) {
      double *rs = ps + r0, *rx = px + r0, *ry = py + r0;
      for (int off = 1; off < 64; off <<= 1) {
        double ts[64], tx[64], ty[64];
        for (int l = 0; l < 64; ++l) {
          ts[l] = rs[l] + rs[l ^ off];
          tx[l] = rx[l] + rx[l ^ off];
          ty[l] = ry[l] + ry[l ^ off];
        }
        memcpy(rs, ts, sizeof ts), memcpy(rx, tx, sizeof tx), memcpy(ry, ty, sizeof ty);
      }
      if (r0 == 0) {
        sum = rs[0], sx = rx[0], sy = ry[0];
      } else {
        sum += rs[0], sx += rx[0], sy += ry[0];
      }
    }
  }
  e->x = sum > 0.0 ? sx / sum : (count > 0 ? (double)(pix[0] % w) : 0.0);
  e->y = sum > 0.0 ? sy / sum : (count > 0 ? (double)(pix[0] / w) : 0.0);
  e->peak_snr = peak;
  e->mask_flags = flags;
  const double sigma = p->expected_fwhm / kFwhmPerSigma;
  const double thr = polarity > 0 ? p->threshold_sigma : p->threshold_sigma_dipole;
  const double ratio = thr > 0.0 ? fabs(e->peak_snr) / thr : 0.0;
  const double lt = ratio > 0.0 ? log(ratio * ratio) : 0.0;
  e->expected_n_pix = kPi * sigma * sigma * (lt > 0.0 ? lt : 0.0);
  e->fwhm_ratio = e->expected_n_pix > 0.0 ? e->n_pix / e->expected_n_pix : INFINITY;
  e->fwhm_consistent = e->fwhm_ratio >= p->fwhm_tolerance_lo &&
                       e->fwhm_ratio <= p->fwhm_tolerance_hi;
  int radius = p->aperture_radius;
  if (radius <= 0) {
    const long r = lround(3.0 * sigma);
    radius = r > 1 ? (int)r : 1;
  }
  e->flux = aperture_flux(diff, w, h, e->x, e->y, radius);
  e->quality = (uint8_t)((e->fwhm_consistent ? 0 : 1) | (e->mask_flags ? 4 : 0));
}

/* src/catalog.cpp:465-488: weighted 1-parameter PSF fit around (cx, cy). */
static int fit_shape(const float *d, const float *var, const uint8_t *mask, int w, int h,
                     double cx, double cy, const float *psf, int ps, or_catalog_entry *e) {
  const int r = ps / 2;
  const int x0 = (int)lround(cx), y0 = (int)lround(cy);
  double swdp = 0.0, swpp = 0.0;
  int used = 0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x0 + dx;
      if (xx < 0 || xx >= w) continue;
      const size_t i = (size_t)yy * w + xx;
      if (mask && mask[i] != 0) continue;
      const float v = var[i], dv = d[i];
      if (!(v > 0.0f) || !isfinite(v) || !isfinite(dv)) continue;
      const float pv = psf[(size_t)(dy + r) * ps + (dx + r)];
      const double wgt = 1.0 / (double)v;
      swdp = fma(wgt * (double)dv, (double)pv, swdp);
      swpp = fma(wgt * (double)pv, (double)pv, swpp);
      ++used;
    }
  }
  if (used < 3 || !(swpp > 0.0)) return 0;
  const double amp = swdp / swpp;
  double chi2 = 0.0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x0 + dx;
      if (xx < 0 || xx >= w) continue;
      const size_t i = (size_t)yy * w + xx;
      if (mask && mask[i] != 0) continue;
      const float v = var[i], dv = d[i];
      if (!(v > 0.0f) || !isfinite(v) || !isfinite(dv)) continue;
      const float pv = psf[(size_t)(dy + r) * ps + (dx + r)];
      const double res = fma(-amp, (double)pv, (double)dv);
      chi2 += res * res / (double)v;
    }
  }
  e->psf_amplitude = amp;
  e->psf_amplitude_err = sqrt(1.0 / swpp);
  e->psf_chi2 = chi2;
  e->psf_chi2_dof = used - 1;
  return 1;
}

/* src/catalog.cpp:534-545 */
static double nearest_bright(double x, double y, const double *bx, const double *by, int n) {
  double best2 = INFINITY;
  for (int i = 0; i < n; ++i) {
    const double dx = x - bx[i], dy = y - by[i];
    const double d2 = fma(dx, dx, dy * dy);
    if (d2 < best2) best2 = d2;
  }
  return sqrt(best2);
}

/* Components of `labels` (1..nc) summarized in label order. */
static void summarize_all(const int32_t *labels, int nc, int w, int h, const float *score,
                          const float *diff, const uint8_t *mask, int polarity,
                          const or_catalog_params *p, or_catalog_entry *out) {
  const long n = (long)w * h;
  long *off = (long *)calloc((size_t)nc + 2, sizeof(long));
  for (long i = 0; i < n; ++i)
    if (labels[i] > 0) off[labels[i] + 1]++;
  for (int l = 1; l <= nc; ++l) off[l + 1] += off[l];
  long *cur = (long *)malloc(sizeof(long) * ((size_t)nc + 2));
  memcpy(cur, off, sizeof(long) * ((size_t)nc + 2));
  long *flat = (long *)malloc(sizeof(long) * (size_t)(off[nc + 1] + 1));
  for (long i = 0; i < n; ++i)
    if (labels[i] > 0) flat[cur[labels[i]]++] = i;
  for (int l = 1; l <= nc; ++l)
    summarize(flat + off[l], off[l + 1] - off[l], score, diff, mask, labels, l, w, h, polarity,
              p, &out[l - 1]);
  free(off), free(cur), free(flat);
}

/* src/catalog.cpp:437-458: labels touched by a seed pixel's 3x3 neighbourhood. */
static void flag_dipoles(const uint8_t *seed, const int32_t *labels, int nc, int w, int h,
                         or_catalog_entry *out) {
  uint8_t *t = (uint8_t *)calloc((size_t)nc + 1, 1);
  for (long i = 0; i < (long)w * h; ++i) {
    if (!seed[i]) continue;
    const int x = (int)(i % w), y = (int)(i / w);
    for (int dy = -1; dy <= 1; ++dy) {
      const int yy = y + dy;
      if (yy < 0 || yy >= h) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx;
        if (xx < 0 || xx >= w) continue;
        const int32_t l = labels[(long)yy * w + xx];
        if (l > 0) t[l] = 1;
      }
    }
  }
  for (int l = 1; l <= nc; ++l) {
    out[l - 1].is_dipole = t[l];
    if (t[l]) out[l - 1].quality |= 2;
  }
  free(t);
}

/* src/catalog.cpp:555-584 */
static void enrich(or_catalog_entry *e, int n, const float *diff, const uint8_t *mask, int w,
                   int h, const or_catalog_params *p, const float *var, const float *psf, int ps,
                   const double *bx, const double *by, int nb) {
  if (n == 0) return;
  if (n > (p->max_shape_fit > 0 ? p->max_shape_fit : 0)) return;
  const int do_shape = p->fit_psf_shape && psf && var && ps > 0;
  const int do_bright = bx && by && nb > 0;
  if (!do_shape && !do_bright) return;
  const double brad = p->bright_star_radius > 0.0 ? p->bright_star_radius : 3.0 * p->expected_fwhm;
  for (int i = 0; i < n; ++i) {
    if (do_shape && fit_shape(diff, var, mask, w, h, e[i].x, e[i].y, psf, ps, &e[i])) {
      e[i].shape_consistent =
          e[i].psf_chi2_dof > 0 ? (e[i].psf_chi2 / e[i].psf_chi2_dof <= p->shape_chi2_tolerance)
                                : 1;
      if (!e[i].shape_consistent) e[i].quality |= 8;
    }
    if (do_bright) {
      e[i].bright_star_dist = nearest_bright(e[i].x, e[i].y, bx, by, nb);
      e[i].near_bright_star = e[i].bright_star_dist <= brad;
      if (e[i].near_bright_star) e[i].quality |= 16;
    }
  }
}

/* src/catalog.cpp:600-697.  Writes min(n, capacity) entries; returns n. */
int or_build_catalog(const float *score, const float *diff, const uint8_t *mask, int w, int h,
                     const or_catalog_params *p, int return_negative, const float *var,
                     const float *psf, int ps, const double *bx, const double *by, int nb,
                     or_catalog_entry *out, int capacity) {
  const long n = (long)w * h;
  const uint8_t *seed_m = p->exclude_bad_pixels ? mask : NULL;
  uint8_t *pos = (uint8_t *)malloc((size_t)n + 1), *neg = (uint8_t *)malloc((size_t)n + 1);
  or_seed_mask(score, seed_m, n, p->threshold_sigma, 1, pos);
  or_seed_mask(score, seed_m, n, p->threshold_sigma_dipole, -1, neg);
  int32_t *lab = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
  const uint8_t *seed = return_negative ? neg : pos;
  const uint8_t *other = return_negative ? pos : neg;
  const int nc = or_label_seed(seed, w, h, lab);
  or_catalog_entry *e = (or_catalog_entry *)calloc((size_t)nc + 1, sizeof(or_catalog_entry));
  summarize_all(lab, nc, w, h, score, diff, mask, return_negative ? -1 : 1, p, e);
  flag_dipoles(other, lab, nc, w, h, e);
  enrich(e, nc, diff, mask, w, h, p, var, psf, ps, bx, by, nb);
  memcpy(out, e, sizeof(or_catalog_entry) * (size_t)(nc < capacity ? nc : capacity));
  free(pos), free(neg), free(lab), free(e);
  return nc;
}
