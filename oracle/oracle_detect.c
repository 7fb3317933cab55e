/* CPU oracle, TEST INFRASTRUCTURE ONLY: the score image and the candidate
 * labelling of delta's detection stage.
 *
 *   or_seed_mask        src/catalog.cpp:61-80
 *   or_label            src/catalog.cpp:105-261 (8-connected labels numbered in
 *                       raster order of each component's first pixel)
 *   or_matched_filter   src/noise.cpp:535-632
 *
 * The labelling here is the single-block case of the reference's
 * block-parallel union-find: one two-pass union-find over the frame, roots
 * renumbered densely in raster scan.  The reference's phase-1 ids, block
 * offsets and final renumbering all preserve the raster order of first
 * pixels, so the labels agree for any thread count (pinned against the
 * reference's own compiled catalog.cpp and scipy.ndimage.label in tests).
 *
 * The matched filter keeps the reference's per-pixel operation order:
 * taps ly-outer, lx-inner, zero taps skipped, fused multiply-adds (the
 * reference is built with FMA contraction), score = acc / sqrt(v * sum psf^2).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void or_seed_mask(const float *score, const uint8_t *mask, long n, double threshold,
                  int polarity, uint8_t *out) {
  const float thr = (float)threshold;
  for (long i = 0; i < n; ++i) {
    const float v = score[i];
    int pass = isfinite(v) && (polarity > 0 ? v >= thr : v <= -thr);
    if (mask) pass = pass && mask[i] == 0;
    out[i] = pass ? 1 : 0;
  }
}

static int32_t uf_find(int32_t *p, int32_t i) {
  while (p[i] != i) {
    p[i] = p[p[i]];
    i = p[i];
  }
  return i;
}

/* Labels of the seed image (0 = background); returns the component count. */
int or_label_seed(const uint8_t *seed, int w, int h, int32_t *labels) {
  const long n = (long)w * h;
  int32_t *par = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  for (long i = 0; i < n; ++i) par[i] = (int32_t)i;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const long i = (long)y * w + x;
      if (!seed[i]) continue;
      /* 8-neighbours already visited in raster order: W, NW, N, NE */
      const long nb[4] = {x > 0 ? i - 1 : -1, (y > 0 && x > 0) ? i - w - 1 : -1,
                          y > 0 ? i - w : -1, (y > 0 && x + 1 < w) ? i - w + 1 : -1};
      for (int q = 0; q < 4; ++q) {
        if (nb[q] < 0 || !seed[nb[q]]) continue;
        int32_t a = uf_find(par, (int32_t)i), b = uf_find(par, (int32_t)nb[q]);
        if (a == b) continue;
        if (a < b) par[b] = a; else par[a] = b;  /* root = smallest index */
      }
    }
  int32_t next = 0;
  for (long i = 0; i < n; ++i) {
    if (!seed[i]) {
      labels[i] = 0;
      continue;
    }
    const int32_t r = uf_find(par, (int32_t)i);
    if (r == i) labels[i] = ++next;   /* first pixel of its component */
    else labels[i] = labels[r];       /* r < i: already numbered */
  }
  free(par);
  return next;
}

int or_label(const float *score, const uint8_t *mask, int w, int h, double threshold,
             int polarity, int32_t *labels) {
  const long n = (long)w * h;
  uint8_t *seed = (uint8_t *)malloc(n > 0 ? n : 1);
  or_seed_mask(score, mask, n, threshold, polarity, seed);
  const int c = or_label_seed(seed, w, h, labels);
  free(seed);
  return c;
}

void or_matched_filter(const float *img, const float *psf, int ps, const float *var, int w,
                       int h, float *out) {
  const int r = ps / 2;
  float sumpsf2 = 0.0f;
  for (int i = 0; i < ps * ps; ++i) sumpsf2 = fmaf(psf[i], psf[i], sumpsf2);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      float acc = 0.0f;
      for (int ly = 0; ly < ps; ++ly) {
        const int gy = y - r + ly;
        for (int lx = 0; lx < ps; ++lx) {
          const float p = psf[ly * ps + lx];
          if (p == 0.0f) continue;
          const int gx = x - r + lx;
          const float s = (gy < 0 || gy >= h || gx < 0 || gx >= w) ? 0.0f : img[(long)gy * w + gx];
          acc = fmaf(p, s, acc);
        }
      }
      const float v = var[(long)y * w + x];
      out[(long)y * w + x] = v > 0.0f ? acc / sqrtf(v * sumpsf2) : 0.0f;
    }
}
