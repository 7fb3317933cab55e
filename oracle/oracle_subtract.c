/*
 * oracle_subtract.c -- full-frame difference image, propagated variance and
 * mask growth.  TEST INFRASTRUCTURE ONLY (see oracle_core.h).
 *
 * Follows src/subtract.cpp:429-966:
 *   model   D = S - b - sum_n a_n (g_ny *_y (g_nx *_x R_clean))   (:466-639)
 *           x taps ascending, then per component the y taps ascending, then
 *           D -= a_n B_n in component order; products accumulated by FMA.
 *   var     per 64x64 tile K frozen at the tile centre, K^2 box-shrunk to
 *           (1-1e-4) of its sum, flat-window shortcut or dense correlation
 *           (:641-824), then + Var(S).
 *   mask    science flags 1:1, reference flags + saturated reference dilated
 *           by the kernel radius, edge border, non-finite; exact-bad pixels get
 *           the stride-7 median of the good difference (:826-963).
 * Pinning: tests/test_oracle_subtract.py checks the model against a direct
 * numpy 2-D convolution of the factorised model (test_subtract.py:88-106), the
 * self-subtraction identity (:109-125), the constant-kernel variance closed
 * form (:128-154) and the dilation box rule (:173-215).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_core.h"

#define TILE 96

static void model_tile(const coarse_fields *cf, const gh_basis *b, const float *sci,
                       const float *ref, const uint8_t *rmask, float fill, int w,
                       int h, int bx, int by, real *diff, real *rt, real *tx,
                       real *bn, real *kx) {
  const int kh = b->radius, ks = b->ks, nmax = b->n_max;
  const int bw = (w - bx) < TILE ? (w - bx) : TILE;
  const int bh = (h - by) < TILE ? (h - by) : TILE;
  const int nrows = bh + 2 * kh, gbw = bw + 2 * kh;
  const int hy0 = by - kh, gx0 = bx - kh;
  /* gather + sanitise the haloed reference window */
  for (int r = 0; r < nrows; ++r) {
    const int gy = hy0 + r;
    real *row = rt + (size_t)r * gbw;
    for (int c = 0; c < gbw; ++c) {
      const int gx = gx0 + c;
      if (gy < 0 || gy >= h || gx < 0 || gx >= w) {
        row[c] = 0;
        continue;
      }
      const size_t i = (size_t)gy * w + gx;
      const float v = ref[i];
      const int bad = (rmask && rmask[i] != MASK_GOOD) || !isfinite(v);
      row[c] = bad ? (real)fill : (real)v;
    }
  }
  /* shared x-passes */
  for (int nx = 0; nx <= nmax; ++nx) {
    const real *k = kx + (size_t)nx * ks;
    for (int r = 0; r < nrows; ++r) {
      real *o = tx + ((size_t)nx * nrows + r) * bw;
      const real *src = rt + (size_t)r * gbw;
      for (int lx = 0; lx < bw; ++lx) o[lx] = 0;
      for (int j = 0; j < ks; ++j) {
        const real kc = k[j];
        const real *s = src + 2 * kh - j;
        for (int lx = 0; lx < bw; ++lx) o[lx] = RFMA(kc, s[lx], o[lx]);
      }
    }
  }
  /* S - b */
  real dt[TILE * TILE], ab[TILE * TILE];
  cf_fill_tile(cf, b->nc, bx, by, bw, bh, ab);
  for (int ly = 0; ly < bh; ++ly)
    for (int lx = 0; lx < bw; ++lx)
      dt[ly * TILE + lx] = (real)sci[(size_t)(by + ly) * w + bx + lx] - ab[ly * bw + lx];
  /* component y-pass + accumulate */
  for (int n = 0; n < b->nc; ++n) {
    const real *src = tx + (size_t)b->onx[n] * nrows * bw;
    const real *k = kx + (size_t)b->ony[n] * ks;
    cf_fill_tile(cf, n, bx, by, bw, bh, ab);
    for (int ly = 0; ly < bh; ++ly) {
      const int y = by + ly;
      int jmin = y + kh - h + 1;
      if (jmin < 0) jmin = 0;
      int jmax = y + kh;
      if (jmax > ks - 1) jmax = ks - 1;
      for (int lx = 0; lx < bw; ++lx) bn[lx] = 0;
      for (int j = jmin; j <= jmax; ++j) {
        const real c = k[j];
        const real *srow = src + (size_t)(y + kh - j - hy0) * bw;
        for (int lx = 0; lx < bw; ++lx) bn[lx] = RFMA(c, srow[lx], bn[lx]);
      }
      const real *ar = ab + ly * bw;
      real *dr = dt + ly * TILE;
      for (int lx = 0; lx < bw; ++lx) dr[lx] = RFMA(-ar[lx], bn[lx], dr[lx]);
    }
  }
  for (int ly = 0; ly < bh; ++ly)
    memcpy(diff + (size_t)(by + ly) * w + bx, dt + ly * TILE, sizeof(real) * bw);
}

static void variance_pass(const coarse_fields *cf, const gh_basis *b, const float *vr,
                          const float *vs, int w, int h, real *var) {
  const int nc = b->nc, ks = b->ks, rk = b->radius;
  const size_t kk = (size_t)ks * ks;
  real *phi = (real *)malloc(sizeof(real) * nc * kk);
  for (int n = 0; n < nc; ++n) gh_kernel2d(b, n, phi + n * kk);
  const size_t npix = (size_t)w * h;
  for (size_t i = 0; i < npix; ++i) var[i] = 0;
  if (vr) {
    const int bsz = 64, ibw = bsz + 2 * rk;
    const int ntx = (w + bsz - 1) / bsz, nty = (h + bsz - 1) / bsz;
#pragma omp parallel
    {
      real *k2 = (real *)malloc(sizeof(real) * kk);
      real *kf = (real *)malloc(sizeof(real) * kk);
      real *win = (real *)malloc(sizeof(real) * ibw * ibw);
      real *out = (real *)malloc(sizeof(real) * bsz * bsz);
#pragma omp for schedule(dynamic)
      for (int bi = 0; bi < ntx * nty; ++bi) {
        const int bx = (bi % ntx) * bsz, by = (bi / ntx) * bsz;
        const int bw = (w - bx) < bsz ? (w - bx) : bsz;
        const int bh = (h - by) < bsz ? (h - by) : bsz;
        int cpx = bx + bw / 2, cpy = by + bh / 2;
        if (cpx > w - 1) cpx = w - 1;
        if (cpy > h - 1) cpy = h - 1;
        for (size_t i = 0; i < kk; ++i) k2[i] = 0;
        for (int n = 0; n < nc; ++n) {
          const real an = cf_point(cf, n, cpx, cpy);
          const real *p = phi + n * kk;
          for (size_t i = 0; i < kk; ++i) k2[i] = RFMA(an, p[i], k2[i]);
        }
        double qsum = 0.0;
        for (size_t i = 0; i < kk; ++i) {
          const real v = k2[i];
          kf[kk - 1 - i] = v * v;
          qsum += (double)v * (double)v;
        }
        int hw = 0;
        {
          const double target = (1.0 - (double)1e-4f) * qsum;
          for (; hw < rk; ++hw) {
            double box = 0.0;
            for (int yy = rk - hw; yy <= rk + hw; ++yy)
              for (int xx = rk - hw; xx <= rk + hw; ++xx) box += (double)kf[yy * ks + xx];
            if (box >= target) break;
          }
        }
        const int lo = rk - hw, hi = rk + hw;
        const int iwh = bh + 2 * rk, iww = bw + 2 * rk;
        real wmin = INFINITY, wmax = -INFINITY;
        for (int iy = 0; iy < iwh; ++iy) {
          const int gy = by - rk + iy;
          for (int ix = 0; ix < iww; ++ix) {
            const int gx = bx - rk + ix;
            const real v = (gy < 0 || gy >= h || gx < 0 || gx >= w)
                               ? (real)0
                               : (real)vr[(size_t)gy * w + gx];
            win[iy * ibw + ix] = v;
            if (v < wmin) wmin = v;
            if (v > wmax) wmax = v;
          }
        }
        if (wmax > 0 && (wmax - wmin) <= (real)1e-3f * wmax) {
          const real q = (real)qsum;
          for (int oy = 0; oy < bh; ++oy)
            for (int ox = 0; ox < bw; ++ox) {
              real *vo = var + (size_t)(by + oy) * w + bx + ox;
              *vo = RFMA(q, win[(oy + rk) * ibw + rk + ox], *vo);
            }
          continue;
        }
        for (int i = 0; i < bh * bw; ++i) out[i] = 0;
        for (int ly = lo; ly <= hi; ++ly)
          for (int lx = lo; lx <= hi; ++lx) {
            const real c = kf[ly * ks + lx];
            if (c == 0) continue;
            for (int oy = 0; oy < bh; ++oy) {
              const real *src = win + (oy + ly) * ibw + lx;
              real *dst = out + oy * bw;
              for (int ox = 0; ox < bw; ++ox) dst[ox] = RFMA(c, src[ox], dst[ox]);
            }
          }
        for (int oy = 0; oy < bh; ++oy)
          for (int ox = 0; ox < bw; ++ox) var[(size_t)(by + oy) * w + bx + ox] += out[oy * bw + ox];
      }
      free(k2);
      free(kf);
      free(win);
      free(out);
    }
  }
  if (vs)
    for (size_t i = 0; i < npix; ++i) var[i] += (real)vs[i];
  free(phi);
}

static void dilate(uint8_t *m, int w, int h, int r) {
  if (r <= 0) return;
  uint8_t *tmp = (uint8_t *)malloc((size_t)w * h);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int x0 = x - r < 0 ? 0 : x - r, x1 = x + r > w - 1 ? w - 1 : x + r;
      uint8_t a = 0;
      for (int xx = x0; xx <= x1; ++xx) a |= m[(size_t)y * w + xx];
      tmp[(size_t)y * w + x] = a;
    }
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; ++y) {
    const int y0 = y - r < 0 ? 0 : y - r, y1 = y + r > h - 1 ? h - 1 : y + r;
    for (int x = 0; x < w; ++x) {
      uint8_t a = 0;
      for (int yy = y0; yy <= y1; ++yy) a |= tmp[(size_t)yy * w + x];
      m[(size_t)y * w + x] = a;
    }
  }
  free(tmp);
}

static void mask_pass(const float *sci, const float *ref, const uint8_t *sm,
                      const uint8_t *rm, double saturation, int w, int h, int r,
                      real *diff, real *var, uint8_t *out) {
  const size_t n = (size_t)w * h;
  const int has_sat = saturation > 0.0;
  const float sat = (float)saturation;
  const int grow = rm || has_sat;
  uint8_t *eb = (uint8_t *)malloc(n);
  uint8_t *g = grow ? (uint8_t *)malloc(n) : NULL;
  for (size_t i = 0; i < n; ++i) {
    uint8_t d = 0, gg = 0, e = 0;
    if (sm && sm[i]) {
      d |= sm[i];
      e = 1;
    }
    if (rm && rm[i]) {
      gg |= rm[i];
      e = 1;
    }
    if (has_sat) {
      if (ref[i] >= sat) {
        gg |= MASK_SAT;
        e = 1;
      }
      if (sci[i] >= sat) {
        d |= MASK_SAT;
        e = 1;
      }
    }
    out[i] = d;
    if (grow) g[i] = gg;
    eb[i] = e;
  }
  if (grow) dilate(g, w, h, r);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const size_t i = (size_t)y * w + x;
      if (grow) out[i] |= g[i];
      if (x < r || x >= w - r || y < r || y >= h - r) out[i] |= MASK_EDGE;
      if (!isfinite(diff[i])) {
        out[i] |= MASK_NONFINITE;
        eb[i] = 1;
      }
    }
  const size_t ns = (n + 6) / 7;
  real *s = (real *)malloc(sizeof(real) * (ns ? ns : 1));
  size_t m = 0;
  for (size_t j = 0; j < ns; ++j) {
    const size_t i = 7 * j;
    if (out[i] == MASK_GOOD && isfinite(diff[i])) s[m++] = diff[i];
  }
  real fill = 0;
#ifdef ORACLE_F64
  if (m) fill = select_kth_d(s, m, m / 2);
#else
  if (m) fill = select_kth_f(s, m, m / 2);
#endif
  free(s);
  for (size_t i = 0; i < n; ++i)
    if (out[i] != MASK_GOOD && eb[i]) {
      diff[i] = fill;
      if (var) var[i] = 0;
    }
  free(eb);
  free(g);
}

/* Full subtract (subtract.cpp:429-966).  Returns 0, or <0 on invalid input.
 * `var` must be non-NULL iff svar or rvar is given; `mask` non-NULL iff a mask
 * or a positive saturation is given. */
OR_EXPORT int or_subtract(int w, int h, const float *sci, const float *ref,
                          const float *svar, const float *rvar, const uint8_t *smask,
                          const uint8_t *rmask, const double *knots, int k,
                          const double *theta, double beta, int n_max, int radius,
                          double saturation, real *diff, real *var, uint8_t *mask) {
  gh_basis b;
  tps_basis t;
  coarse_fields cf;
  if (gh_init(&b, beta, n_max, radius)) return -1;
  if (tps_init(&t, knots, k)) {
    gh_free(&b);
    return -2;
  }
  cf_build(&cf, &t, theta, b.nc, w, h);
  const float fill = fill_value_f(ref, rmask, (size_t)w * h);
  real *kx = (real *)malloc(sizeof(real) * (b.n_max + 1) * b.ks);
  for (int i = 0; i < (b.n_max + 1) * b.ks; ++i) kx[i] = (real)b.b1d[i];
  const int ntx = (w + TILE - 1) / TILE, nty = (h + TILE - 1) / TILE;
  const int nrows = TILE + 2 * b.radius, gbw = TILE + 2 * b.radius;
#pragma omp parallel
  {
    real *rt = (real *)malloc(sizeof(real) * nrows * gbw);
    real *tx = (real *)malloc(sizeof(real) * (b.n_max + 1) * nrows * TILE);
    real bn[TILE];
#pragma omp for schedule(dynamic)
    for (int ti = 0; ti < ntx * nty; ++ti)
      model_tile(&cf, &b, sci, ref, rmask, fill, w, h, (ti % ntx) * TILE,
                 (ti / ntx) * TILE, diff, rt, tx, bn, kx);
    free(rt);
    free(tx);
  }
  if (var) variance_pass(&cf, &b, rvar, svar, w, h, var);
  if (mask) mask_pass(sci, ref, smask, rmask, saturation, w, h, b.radius, diff, var, mask);
  free(kx);
  cf_free(&cf);
  tps_free(&t);
  gh_free(&b);
  return 0;
}

/* Small exported helpers used by the oracle's own tests. */
OR_EXPORT void or_gh_basis1d(double beta, int n_max, int radius, double *out) {
  gh_basis b;
  if (gh_init(&b, beta, n_max, radius)) return;
  memcpy(out, b.b1d, sizeof(double) * (n_max + 1) * b.ks);
  gh_free(&b);
}

OR_EXPORT int or_gh_default_radius(double beta, int n_max) {
  const double scale = beta * sqrt((double)(n_max + 1));   /* src/basis.cpp:10-11 */
  const int r = (int)ceil(3.0 * scale);
  return r > 3 ? r : 3;
}

OR_EXPORT int or_tps_setup(const double *knots, int k, double *norm3, double *Z,
                           double *pen) {
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1;
  norm3[0] = t.cx;
  norm3[1] = t.cy;
  norm3[2] = t.scale;
  if (k > 3) memcpy(Z, t.Z, sizeof(double) * k * (k - 3));
  memcpy(pen, t.pen, sizeof(double) * k * k);
  tps_free(&t);
  return 0;
}

OR_EXPORT int or_tps_design(const double *knots, int k, const double *pts, int m,
                            double *D) {
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1;
  for (int i = 0; i < m; ++i) tps_design_row(&t, pts[2 * i], pts[2 * i + 1], D + (size_t)i * k);
  tps_free(&t);
  return 0;
}

OR_EXPORT double or_tps_min_spacing(const double *knots, int k) {
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1.0;
  const double s = tps_min_spacing(&t);
  tps_free(&t);
  return s;
}

/* Full-resolution fields a_n(x,y), b(x,y) through the same lattice/ramp the
 * subtract uses (evaluate_fields, subtract.cpp:320-427). out: nf x h x w. */
OR_EXPORT int or_fields(int w, int h, const double *knots, int k, const double *theta,
                        int nc, real *out) {
  tps_basis t;
  coarse_fields cf;
  if (tps_init(&t, knots, k)) return -1;
  cf_build(&cf, &t, theta, nc, w, h);
  for (int f = 0; f <= nc; ++f)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) out[((size_t)f * h + y) * w + x] = cf_point(&cf, f, x, y);
  const int ex = cf.exact;
  cf_free(&cf);
  tps_free(&t);
  return ex;
}

/* The coarse field lattice (src/subtract.cpp:113-316) for pinning the ramp:
 * info = [exact, stride, ncx, ncy]; when the caller's capacities suffice
 * (cap_lat >= nf*ncy*ncx, cap_x >= ncx, cap_y >= ncy) the FP64 node values
 * (nf x ncy x ncx) and the node coordinates are copied out.  Returns 0, or 1
 * when only info was written. */
OR_EXPORT int or_field_lattice(int w, int h, const double *knots, int k, const double *theta,
                               int nc, int *info, double *lat, long cap_lat, int *gx, int cap_x,
                               int *gy, int cap_y) {
  tps_basis t;
  coarse_fields cf;
  if (tps_init(&t, knots, k)) return -1;
  cf_build(&cf, &t, theta, nc, w, h);
  info[0] = cf.exact, info[1] = cf.stride, info[2] = cf.ncx, info[3] = cf.ncy;
  int rc = 1;
  const long need = (long)(nc + 1) * cf.ncy * cf.ncx;
  if (!cf.exact && need <= cap_lat && cf.ncx <= cap_x && cf.ncy <= cap_y) {
    memcpy(lat, cf.coarse, sizeof(double) * need);
    memcpy(gx, cf.gx, sizeof(int) * cf.ncx);
    memcpy(gy, cf.gy, sizeof(int) * cf.ncy);
    rc = 0;
  }
  cf_free(&cf);
  tps_free(&t);
  return rc;
}

/* Photometric scale sum_c (float)S_c a_c(x, y) over the frame, accumulated in
 * `real` in component order (src/fit.cpp:513-530; `out[i] += s * a[i]`
 * written with the fusion GCC applies to that statement). */
OR_EXPORT int or_photometric_scale(int w, int h, const double *knots, int k, const double *theta,
                                   const double *sums, int nc, real *out) {
  tps_basis t;
  coarse_fields cf;
  if (tps_init(&t, knots, k)) return -1;
  cf_build(&cf, &t, theta, nc, w, h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      real acc = 0;
      for (int c = 0; c < nc; ++c) {
#ifdef ORACLE_F64
        acc = fma((real)sums[c], cf_point(&cf, c, x, y), acc);
#else
        acc = fmaf((real)sums[c], cf_point(&cf, c, x, y), acc);
#endif
      }
      out[(size_t)y * w + x] = acc;
    }
  cf_free(&cf);
  tps_free(&t);
  return 0;
}

OR_EXPORT float or_fill_value(const float *ref, const uint8_t *mask, long n) {
  return fill_value_f(ref, mask, (size_t)n);
}
