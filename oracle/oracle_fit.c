/*
 * oracle_fit.c -- kernel fit: stamp gather, penalised GLS normal equations,
 * GCV lambda search with FP64 Cholesky (LDLT fallback), IRLS reweighting and
 * robust stamp clipping.  TEST INFRASTRUCTURE ONLY (see oracle_core.h).
 *
 * Follows src/fit.cpp:63-511 (cv_folds = 0 route) and src/solve.cpp:47-262:
 *   gather   per stamp the haloed, sanitised reference patch is convolved with
 *            every basis component (B_n in `real`, float in liboracle32 as
 *            delta does it); pixels are kept when unmasked in both images,
 *            finite in science, var_t + var_c > 0 (when variances exist) and
 *            every B_n finite; merged in stamp order (fit.cpp:117-229).
 *   assemble Xs = sqrt(w) [diag(B_0) D | ... | diag(B_{nc-1}) D | D],
 *            M = Xs^T Xs, rhs = Xs^T sqrt(w) d, ywy = sum w d^2 (solve.cpp:70-145).
 *   solve_at A = M + lambda P0; LLT, else LDLT with diagonal pivoting;
 *            edf = P - lambda ||L^-1 F||_F^2 with F F^T = P0 (LDLT: tr(A^-1 M));
 *            rss = ywy - 2 theta.rhs + theta^T M theta (>= 0);
 *            gcv = n rss / (n - edf)^2, +inf when n <= edf (solve.cpp:147-193).
 *   grid     first strict minimum of the GCV curve (solve.cpp:247-256).
 *   IRLS     chi2_s = sum w r^2 / n_s; w = 1 / (var_t + K^2 (x) var_c) with K
 *            frozen at the stamp centre, masked / non-positive var_c skipped;
 *            clip stamps above median + clip_sigma 1.4826 MAD worst-first down to
 *            the floor; stop when nothing is rejected after the minimum pass
 *            (fit.cpp:244-491).
 * Pinning: tests/test_oracle_fit.py -- theta against numpy's dense penalised
 * normal equations (test_solve.py:88-111), noiseless exact recovery
 * (test_solve.py:24-41), the edf limits P and 3(nc+1) (test_solve.py:59-85),
 * GCV argmin consistency (:114-133), LLT == LDLT on an SPD system, the gather
 * against B_n from the reference's compiled basis_convolve, and the fit on a
 * noiseless synthetic pair recovering the generating kernel.
 */
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_core.h"

/* ------------------------------------------------------------------------ */
/* small dense linear algebra (FP64)                                          */
/* ------------------------------------------------------------------------ */

/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (in place);
 * eigenvalues in w, eigenvectors in the columns of v. */
static void jacobi_eig(double *a, int n, double *w, double *v) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) v[i * n + j] = i == j;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[k * n + p], akq = a[k * n + q];
          a[k * n + p] = c * akp - s * akq;
          a[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[p * n + k], aqk = a[q * n + k];
          a[p * n + k] = c * apk - s * aqk;
          a[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[k * n + p], vkq = v[k * n + q];
          v[k * n + p] = c * vkp - s * vkq;
          v[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = a[i * n + i];
}

/* Half factor of the per-field penalty: Fk = V sqrt(W) over eigenvalues
 * w > k eps max|w| (solve.cpp:47-65).  Returns the kept rank r; fk is k x r. */
static int pen_half(const double *pen, int k, double *fk) {
  double *a = (double *)malloc(sizeof(double) * k * k);
  double *v = (double *)malloc(sizeof(double) * k * k);
  double *w = (double *)malloc(sizeof(double) * k);
  memcpy(a, pen, sizeof(double) * k * k);
  jacobi_eig(a, k, w, v);
  double wmax = 0.0;
  for (int i = 0; i < k; ++i) wmax = fmax(wmax, fabs(w[i]));
  const double tol = wmax * k * DBL_EPSILON;
  int r = 0;
  for (int i = 0; i < k; ++i)
    if (w[i] > tol) {
      const double sw = sqrt(w[i]);
      for (int j = 0; j < k; ++j) fk[j * k + r] = v[j * k + i] * sw;
      ++r;
    }
  free(a);
  free(v);
  free(w);
  return r;
}

/* In-place lower Cholesky of an n x n row-major matrix; 0 on success, else
 * the failing pivot + 1 (a non-positive or non-finite diagonal, as LLT). */
static int cholesky(double *a, int n) {
  for (int j = 0; j < n; ++j) {
    double d2 = 0.0;
    const double *rjj = a + j * n;
#pragma omp simd reduction(+ : d2)
    for (int k = 0; k < j; ++k) d2 += rjj[k] * rjj[k];
    double d = a[j * n + j] - d2;
    if (!(d > 0.0)) return j + 1;
    d = sqrt(d);
    a[j * n + j] = d;
    const double inv = 1.0 / d;
    for (int i = j + 1; i < n; ++i) {
      const double *ri = a + i * n, *rj = a + j * n;
      double t = 0.0;
#pragma omp simd reduction(+ : t)
      for (int k = 0; k < j; ++k) t += ri[k] * rj[k];
      a[i * n + j] = (a[i * n + j] - t) * inv;
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) a[i * n + j] = 0.0;
  return 0;
}

/* x = L^-1 x  (forward),  x = L^-T x (backward); L lower row-major */
static void fwd(const double *l, int n, double *x) {
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    for (int k = 0; k < i; ++k) s -= l[i * n + k] * x[k];
    x[i] = s / l[i * n + i];
  }
}
static void bwd(const double *l, int n, double *x) {
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < n; ++k) s -= l[k * n + i] * x[k];
    x[i] = s / l[i * n + i];
  }
}

/* LDLT with diagonal (largest |d|) pivoting, Eigen's unblocked algorithm:
 * a is overwritten with unit-lower L and D on the diagonal; perm records the
 * transposition at each step. */
static void ldlt(double *a, int n, int *perm) {
  double *tmp = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < n; ++k) {
    int big = k;
    double bv = fabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[i * n + i]) > bv) {
        bv = fabs(a[i * n + i]);
        big = i;
      }
    perm[k] = big;
    if (big != k) { /* symmetric swap of rows/cols k and big (full storage) */
      for (int j = 0; j < n; ++j) {
        const double t = a[k * n + j];
        a[k * n + j] = a[big * n + j];
        a[big * n + j] = t;
      }
      for (int j = 0; j < n; ++j) {
        const double t = a[j * n + k];
        a[j * n + k] = a[j * n + big];
        a[j * n + big] = t;
      }
    }
    if (k > 0) {
      for (int j = 0; j < k; ++j) tmp[j] = a[j * n + j] * a[k * n + j];
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += a[k * n + j] * tmp[j];
      a[k * n + k] -= s;
      for (int i = k + 1; i < n; ++i) {
        double t = 0.0;
        for (int j = 0; j < k; ++j) t += a[i * n + j] * tmp[j];
        a[i * n + k] -= t;
      }
    }
    const double akk = a[k * n + k];
    if (fabs(akk) > 0.0)
      for (int i = k + 1; i < n; ++i) a[i * n + k] /= akk;
  }
  free(tmp);
}

static void ldlt_solve(const double *a, int n, const int *perm, double *x) {
  for (int k = 0; k < n; ++k)
    if (perm[k] != k) {
      const double t = x[k];
      x[k] = x[perm[k]];
      x[perm[k]] = t;
    }
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    for (int k = 0; k < i; ++k) s -= a[i * n + k] * x[k];
    x[i] = s;
  }
  for (int i = 0; i < n; ++i) x[i] = fabs(a[i * n + i]) > DBL_MIN ? x[i] / a[i * n + i] : 0.0;
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < n; ++k) s -= a[k * n + i] * x[k];
    x[i] = s;
  }
  for (int k = n - 1; k >= 0; --k)
    if (perm[k] != k) {
      const double t = x[k];
      x[k] = x[perm[k]];
      x[perm[k]] = t;
    }
}

/* ------------------------------------------------------------------------ */
/* normal equations and the GCV search                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, nc, k, p, rank;
  double *m, *rhs, *fk; /* p x p, p, k x rank */
  double ywy;
  const double *pen;
} gls_system;

/* M = Xs^T Xs over the rows with nonzero weight, in 2048-row chunks. */
static void assemble(gls_system *s, int nrow, const double *D, const double *B,
                     const double *w, const double *d) {
  const int k = s->k, nc = s->nc, p = s->p;
  memset(s->m, 0, sizeof(double) * p * p);
  memset(s->rhs, 0, sizeof(double) * p);
  const int chunk = 2048;
  const int nchunks = (nrow + chunk - 1) / chunk;
  double ywy = 0.0;
  for (int i = 0; i < nrow; ++i) ywy += w[i] * d[i] * d[i];
  s->ywy = ywy;
#pragma omp parallel
  {
    double *xt = (double *)malloc(sizeof(double) * (size_t)p * chunk); /* p x chunk */
    double *ml = (double *)calloc((size_t)p * p, sizeof(double));
    double *rl = (double *)calloc(p, sizeof(double));
    double *ds = (double *)malloc(sizeof(double) * chunk);
#pragma omp for schedule(dynamic)
    for (int c = 0; c < nchunks; ++c) {
      const int r0 = c * chunk, nr = (nrow - r0) < chunk ? (nrow - r0) : chunk;
      for (int r = 0; r < nr; ++r) {
        const int i = r0 + r;
        const double sw = sqrt(w[i]);
        ds[r] = sw * d[i];
        for (int f = 0; f <= nc; ++f) {
          const double bs = (f < nc ? B[(size_t)i * nc + f] : 1.0) * sw;
          for (int m = 0; m < k; ++m) xt[(size_t)(f * k + m) * chunk + r] = D[(size_t)i * k + m] * bs;
        }
      }
      for (int a = 0; a < p; ++a) {
        const double *xa = xt + (size_t)a * chunk;
        double acc = 0.0;
#pragma omp simd reduction(+ : acc)
        for (int r = 0; r < nr; ++r) acc += xa[r] * ds[r];
        rl[a] += acc;
        int b = 0;
        for (; b + 3 <= a; b += 4) { /* four dot products share the xa loads */
          const double *x0 = xt + (size_t)b * chunk, *x1 = x0 + chunk, *x2 = x1 + chunk,
                       *x3 = x2 + chunk;
          double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma omp simd reduction(+ : t0, t1, t2, t3)
          for (int r = 0; r < nr; ++r) {
            t0 += xa[r] * x0[r];
            t1 += xa[r] * x1[r];
            t2 += xa[r] * x2[r];
            t3 += xa[r] * x3[r];
          }
          ml[(size_t)a * p + b] += t0;
          ml[(size_t)a * p + b + 1] += t1;
          ml[(size_t)a * p + b + 2] += t2;
          ml[(size_t)a * p + b + 3] += t3;
        }
        for (; b <= a; ++b) {
          const double *xb = xt + (size_t)b * chunk;
          double t = 0.0;
#pragma omp simd reduction(+ : t)
          for (int r = 0; r < nr; ++r) t += xa[r] * xb[r];
          ml[(size_t)a * p + b] += t;
        }
      }
    }
#pragma omp critical
    {
      for (int a = 0; a < p; ++a) {
        s->rhs[a] += rl[a];
        for (int b = 0; b <= a; ++b) s->m[(size_t)a * p + b] += ml[(size_t)a * p + b];
      }
    }
    free(xt);
    free(ml);
    free(rl);
    free(ds);
  }
  for (int a = 0; a < p; ++a)
    for (int b = a + 1; b < p; ++b) s->m[(size_t)a * p + b] = s->m[(size_t)b * p + a];
}

typedef struct {
  double lambda, gcv, edf, rss;
  int ldlt;
} gls_result;

static gls_result solve_at(const gls_system *s, double lambda, double *theta) {
  const int p = s->p, k = s->k, nc = s->nc, r = s->rank;
  double *a = (double *)malloc(sizeof(double) * (size_t)p * p);
  memcpy(a, s->m, sizeof(double) * (size_t)p * p);
  for (int f = 0; f <= nc; ++f)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) a[(size_t)(f * k + i) * p + f * k + j] += lambda * s->pen[i * k + j];
  gls_result res = {lambda, 0, 0, 0, 0};
  memcpy(theta, s->rhs, sizeof(double) * p);
  if (cholesky(a, p) == 0) {
    fwd(a, p, theta);
    bwd(a, p, theta);
    if (r > 0) {
      /* ||L^-1 F||_F^2 over the (nc+1) block columns of F */
      double *col = (double *)malloc(sizeof(double) * p);
      double sq = 0.0;
      for (int f = 0; f <= nc; ++f)
        for (int c = 0; c < r; ++c) {
          memset(col, 0, sizeof(double) * p);
          for (int i = 0; i < k; ++i) col[f * k + i] = s->fk[i * k + c];
          /* forward solve only from the block start (leading rows are zero) */
          for (int i = f * k; i < p; ++i) {
            const double *ai = a + (size_t)i * p;
            double t = 0.0;
#pragma omp simd reduction(+ : t)
            for (int q = f * k; q < i; ++q) t += ai[q] * col[q];
            col[i] = (col[i] - t) / ai[i];
            sq += col[i] * col[i];
          }
        }
      free(col);
      res.edf = (double)p - lambda * sq;
    } else {
      double tr = 0.0;
      double *col = (double *)malloc(sizeof(double) * p);
      for (int j = 0; j < p; ++j) {
        for (int i = 0; i < p; ++i) col[i] = s->m[(size_t)i * p + j];
        fwd(a, p, col);
        bwd(a, p, col);
        tr += col[j];
      }
      free(col);
      res.edf = tr;
    }
  } else {
    res.ldlt = 1;
    memcpy(a, s->m, sizeof(double) * (size_t)p * p);
    for (int f = 0; f <= nc; ++f)
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) a[(size_t)(f * k + i) * p + f * k + j] += lambda * s->pen[i * k + j];
    int *perm = (int *)malloc(sizeof(int) * p);
    ldlt(a, p, perm);
    memcpy(theta, s->rhs, sizeof(double) * p);
    ldlt_solve(a, p, perm, theta);
    double tr = 0.0;
    double *col = (double *)malloc(sizeof(double) * p);
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i < p; ++i) col[i] = s->m[(size_t)i * p + j];
      ldlt_solve(a, p, perm, col);
      tr += col[j];
    }
    free(col);
    free(perm);
    res.edf = tr;
  }
  double tr_ = 0.0, tmt = 0.0;
  for (int i = 0; i < p; ++i) tr_ += theta[i] * s->rhs[i];
  for (int i = 0; i < p; ++i) {
    double mi = 0.0;
    const double *mr = s->m + (size_t)i * p;
#pragma omp simd reduction(+ : mi)
    for (int j = 0; j < p; ++j) mi += mr[j] * theta[j];
    tmt += theta[i] * mi;
  }
  double rss = s->ywy - 2.0 * tr_ + tmt;
  if (rss < 0.0) rss = 0.0;
  res.rss = rss;
  const double den = (double)s->n - res.edf;
  res.gcv = den > 0.0 ? (double)s->n * rss / (den * den) : INFINITY;
  free(a);
  return res;
}

/* GCV over the grid; theta of the first strict minimum. */
static gls_result solve_gcv(const gls_system *s, const double *grid, int ng, double *theta,
                            double *curve) {
  const int p = s->p;
  double *th = (double *)malloc(sizeof(double) * (size_t)p * ng);
  gls_result *rs = (gls_result *)malloc(sizeof(gls_result) * ng);
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < ng; ++i) rs[i] = solve_at(s, grid[i], th + (size_t)i * p);
  int best = 0;
  double bg = INFINITY;
  for (int i = 0; i < ng; ++i) {
    curve[i] = rs[i].gcv;
    if (rs[i].gcv < bg) {
      bg = rs[i].gcv;
      best = i;
    }
  }
  memcpy(theta, th + (size_t)best * p, sizeof(double) * p);
  gls_result out = rs[best];
  free(th);
  free(rs);
  return out;
}

/* ------------------------------------------------------------------------ */
/* cross-validated lambda (src/solve.cpp:271-633)                            */
/* ------------------------------------------------------------------------ */
typedef struct {
  int G;
  double **mf, **rf, *ysq; /* per fold: p x p, p, 1 */
} cv_parts;

static void cv_alloc(cv_parts *f, int G, int p) {
  f->G = G;
  f->mf = (double **)malloc(sizeof(double *) * G);
  f->rf = (double **)malloc(sizeof(double *) * G);
  f->ysq = (double *)calloc(G, sizeof(double));
  for (int g = 0; g < G; ++g) {
    f->mf[g] = (double *)calloc((size_t)p * p, sizeof(double));
    f->rf[g] = (double *)calloc(p, sizeof(double));
  }
}

static void cv_free(cv_parts *f) {
  for (int g = 0; g < f->G; ++g) free(f->mf[g]), free(f->rf[g]);
  free(f->mf), free(f->rf), free(f->ysq);
}

/* Held-out weighted SSE of fold g at lambda (src/solve.cpp:317-333). */
static double cv_error(const cv_parts *f, const double *mt, const double *rt, int g,
                       const gls_system *s, double lambda) {
  const int p = s->p, k = s->k, nc = s->nc;
  double *a = (double *)malloc(sizeof(double) * (size_t)p * p);
  double *th = (double *)malloc(sizeof(double) * p);
  memcpy(a, mt, sizeof(double) * (size_t)p * p);
  for (int b = 0; b <= nc; ++b)
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) a[(size_t)(b * k + i) * p + b * k + j] += lambda * s->pen[i * k + j];
  memcpy(th, rt, sizeof(double) * p);
  if (cholesky(a, p) == 0) {
    fwd(a, p, th);
    bwd(a, p, th);
  } else {
    memcpy(a, mt, sizeof(double) * (size_t)p * p);
    for (int b = 0; b <= nc; ++b)
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) a[(size_t)(b * k + i) * p + b * k + j] += lambda * s->pen[i * k + j];
    int *perm = (int *)malloc(sizeof(int) * p);
    ldlt(a, p, perm);
    memcpy(th, rt, sizeof(double) * p);
    ldlt_solve(a, p, perm, th);
    free(perm);
  }
  double dr = 0.0, tmt = 0.0;
  for (int i = 0; i < p; ++i) dr += th[i] * f->rf[g][i];
  for (int i = 0; i < p; ++i) {
    double mi = 0.0;
    for (int j = 0; j < p; ++j) mi += f->mf[g][(size_t)i * p + j] * th[j];
    tmt += th[i] * mi;
  }
  free(a), free(th);
  const double e = f->ysq[g] - 2.0 * dr + tmt;
  return e > 0.0 ? e : 0.0;
}

/* cv_finish: the full system is the sum of the folds; lambda by the
 * coarse-to-fine (cold) or descending (warm) search over the CV-error curve;
 * the final fit on all rows at that lambda.  curve: +inf where unevaluated. */
static gls_result cv_finish(const cv_parts *f, gls_system *s, const double *grid, int ng,
                            int warm, double *theta, double *curve) {
  const int p = s->p, G = f->G;
  memset(s->m, 0, sizeof(double) * (size_t)p * p);
  memset(s->rhs, 0, sizeof(double) * p);
  s->ywy = 0.0;
  for (int g = 0; g < G; ++g) {
    for (size_t e = 0; e < (size_t)p * p; ++e) s->m[e] += f->mf[g][e];
    for (int i = 0; i < p; ++i) s->rhs[i] += f->rf[g][i];
    s->ywy += f->ysq[g];
  }
  double **mt = (double **)malloc(sizeof(double *) * G), **rt = (double **)malloc(sizeof(double *) * G);
  for (int g = 0; g < G; ++g) {
    mt[g] = (double *)malloc(sizeof(double) * (size_t)p * p);
    rt[g] = (double *)malloc(sizeof(double) * p);
    for (size_t e = 0; e < (size_t)p * p; ++e) mt[g][e] = s->m[e] - f->mf[g][e];
    for (int i = 0; i < p; ++i) rt[g][i] = s->rhs[i] - f->rf[g][i];
  }
  char *done = (char *)calloc(ng, 1);
  for (int l = 0; l < ng; ++l) curve[l] = INFINITY;
  int *idx = (int *)malloc(sizeof(int) * (ng + 4));
  double *err = (double *)malloc(sizeof(double) * (size_t)(ng + 4) * G);
#define CV_EVAL(m_)                                                              \
  do {                                                                           \
    const int m__ = (m_);                                                        \
    _Pragma("omp parallel for collapse(2) schedule(dynamic)")                    \
    for (int ii = 0; ii < m__; ++ii)                                             \
      for (int g = 0; g < G; ++g)                                                \
        err[(size_t)ii * G + g] = cv_error(f, mt[g], rt[g], g, s, grid[idx[ii]]); \
    for (int ii = 0; ii < m__; ++ii) {                                           \
      double t = 0.0;                                                            \
      for (int g = 0; g < G; ++g) t += err[(size_t)ii * G + g];                  \
      curve[idx[ii]] = t;                                                        \
      done[idx[ii]] = 1;                                                         \
    }                                                                            \
  } while (0)
  if (warm >= 0 && warm < ng) {
    int m = 0;
    for (int l = warm > 0 ? warm - 1 : 0; l <= (warm + 1 < ng ? warm + 1 : ng - 1); ++l) idx[m++] = l;
    CV_EVAL(m);
    for (;;) {
      int b = -1;
      double bc = INFINITY;
      for (int l = 0; l < ng; ++l)
        if (done[l] && curve[l] < bc) bc = curve[l], b = l;
      m = 0;
      if (b > 0 && !done[b - 1]) idx[m++] = b - 1;
      if (b >= 0 && b < ng - 1 && !done[b + 1]) idx[m++] = b + 1;
      if (m == 0) break;
      CV_EVAL(m);
    }
  } else {
    int m = 0;
    for (int l = 0; l < ng; l += 2) idx[m++] = l;
    if (idx[m - 1] != ng - 1) idx[m++] = ng - 1;
    const int mc = m;
    CV_EVAL(m);
    int cb = idx[0];
    for (int q = 0; q < mc; ++q)
      if (curve[idx[q]] < curve[cb]) cb = idx[q];
    m = 0;
    for (int l = cb - 2 > 0 ? cb - 2 : 0; l <= (cb + 2 < ng ? cb + 2 : ng - 1); ++l)
      if (!done[l]) idx[m++] = l;
    if (m) CV_EVAL(m);
  }
#undef CV_EVAL
  int best = 0;
  double bc = INFINITY;
  for (int l = 0; l < ng; ++l)
    if (done[l] && curve[l] < bc) bc = curve[l], best = l;
  gls_result r = solve_at(s, grid[best], theta);
  for (int g = 0; g < G; ++g) free(mt[g]), free(rt[g]);
  free(mt), free(rt), free(done), free(idx), free(err);
  return r;
}

/* Exact per-row folds (src/solve.cpp:434-562): M_g, rhs_g, ysq_g over the
 * rows of fold g. */
static void cv_parts_rows(cv_parts *f, const gls_system *s, int nrow, const double *D,
                          const double *B, const double *w, const double *d, const int *group) {
  const int k = s->k, nc = s->nc, p = s->p;
  double *Dg = (double *)malloc(sizeof(double) * (size_t)nrow * k);
  double *Bg = (double *)malloc(sizeof(double) * (size_t)nrow * (nc ? nc : 1));
  double *wg = (double *)malloc(sizeof(double) * nrow), *dg = (double *)malloc(sizeof(double) * nrow);
  for (int g = 0; g < f->G; ++g) {
    int m = 0;
    for (int i = 0; i < nrow; ++i) {
      if (group[i] != g) continue;
      memcpy(Dg + (size_t)m * k, D + (size_t)i * k, sizeof(double) * k);
      if (nc) memcpy(Bg + (size_t)m * nc, B + (size_t)i * nc, sizeof(double) * nc);
      wg[m] = w[i];
      dg[m] = d[i];
      ++m;
    }
    gls_system t = *s;
    t.m = f->mf[g];
    t.rhs = f->rf[g];
    assemble(&t, m, Dg, Bg, wg, dg);
    double y = 0.0;
    for (int i = 0; i < m; ++i) {
      const double ds = sqrt(wg[i]) * dg[i];
      y += ds * ds;
    }
    f->ysq[g] = y;
  }
  (void)p;
  free(Dg), free(Bg), free(wg), free(dg);
}

/* Per-stamp folds (src/solve.cpp:564-633): stamp moments A_s, r_s, ds2_s and
 * M_g = sum_s A_s (x) d_s d_s^T with the design frozen at each stamp's pixel
 * centroid.  stamp_id: local stamp per row; fold_of: per local stamp. */
static void cv_parts_stamped(cv_parts *f, const gls_system *s, const tps_basis *t, int nrow,
                             const int *px, const int *py, const double *B, const double *w,
                             const double *d, const int *stamp_id, const int *fold_of, int ns) {
  const int k = s->k, nc = s->nc, nf = nc + 1, p = s->p;
  double *A = (double *)calloc((size_t)ns * nf * nf, sizeof(double));
  double *R = (double *)calloc((size_t)ns * nf, sizeof(double));
  double *Y = (double *)calloc(ns, sizeof(double));
  double *cx = (double *)calloc(ns, sizeof(double)), *cy = (double *)calloc(ns, sizeof(double));
  int *cnt = (int *)calloc(ns, sizeof(int));
  double b[80];
  for (int i = 0; i < nrow; ++i) {
    const int st = stamp_id[i];
    const double wi = w[i], ti = d[i];
    for (int c = 0; c < nc; ++c) b[c] = B[(size_t)i * nc + c];
    b[nc] = 1.0;
    double *as = A + (size_t)st * nf * nf;
    for (int c = 0; c < nf; ++c) {
      const double wc = wi * b[c];
      for (int r = c; r < nf; ++r) as[r * nf + c] = fma(wc, b[r], as[r * nf + c]);
    }
    const double wt = wi * ti;
    for (int c = 0; c < nf; ++c) R[(size_t)st * nf + c] = fma(wt, b[c], R[(size_t)st * nf + c]);
    Y[st] = fma(wt, ti, Y[st]);
    cx[st] += px[i];
    cy[st] += py[i];
    ++cnt[st];
  }
  double *dsg = (double *)malloc(sizeof(double) * (size_t)ns * k);
  for (int st = 0; st < ns; ++st) {
    double *as = A + (size_t)st * nf * nf;
    for (int c = 0; c < nf; ++c)
      for (int r = 0; r < c; ++r) as[r * nf + c] = as[c * nf + r];
    tps_design_row(t, cx[st] / cnt[st], cy[st] / cnt[st], dsg + (size_t)st * k);
  }
  for (int g = 0; g < f->G; ++g) {
    double *m = f->mf[g], *r = f->rf[g];
    double y = 0.0;
    for (int st = 0; st < ns; ++st) {
      if (fold_of[st] != g) continue;
      const double *ds = dsg + (size_t)st * k, *as = A + (size_t)st * nf * nf;
      for (int c = 0; c < nf; ++c) {
        for (int cp = 0; cp < nf; ++cp) {
          const double acp = as[c * nf + cp];
          for (int a = 0; a < k; ++a)
            for (int bb = 0; bb < k; ++bb) {
              double *e = &m[(size_t)(c * k + a) * p + cp * k + bb];
              *e = fma(acp, ds[a] * ds[bb], *e);
            }
        }
        for (int a = 0; a < k; ++a) r[c * k + a] = fma(R[(size_t)st * nf + c], ds[a], r[c * k + a]);
      }
      y += Y[st];
    }
    f->ysq[g] = y;
  }
  free(A), free(R), free(Y), free(cx), free(cy), free(cnt), free(dsg);
}

static void system_init(gls_system *s, const tps_basis *t, int nc, int n) {
  s->n = n;
  s->nc = nc;
  s->k = t->k;
  s->p = (nc + 1) * t->k;
  s->pen = t->pen;
  s->m = (double *)malloc(sizeof(double) * (size_t)s->p * s->p);
  s->rhs = (double *)malloc(sizeof(double) * s->p);
  s->fk = (double *)malloc(sizeof(double) * t->k * t->k);
  s->rank = t->k > 1 ? pen_half(t->pen, t->k, s->fk) : 0;
}

static void system_free(gls_system *s) {
  free(s->m);
  free(s->rhs);
  free(s->fk);
}

/* Penalised GLS with GCV on given points (the solve_gls_gcv seam,
 * src/bindings.cpp:540-551).  scal: [lambda, gcv, edf, rss]. */
OR_EXPORT int or_solve_gls_gcv(const double *knots, int k, const double *pts, const double *target,
                               const double *weights, const double *bn, int n, int nc,
                               const double *grid, int ng, double *theta, double *scal,
                               double *curve) {
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1;
  if (ng < 1) {
    tps_free(&t);
    return -2;
  }
  double *D = (double *)malloc(sizeof(double) * (size_t)n * k);
  for (int i = 0; i < n; ++i) tps_design_row(&t, pts[2 * i], pts[2 * i + 1], D + (size_t)i * k);
  gls_system s;
  system_init(&s, &t, nc, n);
  assemble(&s, n, D, bn, weights, target);
  gls_result r = solve_gcv(&s, grid, ng, theta, curve);
  scal[0] = r.lambda;
  scal[1] = r.gcv;
  scal[2] = r.edf;
  scal[3] = r.rss;
  system_free(&s);
  free(D);
  tps_free(&t);
  return 0;
}

/* Penalised GLS with leave-group-out CV over the lambda grid on given points
 * (the solve_gls_cv seam, src/bindings.cpp:553-566; src/solve.cpp:420-455);
 * n_groups < 2 is GCV.  curve: the CV error (+inf where not evaluated). */
OR_EXPORT int or_solve_gls_cv(const double *knots, int k, const double *pts, const double *target,
                              const double *weights, const double *bn, int n, int nc,
                              const double *grid, int ng, const int *group, int n_groups,
                              int warm, double *theta, double *scal, double *curve) {
  if (n_groups < 2)
    return or_solve_gls_gcv(knots, k, pts, target, weights, bn, n, nc, grid, ng, theta, scal,
                            curve);
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1;
  if (ng < 1) {
    tps_free(&t);
    return -2;
  }
  for (int i = 0; i < n; ++i)
    if (group[i] < 0 || group[i] >= n_groups) {
      tps_free(&t);
      return -3;
    }
  double *D = (double *)malloc(sizeof(double) * (size_t)n * k);
  for (int i = 0; i < n; ++i) tps_design_row(&t, pts[2 * i], pts[2 * i + 1], D + (size_t)i * k);
  gls_system s;
  system_init(&s, &t, nc, n);
  cv_parts f;
  cv_alloc(&f, n_groups, s.p);
  cv_parts_rows(&f, &s, n, D, bn, weights, target, group);
  gls_result r = cv_finish(&f, &s, grid, ng, warm, theta, curve);
  scal[0] = r.lambda;
  scal[1] = r.gcv;
  scal[2] = r.edf;
  scal[3] = r.rss;
  cv_free(&f);
  system_free(&s);
  free(D);
  tps_free(&t);
  return 0;
}

/* Normal equations only (for checking the device assembly). */
OR_EXPORT int or_assemble(const double *knots, int k, const double *pts, const double *target,
                          const double *weights, const double *bn, int n, int nc, double *m,
                          double *rhs, double *ywy) {
  tps_basis t;
  if (tps_init(&t, knots, k)) return -1;
  double *D = (double *)malloc(sizeof(double) * (size_t)n * k);
  for (int i = 0; i < n; ++i) tps_design_row(&t, pts[2 * i], pts[2 * i + 1], D + (size_t)i * k);
  gls_system s;
  system_init(&s, &t, nc, n);
  assemble(&s, n, D, bn, weights, target);
  memcpy(m, s.m, sizeof(double) * (size_t)s.p * s.p);
  memcpy(rhs, s.rhs, sizeof(double) * s.p);
  *ywy = s.ywy;
  system_free(&s);
  free(D);
  tps_free(&t);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* stamp gather (fit.cpp:117-229)                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int *px, *py, *stamp;
  double *target, *vt, *vc, *bn; /* bn: n x nc */
  int ns;                        /* contributing stamps */
  int *scx, *scy;
} gather_out;

static void conv_patch(const real *in, int pw, int ph, const gh_basis *b, real *tx, real *out) {
  /* basis_convolve on a zero-padded patch: shared x-pass per nx, y-pass per
   * component, taps ascending (convolve.cpp:24-113). */
  const int ks = b->ks, hh = b->radius;
  for (int nx = 0; nx <= b->n_max; ++nx) {
    real kx[128];
    for (int j = 0; j < ks; ++j) kx[j] = (real)b->b1d[nx * ks + j];
    real *o = tx + (size_t)nx * pw * ph;
    for (int y = 0; y < ph; ++y)
      for (int x = 0; x < pw; ++x) {
        real acc = 0;
        for (int j = 0; j < ks; ++j) {
          const int sx = x + hh - j;
          if (sx >= 0 && sx < pw) acc = RFMA(kx[j], in[y * pw + sx], acc);
        }
        o[y * pw + x] = acc;
      }
  }
  for (int c = 0; c < b->nc; ++c) {
    real ky[128];
    for (int j = 0; j < ks; ++j) ky[j] = (real)b->b1d[b->ony[c] * ks + j];
    const real *t = tx + (size_t)b->onx[c] * pw * ph;
    real *o = out + (size_t)c * pw * ph;
    for (int y = 0; y < ph; ++y)
      for (int x = 0; x < pw; ++x) {
        real acc = 0;
        for (int j = 0; j < ks; ++j) {
          const int sy = y + hh - j;
          if (sy >= 0 && sy < ph) acc = RFMA(ky[j], t[sy * pw + x], acc);
        }
        o[y * pw + x] = acc;
      }
  }
}

/* Whole-image basis_convolve (src/convolve.cpp:90-113, binding
 * src/bindings.cpp:209-227): out[c][y][x], zero padding in both passes. */
OR_EXPORT int or_basis_convolve(const float *img, int w, int h, double beta, int n_max, int radius,
                                real *out) {
  gh_basis b;
  if (gh_init(&b, beta, n_max, radius)) return -1;
  const size_t np = (size_t)w * h;
  real *in = (real *)malloc(sizeof(real) * np);
  real *tx = (real *)malloc(sizeof(real) * np * (size_t)(n_max + 1));
  if (!in || !tx) {
    free(in);
    free(tx);
    gh_free(&b);
    return -2;
  }
  for (size_t i = 0; i < np; ++i) in[i] = (real)img[i];
  conv_patch(in, w, h, &b, tx, out);
  free(in);
  free(tx);
  gh_free(&b);
  return 0;
}

static void gather(int w, int h, const float *sci, const float *ref, const float *svar,
                   const float *rvar, const uint8_t *smask, const uint8_t *rmask,
                   const int *sx, const int *sy, int ns, int sr, const gh_basis *b,
                   gather_out *g) {
  const int nc = b->nc, r = b->radius;
  const float fill = fill_value_f(ref, rmask, (size_t)w * h);
  const int have_var = svar || rvar;
  const int cap = ns * (2 * sr + 1) * (2 * sr + 1);
  memset(g, 0, sizeof(*g));
  g->px = (int *)malloc(sizeof(int) * (cap ? cap : 1));
  g->py = (int *)malloc(sizeof(int) * (cap ? cap : 1));
  g->stamp = (int *)malloc(sizeof(int) * (cap ? cap : 1));
  g->target = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  g->vt = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  g->vc = (double *)malloc(sizeof(double) * (cap ? cap : 1));
  g->bn = (double *)malloc(sizeof(double) * (size_t)(cap ? cap : 1) * nc);
  g->scx = (int *)malloc(sizeof(int) * (ns ? ns : 1));
  g->scy = (int *)malloc(sizeof(int) * (ns ? ns : 1));
  /* per-stamp work in parallel, merged serially in stamp order */
  int *cnt = (int *)calloc(ns ? ns : 1, sizeof(int));
  int **idx = (int **)calloc(ns ? ns : 1, sizeof(int *));
  double **bnv = (double **)calloc(ns ? ns : 1, sizeof(double *));
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < ns; ++s) {
    const int cx = sx[s], cy = sy[s];
    const int x0 = cx - sr < 0 ? 0 : cx - sr, x1 = cx + sr > w - 1 ? w - 1 : cx + sr;
    const int y0 = cy - sr < 0 ? 0 : cy - sr, y1 = cy + sr > h - 1 ? h - 1 : cy + sr;
    if (x0 > x1 || y0 > y1) continue;
    const int ex0 = x0 - r < 0 ? 0 : x0 - r, ex1 = x1 + r > w - 1 ? w - 1 : x1 + r;
    const int ey0 = y0 - r < 0 ? 0 : y0 - r, ey1 = y1 + r > h - 1 ? h - 1 : y1 + r;
    const int pw = ex1 - ex0 + 1, ph = ey1 - ey0 + 1;
    real *patch = (real *)malloc(sizeof(real) * pw * ph);
    real *tx = (real *)malloc(sizeof(real) * (size_t)(b->n_max + 1) * pw * ph);
    real *bn = (real *)malloc(sizeof(real) * (size_t)nc * pw * ph);
    for (int yy = 0; yy < ph; ++yy)
      for (int xx = 0; xx < pw; ++xx) {
        const size_t i = (size_t)(ey0 + yy) * w + ex0 + xx;
        const float v = ref[i];
        const int bad = (rmask && rmask[i] != MASK_GOOD) || !isfinite(v);
        patch[yy * pw + xx] = bad ? (real)fill : (real)v;
      }
    conv_patch(patch, pw, ph, b, tx, bn);
    const int npx = (x1 - x0 + 1) * (y1 - y0 + 1);
    idx[s] = (int *)malloc(sizeof(int) * npx);
    bnv[s] = (double *)malloc(sizeof(double) * (size_t)npx * nc);
    int m = 0;
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) {
        const size_t i = (size_t)y * w + x;
        const size_t li = (size_t)(y - ey0) * pw + (x - ex0);
        if (smask && smask[i] != MASK_GOOD) continue;
        if (rmask && rmask[i] != MASK_GOOD) continue;
        if (!isfinite(sci[i])) continue;
        const double vt = svar ? (double)svar[i] : 0.0;
        const double vc = rvar ? (double)rvar[i] : 0.0;
        if (have_var && !(vt + vc > 0.0)) continue;
        int fin = 1;
        for (int c = 0; c < nc; ++c)
          if (!isfinite(bn[(size_t)c * pw * ph + li])) {
            fin = 0;
            break;
          }
        if (!fin) continue;
        idx[s][m] = (int)i;
        for (int c = 0; c < nc; ++c) bnv[s][(size_t)m * nc + c] = (double)bn[(size_t)c * pw * ph + li];
        ++m;
      }
    cnt[s] = m;
    free(patch);
    free(tx);
    free(bn);
  }
  int n = 0, nsu = 0;
  for (int s = 0; s < ns; ++s) {
    if (cnt[s] > 0) {
      g->scx[nsu] = sx[s];
      g->scy[nsu] = sy[s];
      for (int m = 0; m < cnt[s]; ++m) {
        const int i = idx[s][m];
        g->px[n] = i % w;
        g->py[n] = i / w;
        g->stamp[n] = nsu;
        g->target[n] = (double)sci[i];
        g->vt[n] = svar ? (double)svar[i] : 0.0;
        g->vc[n] = rvar ? (double)rvar[i] : 0.0;
        memcpy(g->bn + (size_t)n * nc, bnv[s] + (size_t)m * nc, sizeof(double) * nc);
        ++n;
      }
      ++nsu;
    }
    free(idx[s]);
    free(bnv[s]);
  }
  free(cnt);
  free(idx);
  free(bnv);
  g->n = n;
  g->ns = nsu;
}

static void gather_free(gather_out *g) {
  free(g->px);
  free(g->py);
  free(g->stamp);
  free(g->target);
  free(g->vt);
  free(g->vc);
  free(g->bn);
  free(g->scx);
  free(g->scy);
}

/* Gathered stamp pixels for inspection: returns n; arrays sized by caller
 * (capacity n_stamps * (2 sr + 1)^2). */
OR_EXPORT int or_gather(int w, int h, const float *sci, const float *ref, const float *svar,
                        const float *rvar, const uint8_t *smask, const uint8_t *rmask,
                        const int *sx, const int *sy, int ns, int sr, double beta, int n_max,
                        int radius, int *px, int *py, int *stamp, double *target, double *bn) {
  gh_basis b;
  if (gh_init(&b, beta, n_max, radius)) return -1;
  gather_out g;
  gather(w, h, sci, ref, svar, rvar, smask, rmask, sx, sy, ns, sr, &b, &g);
  memcpy(px, g.px, sizeof(int) * g.n);
  memcpy(py, g.py, sizeof(int) * g.n);
  memcpy(stamp, g.stamp, sizeof(int) * g.n);
  memcpy(target, g.target, sizeof(double) * g.n);
  memcpy(bn, g.bn, sizeof(double) * (size_t)g.n * b.nc);
  const int n = g.n;
  gather_free(&g);
  gh_free(&b);
  return n;
}

/* ------------------------------------------------------------------------ */
/* the fit (fit.cpp:63-511, cv_folds = 0)                                     */
/* ------------------------------------------------------------------------ */
static int cmp_dbl(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double median_copy(const double *v, int n) {
  if (n == 0) return NAN;
  double *t = (double *)malloc(sizeof(double) * n);
  memcpy(t, v, sizeof(double) * n);
  const double m = select_kth_d(t, n, n / 2);
  free(t);
  return m;
}

/* IRLS weights of the rows (src/fit.cpp:38-59,411-446): with the fields a_s
 * of each active stamp taken at its centre (scx, scy), K_s = sum_n a_n phi_n in
 * FP32, elementwise squared; a row at (x, y) of stamp s gets
 * 1 / (vt + sum_{du,dv} K_s^2[dv+r][du+r] var_c(x-du, y-dv)) over in-frame,
 * unmasked, positive finite var_c, accumulated in FP64 in (dv, du) order; a row
 * whose total is not positive keeps its weight.  phi holds the nc float
 * footprints ks x ks. */
static void irls_weights(int w, int h, const float *rvar, const uint8_t *rmask,
                         const tps_basis *t, const double *theta, int nc, const real *phi,
                         int ks, int S, const int *scx, const int *scy, const uint8_t *active,
                         int npix, const int *px, const int *py, const int *stamp,
                         const double *vt, double *wts) {
  const int k = t->k, rr = ks / 2;
  const size_t kk = (size_t)ks * ks;
  real **k2s = (real **)calloc(S ? S : 1, sizeof(real *));
  for (int s = 0; s < S; ++s) {
    if (!active[s]) continue;
    double row[256], a[80];
    tps_design_row(t, scx[s], scy[s], row);
    for (int fi = 0; fi <= nc; ++fi) {
      double v = 0.0;
      for (int m = 0; m < k; ++m) v += row[m] * theta[fi * k + m];
      a[fi] = v;
    }
    real *k2 = (real *)calloc(kk, sizeof(real));
    for (int c = 0; c < nc; ++c) {
      const real an = (real)a[c];
      for (size_t u = 0; u < kk; ++u) k2[u] = RFMA(an, phi[c * kk + u], k2[u]);
    }
    for (size_t u = 0; u < kk; ++u) k2[u] = k2[u] * k2[u];
    k2s[s] = k2;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int j = 0; j < npix; ++j) {
    const int s = stamp[j];
    if (!k2s[s]) continue;
    const int x = px[j], y = py[j];
    const real *k2 = k2s[s];
    double vc = 0.0;
    for (int dv = -rr; dv <= rr; ++dv) {
      const int yy = y - dv;
      if (yy < 0 || yy >= h) continue;
      for (int du = -rr; du <= rr; ++du) {
        const int xx = x - du;
        if (xx < 0 || xx >= w) continue;
        const size_t q = (size_t)yy * w + xx;
        if (rmask && rmask[q] != MASK_GOOD) continue;
        const float v = rvar[q];
        if (!(v > 0.0f) || !isfinite(v)) continue;
        vc = fma((double)k2[(dv + rr) * ks + du + rr], (double)v, vc);
      }
    }
    const double rv = vt[j] + vc;
    if (rv > 0.0) wts[j] = 1.0 / rv;
  }
  for (int s = 0; s < S; ++s) free(k2s[s]);
  free(k2s);
}

/* The robust per-stamp clip of one IRLS pass (src/fit.cpp:456-483): over the
 * stamps that are active (accepted with pixels), thr = median + clip_sigma *
 * 1.4826 * MAD (upper medians, element n/2); candidates with chi2 > thr,
 * worst first (ties by index), are rejected while more than floor_st active
 * stamps remain.  Nothing is clipped when at most floor_st are active.  Writes
 * the rejected stamp indices in order; returns their count. */
static int clip_stamps(const double *chi2, const uint8_t *active, int S, int floor_st,
                       double clip_sigma, int *rej) {
  double *act = (double *)malloc(sizeof(double) * (S ? S : 1));
  int na = 0, nrej = 0;
  for (int s = 0; s < S; ++s)
    if (active[s]) act[na++] = chi2[s];
  if (na > floor_st) {
    const double med = median_copy(act, na);
    for (int q = 0; q < na; ++q) act[q] = fabs(act[q] - med);
    const double thr = med + clip_sigma * 1.4826 * median_copy(act, na);
    int nca = 0;
    int *cand = (int *)malloc(sizeof(int) * S);
    for (int s = 0; s < S; ++s)
      if (active[s] && chi2[s] > thr) cand[nca++] = s;
    /* worst first (stable for equal chi2 by index) */
    for (int a1 = 1; a1 < nca; ++a1) {
      const int v = cand[a1];
      int j = a1 - 1;
      while (j >= 0 && chi2[cand[j]] < chi2[v]) {
        cand[j + 1] = cand[j];
        --j;
      }
      cand[j + 1] = v;
    }
    for (int q = 0; q < nca; ++q) {
      if (na - nrej <= floor_st) break;
      rej[nrej++] = cand[q];
    }
    free(cand);
  }
  free(act);
  return nrej;
}

/* Exported for pinning: irls_weights on caller rows (weights updated in place). */
OR_EXPORT int or_irls_weights(int w, int h, const float *rvar, const uint8_t *rmask,
                              const double *knots, int k, const double *theta, double beta,
                              int n_max, int radius, int S, const int *scx, const int *scy,
                              const uint8_t *active, int npix, const int *px, const int *py,
                              const int *stamp, const double *vt, double *wts) {
  gh_basis b;
  tps_basis t;
  if (gh_init(&b, beta, n_max, radius)) return -1;
  if (tps_init(&t, knots, k)) {
    gh_free(&b);
    return -1;
  }
  const size_t kk = (size_t)b.ks * b.ks;
  real *phi = (real *)malloc(sizeof(real) * b.nc * kk);
  for (int c = 0; c < b.nc; ++c) gh_kernel2d(&b, c, phi + c * kk);
  irls_weights(w, h, rvar, rmask, &t, theta, b.nc, phi, b.ks, S, scx, scy, active, npix, px, py,
               stamp, vt, wts);
  free(phi);
  tps_free(&t);
  gh_free(&b);
  return 0;
}

/* Exported for pinning: clip_stamps. */
OR_EXPORT int or_clip_stamps(const double *chi2, const uint8_t *active, int S, int floor_st,
                             double clip_sigma, int *rej) {
  return clip_stamps(chi2, active, S, floor_st, clip_sigma, rej);
}

/* scal: [lambda, gcv, edf, rss, reduced_chi2]; ints: [n_pixels, n_used,
 * n_total, n_rejected, passes]; per-stamp arrays have capacity ns.
 * Returns 0, -1 bad arguments, -3 no usable stamp pixels, -4 all clipped. */
OR_EXPORT int or_fit_kernel(int w, int h, const float *sci, const float *ref, const float *svar,
                            const float *rvar, const uint8_t *smask, const uint8_t *rmask,
                            const double *knots, int k, const int *sx, const int *sy, int ns,
                            int sr, double beta, int n_max, int radius, const double *grid,
                            int ng, double clip_sigma, int clip_iterations, int min_stamps,
                            int cv_folds, int stamp_approx, double *theta, double *scal,
                            double *curve, int *ints, int *st_x, int *st_y, double *st_chi2,
                            uint8_t *st_acc) {
  gh_basis b;
  tps_basis t;
  if (gh_init(&b, beta, n_max, radius)) return -1;
  if (tps_init(&t, knots, k)) {
    gh_free(&b);
    return -1;
  }
  if (ng < 1) {
    gh_free(&b);
    tps_free(&t);
    return -1;
  }
  const int nc = b.nc, ks = b.ks;
  gather_out g;
  gather(w, h, sci, ref, svar, rvar, smask, rmask, sx, sy, ns, sr, &b, &g);
  if (g.n == 0) {
    gather_free(&g);
    gh_free(&b);
    tps_free(&t);
    return -3;
  }
  const int N = g.n, S = g.ns;
  const int have_var = svar || rvar, have_cvar = rvar != NULL;
  double *D = (double *)malloc(sizeof(double) * (size_t)N * k);
  for (int i = 0; i < N; ++i) tps_design_row(&t, g.px[i], g.py[i], D + (size_t)i * k);
  double *wts = (double *)malloc(sizeof(double) * N);
  for (int i = 0; i < N; ++i) wts[i] = have_var ? 1.0 / (g.vt[i] + g.vc[i]) : 1.0;
  real *phi = NULL;
  const size_t kk = (size_t)ks * ks;
  if (have_cvar) {
    phi = (real *)malloc(sizeof(real) * nc * kk);
    for (int c = 0; c < nc; ++c) gh_kernel2d(&b, c, phi + c * kk);
  }
  uint8_t *acc = (uint8_t *)malloc(S);
  memset(acc, 1, S);
  double *chi2 = (double *)malloc(sizeof(double) * S);
  for (int s = 0; s < S; ++s) chi2[s] = NAN;
  const int clip_passes = (clip_sigma > 0.0 && clip_iterations > 0) ? clip_iterations : 0;
  const int do_iterate = clip_iterations > 0;
  const int min_pass = (do_iterate && have_var) ? 1 : 0;
  const int hard_max = clip_passes > min_pass ? clip_passes : min_pass;
  int floor_st = min_stamps < S ? min_stamps : S;
  if (floor_st < 1) floor_st = 1;
  /* active-row buffers */
  double *Da = (double *)malloc(sizeof(double) * (size_t)N * k);
  double *Ba = (double *)malloc(sizeof(double) * (size_t)N * nc);
  double *wa = (double *)malloc(sizeof(double) * N);
  double *da = (double *)malloc(sizeof(double) * N);
  int *rows = (int *)malloc(sizeof(int) * N);
  gls_system sys;
  system_init(&sys, &t, nc, 0);
  gls_result gr = {0};
  double red = NAN, design_bytes = 0.0;
  int npix = 0, pass = 0, warm = -1;
  int rc = 0;
  for (int iter = 0;; ++iter) {
    npix = 0;
    for (int i = 0; i < N; ++i)
      if (acc[g.stamp[i]]) rows[npix++] = i;
    if (npix == 0) {
      rc = -4;
      break;
    }
    for (int j = 0; j < npix; ++j) {
      const int i = rows[j];
      memcpy(Da + (size_t)j * k, D + (size_t)i * k, sizeof(double) * k);
      memcpy(Ba + (size_t)j * nc, g.bn + (size_t)i * nc, sizeof(double) * nc);
      wa[j] = wts[i];
      da[j] = g.target[i];
    }
    sys.n = npix;
    if (cv_folds > 1) {
      /* src/fit.cpp:313-383: the per-stamp build when the knots are >= 16
       * stamp widths apart (or forced), else the exact per-row folds; folds
       * are the contributing stamp's index mod cv_folds */
      const double spacing = tps_min_spacing(&t);
      int stamped = spacing > 0.0 && spacing >= 16.0 * (2.0 * sr + 1.0);
      if (stamp_approx >= 0) stamped = stamp_approx != 0;
      cv_parts f;
      cv_alloc(&f, cv_folds, sys.p);
      if (stamped) {
        int *g2l = (int *)malloc(sizeof(int) * (S ? S : 1));
        for (int q = 0; q < S; ++q) g2l[q] = -1;
        int *sid = (int *)malloc(sizeof(int) * npix), *fold_of = (int *)malloc(sizeof(int) * (S ? S : 1));
        int *pxa = (int *)malloc(sizeof(int) * npix), *pya = (int *)malloc(sizeof(int) * npix);
        int nsl = 0;
        for (int j = 0; j < npix; ++j) {
          const int gid = g.stamp[rows[j]];
          if (g2l[gid] < 0) {
            g2l[gid] = nsl;
            fold_of[nsl++] = gid % cv_folds;
          }
          sid[j] = g2l[gid];
          pxa[j] = g.px[rows[j]];
          pya[j] = g.py[rows[j]];
        }
        cv_parts_stamped(&f, &sys, &t, npix, pxa, pya, Ba, wa, da, sid, fold_of, nsl);
        free(g2l), free(sid), free(fold_of), free(pxa), free(pya);
        design_bytes = 0.0;
      } else {
        int *grp = (int *)malloc(sizeof(int) * npix);
        for (int j = 0; j < npix; ++j) grp[j] = g.stamp[rows[j]] % cv_folds;
        cv_parts_rows(&f, &sys, npix, Da, Ba, wa, da, grp);
        free(grp);
        design_bytes = (double)npix * sys.p * sizeof(double);
      }
      gr = cv_finish(&f, &sys, grid, ng, warm, theta, curve);
      cv_free(&f);
      warm = 0;
      for (int l = 1; l < ng; ++l)
        if (curve[l] < curve[warm]) warm = l;
    } else {
      assemble(&sys, npix, Da, Ba, wa, da);
      gr = solve_gcv(&sys, grid, ng, theta, curve);
    }
    ++pass;
    /* per-stamp chi2 under this solution */
    double *sum = (double *)calloc(S, sizeof(double));
    int *cnt = (int *)calloc(S, sizeof(int));
    for (int j = 0; j < npix; ++j) {
      const double *dr = Da + (size_t)j * k;
      double f[80];
      for (int fi = 0; fi <= nc; ++fi) {
        double v = 0.0;
        for (int m = 0; m < k; ++m) v += dr[m] * theta[fi * k + m];
        f[fi] = v;
      }
      double model = f[nc];
      for (int c = 0; c < nc; ++c) model += f[c] * Ba[(size_t)j * nc + c];
      const double res = da[j] - model;
      const int s = g.stamp[rows[j]];
      sum[s] += wa[j] * res * res;
      ++cnt[s];
    }
    for (int s = 0; s < S; ++s) chi2[s] = cnt[s] > 0 ? sum[s] / cnt[s] : NAN;
    const double den = (double)npix - gr.edf;
    red = den > 0.0 ? gr.rss / den : NAN;
    /* IRLS weights: 1 / (var_t + K^2 (x) var_c), K frozen at the stamp centre */
    if (have_var) {
      if (have_cvar) {
        uint8_t *act = (uint8_t *)malloc(S ? S : 1);
        int *pxa = (int *)malloc(sizeof(int) * npix), *pya = (int *)malloc(sizeof(int) * npix);
        int *sta = (int *)malloc(sizeof(int) * npix);
        double *vta = (double *)malloc(sizeof(double) * npix), *wta = (double *)malloc(sizeof(double) * npix);
        for (int s = 0; s < S; ++s) act[s] = acc[s] && cnt[s] > 0;
        for (int j = 0; j < npix; ++j) {
          const int i = rows[j];
          pxa[j] = g.px[i], pya[j] = g.py[i], sta[j] = g.stamp[i];
          vta[j] = g.vt[i], wta[j] = wts[i];
        }
        irls_weights(w, h, rvar, rmask, &t, theta, nc, phi, ks, S, g.scx, g.scy, act, npix, pxa,
                     pya, sta, vta, wta);
        for (int j = 0; j < npix; ++j) wts[rows[j]] = wta[j];
        free(act), free(pxa), free(pya), free(sta), free(vta), free(wta);
      } else {
        for (int j = 0; j < npix; ++j) {
          const int i = rows[j];
          if (g.vt[i] > 0.0) wts[i] = 1.0 / g.vt[i];
        }
      }
    }
    /* robust clip */
    int nrej = 0;
    int *rej = (int *)malloc(sizeof(int) * (S ? S : 1));
    if (clip_passes > 0 && iter < clip_passes) {
      uint8_t *act = (uint8_t *)malloc(S ? S : 1);
      for (int s = 0; s < S; ++s) act[s] = acc[s] && cnt[s] > 0;
      nrej = clip_stamps(chi2, act, S, floor_st, clip_sigma, rej);
      free(act);
    }
    free(sum);
    free(cnt);
    const int stop = (nrej == 0 && iter >= min_pass) || iter >= hard_max;
    if (!stop)
      for (int q = 0; q < nrej; ++q) acc[rej[q]] = 0;
    free(rej);
    if (stop) break;
  }
  (void)cmp_dbl;
  if (rc == 0) {
    int used = 0;
    for (int s = 0; s < S; ++s)
      if (acc[s] && isfinite(chi2[s])) ++used;
    scal[0] = gr.lambda;
    scal[1] = gr.gcv;
    scal[2] = gr.edf;
    scal[3] = gr.rss;
    scal[4] = red;
    scal[5] = design_bytes;
    ints[0] = npix;
    ints[1] = used;
    ints[2] = S;
    ints[3] = S - used;
    ints[4] = pass;
    for (int s = 0; s < S; ++s) {
      st_x[s] = g.scx[s];
      st_y[s] = g.scy[s];
      st_chi2[s] = chi2[s];
      st_acc[s] = acc[s];
    }
  }
  system_free(&sys);
  free(Da);
  free(Ba);
  free(wa);
  free(da);
  free(rows);
  free(acc);
  free(chi2);
  free(phi);
  free(wts);
  free(D);
  gather_free(&g);
  tps_free(&t);
  gh_free(&b);
  return rc;
}
