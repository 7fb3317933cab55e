/*
 * oracle_core.h -- shared declarations of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of delta's
 * PSF-matching algorithm.  It is built twice from the same sources:
 *   liboracle32.so  image arithmetic in float, as delta's own CPU path does it
 *                   (FMA contraction made explicit with fmaf);
 *   liboracle64.so  the same algorithm with every image quantity in double,
 *                   the accuracy reference for the GPU path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product (delta_amd/, libdeltagpu.so) never links or calls it.
 */
#ifndef DELTA_ORACLE_CORE_H
#define DELTA_ORACLE_CORE_H

#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifdef ORACLE_F64
typedef double real;
#define RFMA fma
#else
typedef float real;
#define RFMA fmaf
#endif

#define OR_EXPORT __attribute__((visibility("default")))

/* Mask bits (include/delta/image.hpp:14-22). */
enum { MASK_GOOD = 0, MASK_BAD = 1, MASK_SAT = 2, MASK_COSMIC = 4,
       MASK_NONFINITE = 8, MASK_EDGE = 16, MASK_USER = 32 };

/* ---- Gauss-Hermite basis (src/basis.cpp) ---- */
typedef struct {
  double beta;
  int n_max, radius, ks, nc;
  double *b1d;  /* (n_max+1) x ks, FP64 samples */
  int *onx, *ony; /* component -> (nx, ny) */
} gh_basis;

int gh_init(gh_basis *b, double beta, int n_max, int radius);
void gh_free(gh_basis *b);
/* 2-D footprint of component c cast as the reference casts it (basis.cpp:66-79). */
void gh_kernel2d(const gh_basis *b, int c, real *out);

/* ---- thin-plate spline (src/spatial.cpp); k == 1 is the constant mode ---- */
typedef struct {
  int k;
  double cx, cy, scale;
  double *kn;  /* k x 2 normalised knots */
  double *Z;   /* k x (k-3) affine null space, row-major */
  double *pen; /* k x k penalty */
} tps_basis;

int tps_init(tps_basis *t, const double *knots, int k);
void tps_free(tps_basis *t);
void tps_design_row(const tps_basis *t, double px, double py, double *row);
double tps_min_spacing(const tps_basis *t);

/* ---- coarse field lattice (src/subtract.cpp:107-316) ---- */
typedef struct {
  int exact, nf, w, h, ncx, ncy, stride;
  double *coarse; /* nf x ncy x ncx */
  int *gx, *gy, *cellx, *celly;
  real *wx, *wy;
  real *full; /* exact mode: nf x h x w */
} coarse_fields;

int cf_build(coarse_fields *cf, const tps_basis *t, const double *theta, int nc,
             int w, int h);
void cf_free(coarse_fields *cf);
real cf_point(const coarse_fields *cf, int f, int x, int y);
void cf_fill_tile(const coarse_fields *cf, int f, int bx, int by, int bw, int bh,
                  real *dst);

float fill_value_f(const float *ref, const uint8_t *mask, size_t n);
/* k-th smallest of v[0..n) (v is reordered). */
float select_kth_f(float *v, size_t n, size_t k);
double select_kth_d(double *v, size_t n, size_t k);

#endif
