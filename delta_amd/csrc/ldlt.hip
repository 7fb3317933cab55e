// Pivoted LDLT fallback for one grid point of the GCV solve (reference
// src/solve.cpp:177-181, Eigen's LDLT with diagonal pivoting): used only where
// the Cholesky of A = M + lambda P0 meets a non-positive pivot.  One workgroup
// factors A in global memory; theta = A^-1 rhs and edf = tr(A^-1 M) follow by
// P + 1 pivoted solves (zero pivots give zero components, as Eigen's solve).
#include <cfloat>

#include "dg_kernels.hpp"
#include "gcv.hpp"

namespace dg {
namespace {

constexpr int kT = 1024;

__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kT / 64; ++i) t += red[i];
  return t;
}

__global__ __launch_bounds__(kT) void k_ldlt(double* a, int p, const double* M, int ldm,
                                             const double* pen, int k, double lambda, int* perm,
                                             double* tmp) {
  __shared__ double red[kT / 64];
  __shared__ int s_big;
  for (int e = threadIdx.x; e < p * p; e += kT) {
    const int i = e / p, j = e % p;
    double v = M[size_t(i) * ldm + j];
    if (i / k == j / k) v += lambda * pen[(i % k) * k + j % k];
    a[e] = v;
  }
  __syncthreads();
  for (int kk = 0; kk < p; ++kk) {
    if (threadIdx.x == 0) {
      int big = kk;
      double bv = fabs(a[size_t(kk) * p + kk]);
      for (int i = kk + 1; i < p; ++i) {
        const double v = fabs(a[size_t(i) * p + i]);
        if (v > bv) {
          bv = v;
          big = i;
        }
      }
      s_big = big;
      perm[kk] = big;
    }
    __syncthreads();
    const int big = s_big;
    if (big != kk) {
      for (int j = threadIdx.x; j < p; j += kT) {
        const double t = a[size_t(kk) * p + j];
        a[size_t(kk) * p + j] = a[size_t(big) * p + j];
        a[size_t(big) * p + j] = t;
      }
      __syncthreads();
      for (int j = threadIdx.x; j < p; j += kT) {
        const double t = a[size_t(j) * p + kk];
        a[size_t(j) * p + kk] = a[size_t(j) * p + big];
        a[size_t(j) * p + big] = t;
      }
      __syncthreads();
    }
    if (kk > 0) {
      for (int j = threadIdx.x; j < kk; j += kT) tmp[j] = a[size_t(j) * p + j] * a[size_t(kk) * p + j];
      __syncthreads();
      double part = 0.0;
      for (int j = threadIdx.x; j < kk; j += kT) part += a[size_t(kk) * p + j] * tmp[j];
      const double s = block_sum(part, red);
      if (threadIdx.x == 0) a[size_t(kk) * p + kk] -= s;
      for (int i = kk + 1 + threadIdx.x; i < p; i += kT) {
        double t = 0.0;
        for (int j = 0; j < kk; ++j) t += a[size_t(i) * p + j] * tmp[j];
        a[size_t(i) * p + kk] -= t;
      }
      __syncthreads();
    }
    const double akk = a[size_t(kk) * p + kk];
    if (fabs(akk) > 0.0)
      for (int i = kk + 1 + threadIdx.x; i < p; i += kT) a[size_t(i) * p + kk] /= akk;
    __syncthreads();
  }
}

// x <- A^-1 x for `ncol` right-hand sides stored column j at x + j*p (one
// column per thread), A = P^T L D L^T P from k_ldlt.
__global__ void k_ldlt_solve(const double* a, int p, const int* perm, double* x, int ncol) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double* v = x + size_t(c) * p;
  for (int kk = 0; kk < p; ++kk)
    if (perm[kk] != kk) {
      const double t = v[kk];
      v[kk] = v[perm[kk]];
      v[perm[kk]] = t;
    }
  for (int i = 0; i < p; ++i) {
    double s = v[i];
    for (int q = 0; q < i; ++q) s -= a[size_t(i) * p + q] * v[q];
    v[i] = s;
  }
  for (int i = 0; i < p; ++i) {
    const double d = a[size_t(i) * p + i];
    v[i] = fabs(d) > DBL_MIN ? v[i] / d : 0.0;
  }
  for (int i = p - 1; i >= 0; --i) {
    double s = v[i];
    for (int q = i + 1; q < p; ++q) s -= a[size_t(q) * p + i] * v[q];
    v[i] = s;
  }
  for (int kk = p - 1; kk >= 0; --kk)
    if (perm[kk] != kk) {
      const double t = v[kk];
      v[kk] = v[perm[kk]];
      v[perm[kk]] = t;
    }
}

__global__ void k_pack_rhs(const double* M, int ldm, const double* rhs, int p, double* x) {
  // column 0: rhs; column j+1: column j of M
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const size_t tot = size_t(p) * (p + 1);
  if (e >= tot) return;
  const int c = int(e / p), i = int(e % p);
  x[e] = c == 0 ? rhs[i] : M[size_t(i) * ldm + (c - 1)];
}

__global__ void k_trace(const double* x, int p, double* theta, double* edf) {
  __shared__ double red[256];
  double t = 0.0;
  for (int j = threadIdx.x; j < p; j += 256) t += x[size_t(j + 1) * p + j];
  red[threadIdx.x] = t;
  for (int i = threadIdx.x; i < p; i += 256) theta[i] = x[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < 256; ++i) s += red[i];
    *edf = s;
  }
}

}  // namespace

size_t ldlt_scratch_bytes(int p) {
  return Arena::round(sizeof(double) * size_t(p) * p) + Arena::round(sizeof(int) * p) +
         Arena::round(sizeof(double) * p) + Arena::round(sizeof(double) * size_t(p) * (p + 1));
}

void ldlt_solve_point(const double* M, int ldm, const double* rhs, const double* pen, int k,
                      double lambda, int p, double* theta, double* edf, Arena& ar,
                      hipStream_t s) {
  double* a = ar.take<double>(size_t(p) * p);
  int* perm = ar.take<int>(p);
  double* tmp = ar.take<double>(p);
  double* x = ar.take<double>(size_t(p) * (p + 1));
  hipLaunchKernelGGL(k_ldlt, dim3(1), dim3(kT), 0, s, a, p, M, ldm, pen, k, lambda, perm, tmp);
  const size_t tot = size_t(p) * (p + 1);
  hipLaunchKernelGGL(k_pack_rhs, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, s, M, ldm, rhs,
                     p, x);
  hipLaunchKernelGGL(k_ldlt_solve, dim3((p + 1 + 63) / 64), dim3(64), 0, s, a, p, perm, x, p + 1);
  hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, s, x, p, theta, edf);
  DG_LAUNCH_CHECK();
}

}  // namespace dg
