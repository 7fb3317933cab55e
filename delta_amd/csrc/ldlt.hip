// Pivoted LDLT fallback for one grid point of the GCV solve (reference
// src/solve.cpp:177-181, Eigen's LDLT with diagonal pivoting): used only where
// the Cholesky of A = M + lambda P0 meets a non-positive pivot.  Every such
// grid point is factored at once, one workgroup each, in global memory;
// theta = A^-1 rhs and edf = tr(A^-1 M) follow by P + 1 pivoted solves (zero
// pivots give zero components, as Eigen's solve).
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

// Workgroup b factors A_b = M + lambda_{idx[b]} P0 (workspace a + b p^2).
// The pivot is the first index of the largest |diagonal| (Eigen's
// diagonal pivoting), found by a block reduction.
__global__ __launch_bounds__(kT) void k_ldlt(double* a_all, int p, const double* M, int ldm,
                                             const double* pen, int k, const double* lambdas,
                                             const int* idx, int* perm_all, double* tmp_all) {
  __shared__ double red[kT / 64];
  __shared__ double bv_w[kT / 64];
  __shared__ int bi_w[kT / 64];
  __shared__ int s_big;
  const int b = blockIdx.x;
  const double lambda = lambdas[idx[b]];
  double* a = a_all + size_t(b) * p * p;
  int* perm = perm_all + size_t(b) * p;
  double* tmp = tmp_all + size_t(b) * p;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int e = threadIdx.x; e < p * p; e += kT) {
    const int i = e / p, j = e % p;
    double v = M[size_t(i) * ldm + j];
    if (i / k == j / k) v += lambda * pen[(i % k) * k + j % k];
    a[e] = v;
  }
  __syncthreads();
  for (int kk = 0; kk < p; ++kk) {
    // argmax |a_ii|, i >= kk, first index on ties
    double bv = -1.0;
    int bi = kk;
    for (int i = kk + threadIdx.x; i < p; i += kT) {
      const double v = fabs(a[size_t(i) * p + i]);
      if (v > bv) {
        bv = v;
        bi = i;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      bv_w[wv] = bv;
      bi_w[wv] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = bv_w[0];
      int big = bi_w[0];
      for (int w = 1; w < kT / 64; ++w)
        if (bv_w[w] > v || (bv_w[w] == v && bi_w[w] < big)) {
          v = bv_w[w];
          big = bi_w[w];
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
__global__ void k_ldlt_solve(const double* a_all, int p, const int* perm_all, double* x_all,
                             int ncol) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int b = blockIdx.y;
  const double* a = a_all + size_t(b) * p * p;
  const int* perm = perm_all + size_t(b) * p;
  double* v = x_all + (size_t(b) * ncol + c) * p;
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

__global__ void k_pack_rhs(const double* M, int ldm, const double* rhs, int p, double* x_all) {
  // column 0: rhs; column j+1: column j of M
  const size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  const size_t tot = size_t(p) * (p + 1);
  if (e >= tot) return;
  double* x = x_all + size_t(blockIdx.y) * tot;
  const int c = int(e / p), i = int(e % p);
  x[e] = c == 0 ? rhs[i] : M[size_t(i) * ldm + (c - 1)];
}

__global__ void k_trace(const double* x_all, int p, const int* idx, double* theta_all,
                        double* edf_all) {
  __shared__ double red[256];
  const int b = blockIdx.x, l = idx[b];
  const double* x = x_all + size_t(b) * p * (p + 1);
  double* theta = theta_all + size_t(l) * p;
  double* edf = edf_all + l;
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

size_t ldlt_scratch_bytes(int p, int nb) {
  return Arena::round(sizeof(double) * size_t(nb) * p * p) +
         Arena::round(sizeof(int) * size_t(nb) * p) + Arena::round(sizeof(double) * size_t(nb) * p) +
         Arena::round(sizeof(double) * size_t(nb) * p * (p + 1)) + Arena::round(sizeof(int) * nb);
}

void ldlt_solve_batch(const double* M, int ldm, const double* rhs, const double* pen, int k,
                      const double* lambdas, const std::vector<int>& which, int p, double* theta,
                      double* edf, Arena& ar, hipStream_t s) {
  const int nb = static_cast<int>(which.size());
  if (nb == 0) return;
  double* a = ar.take<double>(size_t(nb) * p * p);
  int* perm = ar.take<int>(size_t(nb) * p);
  double* tmp = ar.take<double>(size_t(nb) * p);
  double* x = ar.take<double>(size_t(nb) * p * (p + 1));
  int* idx = ar.take<int>(nb);
  h2d(idx, which.data(), nb, s);
  hipLaunchKernelGGL(k_ldlt, dim3(nb), dim3(kT), 0, s, a, p, M, ldm, pen, k, lambdas,
                     static_cast<const int*>(idx), perm, tmp);
  const size_t tot = size_t(p) * (p + 1);
  hipLaunchKernelGGL(k_pack_rhs, dim3(unsigned((tot + 255) / 256), nb), dim3(256), 0, s, M, ldm,
                     rhs, p, x);
  hipLaunchKernelGGL(k_ldlt_solve, dim3((p + 1 + 63) / 64, nb), dim3(64), 0, s,
                     static_cast<const double*>(a), p, static_cast<const int*>(perm), x, p + 1);
  hipLaunchKernelGGL(k_trace, dim3(nb), dim3(256), 0, s, static_cast<const double*>(x), p,
                     static_cast<const int*>(idx), theta, edf);
  DG_LAUNCH_CHECK();
}

}  // namespace dg
