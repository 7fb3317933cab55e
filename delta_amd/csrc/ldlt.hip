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

// Workgroup b factors A_b = M + lambda_{idx[b]} P0 (workspace a + b p^2).
// The pivot is the first index of the largest |diagonal| (Eigen's
// diagonal pivoting), found by a block reduction.
__global__ __launch_bounds__(kT) void k_ldlt(double* a_all, int p, const double* M, int ldm,
                                             const double* pen, int k, const double* lambdas,
                                             const int* idx, int* perm_all, double* tmp_all) {
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
      // column kk of L D (rows i >= kk, the diagonal included):
      // a[i][kk] -= sum_{j<kk} a[i][j] d_j a[kk][j]; one wave per row, the
      // lanes along j so the row is read in coalesced runs
      for (int j = threadIdx.x; j < kk; j += kT) tmp[j] = a[size_t(j) * p + j] * a[size_t(kk) * p + j];
      __syncthreads();
      for (int i = kk + wv; i < p; i += kT / 64) {
        const double* row = a + size_t(i) * p;
        double t = 0.0;
        for (int j = lane; j < kk; j += 64) t = fma(row[j], tmp[j], t);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane == 0) a[size_t(i) * p + kk] -= t;
      }
      __syncthreads();
    }
    const double akk = a[size_t(kk) * p + kk];
    if (fabs(akk) > 0.0)
      for (int i = kk + 1 + threadIdx.x; i < p; i += kT) a[size_t(i) * p + kk] /= akk;
    __syncthreads();
  }
}

// Swap sequence of the factorisation -> gather permutation: after the swaps
// (v[kk] <-> v[perm[kk]], kk ascending) a vector b reads v[i] = b[pi[i]].
__global__ __launch_bounds__(256) void k_perm_vector(const int* perm_all, int p, int* pi_all) {
  __shared__ int sp[1024], spi[1024];  // p <= 1024 (the GCV solve's limit)
  const int b = blockIdx.x;
  const int* perm = perm_all + size_t(b) * p;
  int* pi = pi_all + size_t(b) * p;
  for (int i = threadIdx.x; i < p; i += blockDim.x) {
    sp[i] = perm[i];
    spi[i] = i;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int kk = 0; kk < p; ++kk) {
      const int o = sp[kk];
      if (o != kk) {
        const int t = spi[kk];
        spi[kk] = spi[o];
        spi[o] = t;
      }
    }
  __syncthreads();
  for (int i = threadIdx.x; i < p; i += blockDim.x) pi[i] = spi[i];
}

// at = strict upper part of L^T (at[j][i] = L[i][j], i > j), so that the
// forward substitution reads contiguous rows.
__global__ void k_ldlt_transpose(const double* a_all, int p, double* at_all) {
  __shared__ double tile[32][33];
  const int b = blockIdx.z;
  const double* a = a_all + size_t(b) * p * p;
  double* at = at_all + size_t(b) * p * p;
  const int i0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int i = i0 + r, j = j0 + threadIdx.x;
    tile[r][threadIdx.x] = (i < p && j < p) ? a[size_t(i) * p + j] : 0.0;
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += 8) {
    const int j = j0 + r, i = i0 + threadIdx.x;
    if (i < p && j < p) at[size_t(j) * p + i] = tile[threadIdx.x][r];
  }
}

// x <- A^-1 x for kSolveCols right-hand sides per workgroup (one wave), with
// A = P^T L D L^T P from k_ldlt; the columns live in LDS.  Both triangular
// solves are right-looking: each step broadcasts one finished entry and
// updates the rest from a contiguous row (of L^T forward, of L backward).
// Zero pivots give zero components, as Eigen's solve.
constexpr int kSolveCols = 8;

__global__ __launch_bounds__(64) void k_ldlt_solve(const double* a_all, const double* at_all,
                                                   int p, const int* pi_all, double* x_all,
                                                   int ncol) {
  extern __shared__ double v_lds[];  // kSolveCols x p
  const int b = blockIdx.y, c0 = blockIdx.x * kSolveCols, lane = threadIdx.x;
  const double* a = a_all + size_t(b) * p * p;
  const double* at = at_all + size_t(b) * p * p;
  const int* pi = pi_all + size_t(b) * p;
  double* xb = x_all + size_t(b) * ncol * p;
  const int nc = min(kSolveCols, ncol - c0);
  for (int c = 0; c < kSolveCols; ++c)
    for (int i = lane; i < p; i += 64)
      v_lds[c * p + i] = c < nc ? xb[size_t(c0 + c) * p + pi[i]] : 0.0;
  __builtin_amdgcn_wave_barrier();
  // L y = P b (unit diagonal)
  for (int q = 0; q + 1 < p; ++q) {
    double xq[kSolveCols];
#pragma unroll
    for (int c = 0; c < kSolveCols; ++c) xq[c] = v_lds[c * p + q];
    const double* row = at + size_t(q) * p;
    for (int i = q + 1 + lane; i < p; i += 64) {
      const double l = row[i];
#pragma unroll
      for (int c = 0; c < kSolveCols; ++c) v_lds[c * p + i] = fma(-l, xq[c], v_lds[c * p + i]);
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int i = lane; i < p; i += 64) {
    const double d = a[size_t(i) * p + i];
#pragma unroll
    for (int c = 0; c < kSolveCols; ++c)
      v_lds[c * p + i] = fabs(d) > DBL_MIN ? v_lds[c * p + i] / d : 0.0;
  }
  __builtin_amdgcn_wave_barrier();
  // L^T x = y
  for (int i = p - 1; i > 0; --i) {
    double xi[kSolveCols];
#pragma unroll
    for (int c = 0; c < kSolveCols; ++c) xi[c] = v_lds[c * p + i];
    const double* row = a + size_t(i) * p;
    for (int q = lane; q < i; q += 64) {
      const double l = row[q];
#pragma unroll
      for (int c = 0; c < kSolveCols; ++c) v_lds[c * p + q] = fma(-l, xi[c], v_lds[c * p + q]);
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int c = 0; c < nc; ++c)
    for (int i = lane; i < p; i += 64) xb[size_t(c0 + c) * p + pi[i]] = v_lds[c * p + i];
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
  return 2 * Arena::round(sizeof(double) * size_t(nb) * p * p) +
         2 * Arena::round(sizeof(int) * size_t(nb) * p) +
         Arena::round(sizeof(double) * size_t(nb) * p) +
         Arena::round(sizeof(double) * size_t(nb) * p * (p + 1)) + Arena::round(sizeof(int) * nb);
}

void ldlt_solve_batch(const double* M, int ldm, const double* rhs, const double* pen, int k,
                      const double* lambdas, const std::vector<int>& which, int p, double* theta,
                      double* edf, Arena& ar, hipStream_t s) {
  const int nb = static_cast<int>(which.size());
  if (nb == 0) return;
  double* a = ar.take<double>(size_t(nb) * p * p);
  double* at = ar.take<double>(size_t(nb) * p * p);
  int* perm = ar.take<int>(size_t(nb) * p);
  int* pi = ar.take<int>(size_t(nb) * p);
  double* tmp = ar.take<double>(size_t(nb) * p);
  double* x = ar.take<double>(size_t(nb) * p * (p + 1));
  int* idx = ar.take<int>(nb);
  h2d(idx, which.data(), nb, s);
  hipLaunchKernelGGL(k_ldlt, dim3(nb), dim3(kT), 0, s, a, p, M, ldm, pen, k, lambdas,
                     static_cast<const int*>(idx), perm, tmp);
  hipLaunchKernelGGL(k_perm_vector, dim3(nb), dim3(256), 0, s, static_cast<const int*>(perm), p,
                     pi);
  hipLaunchKernelGGL(k_ldlt_transpose, dim3((p + 31) / 32, (p + 31) / 32, nb), dim3(32, 8), 0, s,
                     static_cast<const double*>(a), p, at);
  const size_t tot = size_t(p) * (p + 1);
  hipLaunchKernelGGL(k_pack_rhs, dim3(unsigned((tot + 255) / 256), nb), dim3(256), 0, s, M, ldm,
                     rhs, p, x);
  const size_t lds = sizeof(double) * kSolveCols * p;
  set_max_lds(reinterpret_cast<const void*>(&k_ldlt_solve), int(lds));
  hipLaunchKernelGGL(k_ldlt_solve, dim3((p + 1 + kSolveCols - 1) / kSolveCols, nb), dim3(64), lds,
                     s, static_cast<const double*>(a), static_cast<const double*>(at), p,
                     static_cast<const int*>(pi), x, p + 1);
  hipLaunchKernelGGL(k_trace, dim3(nb), dim3(256), 0, s, static_cast<const double*>(x), p,
                     static_cast<const int*>(idx), theta, edf);
  DG_LAUNCH_CHECK();
}

}  // namespace dg
