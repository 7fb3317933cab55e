// Batched FP64 penalised-GLS solve over the whole lambda grid (reference
// src/solve.cpp:147-262).  For every grid point l:
//   A_l = M + lambda_l P0,   L_l = chol(A_l),   theta_l = A_l^-1 rhs,
//   sq_l = ||L_l^-1 F||_F^2 with F F^T = P0,   edf_l = P - lambda_l sq_l.
// The grid is solved in one pass of blocked right-looking kernels (64-wide
// panels), every panel step launching over (lambda x tile): diagonal factor,
// panel triangular solve, trailing lower-triangle update, then the same
// blocked forward substitution applied to [F | rhs].  A lambda whose Cholesky
// meets a non-positive pivot is re-solved by a pivoted LDLT (the reference's
// fallback), with edf = tr(A^-1 M).
#include "dg_kernels.hpp"
#include "gcv.hpp"

namespace dg {
namespace {

constexpr int NB = 64;

__global__ void k_fill_f64(double* p, size_t n, double v) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

struct Sys {
  int p, ld;        // order P; leading dimension of the per-lambda matrices
  int nl;           // number of grid points
  double* A;        // nl x P x ld   (factor in place, lower)
  double* X;        // nl x P x ldx  ([F | rhs], solved in place)
  int ldx, rcols;   // ldx >= rcols + 1; rcols = columns of F
  int* fail;        // nl
  double* sq;       // nl x ntile partial sums of squares (deterministic)
  int k, r;         // per-field knots and penalty rank (F block structure)
};

// A_l = M + lambda_l * blockdiag(pen); X_l = [F | rhs]
__global__ void k_init(Sys s, const double* __restrict__ M, int ldm, const double* __restrict__ rhs,
                       const double* __restrict__ pen, const double* __restrict__ fk,
                       const double* __restrict__ lambdas) {
  const int l = blockIdx.y;
  const double lam = lambdas[l];
  const size_t tot = size_t(s.p) * s.p;
  double* A = s.A + size_t(l) * s.p * s.ld;
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < tot;
       e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / s.p), j = int(e % s.p);
    double v = M[size_t(i) * ldm + j];
    if (i / s.k == j / s.k) v += lam * pen[(i % s.k) * s.k + (j % s.k)];
    A[size_t(i) * s.ld + j] = v;
  }
  double* X = s.X + size_t(l) * s.p * s.ldx;
  const size_t totx = size_t(s.p) * (s.rcols + 1);
  for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < totx;
       e += size_t(gridDim.x) * blockDim.x) {
    const int i = int(e / (s.rcols + 1)), c = int(e % (s.rcols + 1));
    double v;
    if (c == s.rcols) {
      v = rhs[i];
    } else {
      const int fb = c / s.r, cc = c % s.r;  // field block of this column
      v = (i / s.k == fb) ? fk[(i % s.k) * s.k + cc] : 0.0;
    }
    X[size_t(i) * s.ldx + c] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) s.fail[l] = 0;
}

// Factor the nb x nb diagonal block at kb (already updated) in LDS.
__global__ __launch_bounds__(256) void k_potrf_diag(Sys s, int kb) {
  const int l = blockIdx.x;
  if (s.fail[l]) return;
  __shared__ double a[NB][NB + 1];
  __shared__ int bad;
  const int nb = min(NB, s.p - kb);
  double* A = s.A + size_t(l) * s.p * s.ld;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    a[i][j] = A[size_t(kb + i) * s.ld + kb + j];
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    // column j of the block is final except for its own pivot step
    if (threadIdx.x == 0) {
      const double d = a[j][j];
      if (!(d > 0.0)) bad = 1;
      a[j][j] = sqrt(d);
    }
    __syncthreads();
    if (bad) break;
    const double piv = a[j][j];
    for (int i = j + 1 + threadIdx.x; i < nb; i += blockDim.x) a[i][j] /= piv;
    __syncthreads();
    // rank-1 update of the trailing lower triangle
    const int m = nb - j - 1;
    for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
      const int i = j + 1 + e / m, c = j + 1 + e % m;
      if (c <= i) a[i][c] -= a[i][j] * a[c][j];
    }
    __syncthreads();
  }
  if (bad) {
    if (threadIdx.x == 0) s.fail[l] = 1;
    return;
  }
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    A[size_t(kb + i) * s.ld + kb + j] = j <= i ? a[i][j] : 0.0;
  }
}

// Rows below the diagonal block: L_I = A_I L_dd^-T (one 64-row block per WG).
__global__ __launch_bounds__(64) void k_trsm_panel(Sys s, int kb) {
  const int l = blockIdx.y;
  if (s.fail[l]) return;
  const int nb = min(NB, s.p - kb);
  const int r0 = kb + nb + blockIdx.x * 64;
  __shared__ double ld[NB][NB + 1];
  double* A = s.A + size_t(l) * s.p * s.ld;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    ld[i][j] = A[size_t(kb + i) * s.ld + kb + j];
  }
  __syncthreads();
  const int row = r0 + threadIdx.x;
  if (row >= s.p) return;
  double x[NB];
  double* ar = A + size_t(row) * s.ld + kb;
#pragma unroll
  for (int j = 0; j < NB; ++j) x[j] = j < nb ? ar[j] : 0.0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      double t = x[j];
#pragma unroll
      for (int q = 0; q < j; ++q) t -= x[q] * ld[j][q];
      x[j] = t / ld[j][j];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < nb) ar[j] = x[j];
}

// Trailing update of lower tiles (I >= J) below the panel:
// A[I,J] -= L[I,panel] L[J,panel]^T.  Tile index -> (I, J) packed.
__global__ __launch_bounds__(256) void k_syrk(Sys s, int kb) {
  const int l = blockIdx.y;
  if (s.fail[l]) return;
  const int nb = min(NB, s.p - kb);
  const int t0 = kb + nb;
  int t = blockIdx.x, ti = 0;
  while (t > ti) {
    t -= ti + 1;
    ++ti;
  }
  const int tj = t;  // 0 <= tj <= ti
  const int i0 = t0 + ti * 64, j0 = t0 + tj * 64;
  if (i0 >= s.p) return;
  __shared__ double li[64][NB + 1], lj[64][NB + 1];
  double* A = s.A + size_t(l) * s.p * s.ld;
  for (int e = threadIdx.x; e < 64 * nb; e += blockDim.x) {
    const int r = e / nb, c = e % nb;
    li[r][c] = (i0 + r < s.p) ? A[size_t(i0 + r) * s.ld + kb + c] : 0.0;
    lj[r][c] = (j0 + r < s.p) ? A[size_t(j0 + r) * s.ld + kb + c] : 0.0;
  }
  __syncthreads();
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int q = 0; q < nb; ++q) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = li[tr + u][q];
      b[u] = lj[tc + u][q];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + tr + u, j = j0 + tc + v;
      if (i < s.p && j < s.p && j <= i) A[size_t(i) * s.ld + j] -= acc[u][v];
    }
}

// Forward substitution of the panel rows of X: X[kb block] = L_dd^-1 X[kb block]
// (one thread per column of [F | rhs]); accumulates ||.||^2 over F columns.
__global__ __launch_bounds__(64) void k_fwd_panel(Sys s, int kb, int pass) {
  const int l = blockIdx.y;
  if (s.fail[l]) return;
  const int nb = min(NB, s.p - kb);
  __shared__ double ld[NB][NB + 1];
  const double* A = s.A + size_t(l) * s.p * s.ld;
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    const int i = e / nb, j = e % nb;
    ld[i][j] = A[size_t(kb + i) * s.ld + kb + j];
  }
  __syncthreads();
  const int c = blockIdx.x * 64 + threadIdx.x;
  double ssq = 0.0;
  if (c <= s.rcols) {
    // columns of field block fb are zero above row fb*k: skip panels above it
    const bool live = (c == s.rcols) || (kb + nb > (c / s.r) * s.k);
    if (live) {
      double* X = s.X + size_t(l) * s.p * s.ldx;
      double x[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = j < nb ? X[size_t(kb + j) * s.ldx + c] : 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < nb) {
          double t = x[j];
#pragma unroll
          for (int q = 0; q < j; ++q) t -= ld[j][q] * x[q];
          x[j] = t / ld[j][j];
        }
      }
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < nb) {
          X[size_t(kb + j) * s.ldx + c] = x[j];
          ssq = fma(x[j], x[j], ssq);
        }
      if (c == s.rcols) ssq = 0.0;
    }
  }
  // deterministic per-(lambda, panel, column-block) partial
  __shared__ double red[64];
  red[threadIdx.x] = ssq;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < 64; ++i) t += red[i];
    s.sq[(size_t(l) * gridDim.x + blockIdx.x) * 16 + pass] = t;
  }
}

// X[I] -= L[I, panel] X[panel] for 64-row blocks I below the panel and
// 64-column blocks of [F | rhs].
__global__ __launch_bounds__(256) void k_fwd_update(Sys s, int kb) {
  const int l = blockIdx.z;
  if (s.fail[l]) return;
  const int nb = min(NB, s.p - kb);
  const int i0 = kb + nb + blockIdx.y * 64;
  const int c0 = blockIdx.x * 64;
  if (i0 >= s.p || c0 > s.rcols) return;
  // no-op when the panel rows of every column in this block are still zero
  // (columns of field block fb are zero above row fb*k)
  if (c0 + 63 < s.rcols && (c0 / s.r) * s.k >= kb + nb) return;
  __shared__ double li[64][NB + 1], xp[NB][64 + 1];
  const double* A = s.A + size_t(l) * s.p * s.ld;
  double* X = s.X + size_t(l) * s.p * s.ldx;
  for (int e = threadIdx.x; e < 64 * nb; e += blockDim.x) {
    const int r = e / nb, q = e % nb;
    li[r][q] = (i0 + r < s.p) ? A[size_t(i0 + r) * s.ld + kb + q] : 0.0;
  }
  for (int e = threadIdx.x; e < nb * 64; e += blockDim.x) {
    const int q = e / 64, c = e % 64;
    xp[q][c] = (c0 + c <= s.rcols) ? X[size_t(kb + q) * s.ldx + c0 + c] : 0.0;
  }
  __syncthreads();
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int q = 0; q < nb; ++q) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = li[tr + u][q];
      b[u] = xp[q][tc + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v] = fma(a[u], b[v], acc[u][v]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + tr + u, c = c0 + tc + v;
      if (i < s.p && c <= s.rcols) X[size_t(i) * s.ldx + c] -= acc[u][v];
    }
}

// theta = L^-T y (y = the solved rhs column), one WG per lambda; also the
// deterministic sum of the ||L^-1 F||^2 partials.
__global__ __launch_bounds__(256) void k_backsolve(Sys s, int ncolblk, int npanel, double* theta,
                                                   double* sqsum) {
  const int l = blockIdx.x;
  if (s.fail[l]) return;
  extern __shared__ double y[];
  const double* A = s.A + size_t(l) * s.p * s.ld;
  const double* X = s.X + size_t(l) * s.p * s.ldx;
  for (int i = threadIdx.x; i < s.p; i += blockDim.x) y[i] = X[size_t(i) * s.ldx + s.rcols];
  __syncthreads();
  for (int i = s.p - 1; i >= 0; --i) {
    if (threadIdx.x == 0) y[i] /= A[size_t(i) * s.ld + i];
    __syncthreads();
    const double ti = y[i];
    for (int q = threadIdx.x; q < i; q += blockDim.x) y[q] -= A[size_t(i) * s.ld + q] * ti;
    __syncthreads();
  }
  for (int i = threadIdx.x; i < s.p; i += blockDim.x) theta[size_t(l) * s.p + i] = y[i];
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int pnl = 0; pnl < npanel; ++pnl)
      for (int b = 0; b < ncolblk; ++b) t += s.sq[(size_t(l) * ncolblk + b) * 16 + pnl];
    sqsum[l] = t;
  }
}

}  // namespace

size_t gcv_scratch_bytes(int p, int k, int nl) {
  const int ld = p + 1;
  const int r = k;  // upper bound of the penalty rank
  const int rcols = ((p / k)) * r;
  const int ldx = rcols + 2;
  return Arena::round(sizeof(double) * size_t(nl) * p * ld) +
         Arena::round(sizeof(double) * size_t(nl) * p * ldx) + Arena::round(sizeof(int) * nl) +
         Arena::round(sizeof(double) * size_t(nl) * 16 * ((rcols + 64) / 64 + 1)) +
         Arena::round(sizeof(double) * nl);
}

// Solves the grid; theta (nl x P), sq (nl) and fail (nl) are device outputs.
void gcv_solve(const GcvProblem& g, Arena& ar, hipStream_t st) {
  Sys s{};
  s.p = g.p;
  s.ld = g.p + 1;
  s.nl = g.nl;
  s.k = g.k;
  s.r = g.rank;
  s.rcols = g.rank * (g.p / g.k);
  s.ldx = s.rcols + 2;
  s.A = ar.take<double>(size_t(g.nl) * s.p * s.ld);
  s.X = ar.take<double>(size_t(g.nl) * s.p * s.ldx);
  s.fail = g.fail;
  const int npanel = (s.p + NB - 1) / NB;
  if (npanel > 16) invalid("gcv: at most 1024 unknowns are supported");
  const int ncolblk = (s.rcols + 1 + 63) / 64;
  s.sq = ar.take<double>(size_t(g.nl) * 16 * ncolblk);
  hipLaunchKernelGGL(k_init, dim3(64, g.nl), dim3(256), 0, st, s, g.M, g.ldm, g.rhs, g.pen, g.fk,
                     g.lambdas);
  // zero the partial sums (a kernel, not a memset)
  hipLaunchKernelGGL(k_fill_f64, dim3(4), dim3(256), 0, st, s.sq, size_t(g.nl) * 16 * ncolblk, 0.0);
  for (int pnl = 0; pnl < npanel; ++pnl) {
    const int kb = pnl * NB;
    const int nb = std::min(NB, s.p - kb);
    const int below = s.p - kb - nb;
    const int nblk = (below + 63) / 64;
    hipLaunchKernelGGL(k_potrf_diag, dim3(g.nl), dim3(256), 0, st, s, kb);
    if (nblk > 0) {
      hipLaunchKernelGGL(k_trsm_panel, dim3(nblk, g.nl), dim3(64), 0, st, s, kb);
      hipLaunchKernelGGL(k_syrk, dim3(nblk * (nblk + 1) / 2, g.nl), dim3(256), 0, st, s, kb);
    }
    hipLaunchKernelGGL(k_fwd_panel, dim3(ncolblk, g.nl), dim3(64), 0, st, s, kb, pnl);
    if (nblk > 0)
      hipLaunchKernelGGL(k_fwd_update, dim3(ncolblk, nblk, g.nl), dim3(256), 0, st, s, kb);
  }
  hipLaunchKernelGGL(k_backsolve, dim3(g.nl), dim3(256), sizeof(double) * s.p, st, s, ncolblk,
                     npanel, g.theta, g.sq);
  DG_LAUNCH_CHECK();
}

}  // namespace dg
