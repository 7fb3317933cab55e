// Batched FP64 penalised-GLS solve over the whole lambda grid (reference
// src/solve.cpp:147-262).  For every grid point l:
//   A_l = M + lambda_l P0,   L_l = chol(A_l),   theta_l = A_l^-1 rhs,
//   sq_l = ||L_l^-1 F||_F^2 with F F^T = P0,   edf_l = P - lambda_l sq_l.
//
// Right-looking blocked Cholesky with 64-wide panels, run for all grid points
// at once, the trailing matrix updated by groups of kGroup panels (one panel
// per group at P = 725).  The launches over (lambda x tile) are:
//   k_diag    one workgroup per lambda factors the 64x64 diagonal block in
//             LDS, blocked by 16, and forms its triangular inverse Linv
//             (first panel; the workgroup of k_update that updates the next
//             diagonal block factors it in the same launch);
//   k_panel   L_I = A_I Linv^T for the row blocks below (the panel solve as a
//             64x64x64 product) and X_p = Linv X_p for the panel rows of
//             [F | rhs], plus the sum of squares of the solved F entries;
//   k_update  trailing lower-triangle update A[I,J] -= L_I L_J^T and the
//             forward update X[I] -= L_I X_p, by groups of kGroup panels:
//             each panel's column block is brought up to date with the
//             group's earlier panels alone (kColumn) before its panel solve,
//             and everything below the group is updated once with all of
//             them (kTrail, K = 64 kGroup), so each trailing tile is read and
//             written once per group.
// The 64x64x64 products of k_panel and k_update run on v_mfma_f64_16x16x4f64.
// k_backsolve then forms theta = L^-T y block by block with the stored
// inverses.  A lambda whose Cholesky meets a non-positive pivot is flagged
// and re-solved by the pivoted LDLT fallback (ldlt.hip), as the reference.
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "dg_kernels.hpp"
#include "gcv.hpp"

namespace dg {
namespace {

constexpr int NB = 64;
constexpr int kMaxPanels = 16;
constexpr int kGroup = 1;  // panels per trailing update (2 and 3 measured slower at P = 725)
constexpr int kPer = NB * NB / 256;  // tile entries per thread of a 256-thread workgroup

struct Sys {
  int p, ld;        // order P; leading dimension of the per-lambda matrices
  int nl;           // number of grid points
  double* A;        // nl x P x ld   (factor in place, lower)
  double* X;        // nl x P x ldx  ([F | rhs], solved in place)
  double* Linv;     // nl x npanel x 64 x 64 inverses of the diagonal blocks
  int ldx, rcols;   // ldx >= rcols + 1; rcols = columns of F
  int npanel, ncolblk;
  int* fail;        // nl
  double* sq;       // nl x ncolblk x kMaxPanels partial sums of squares
  int k, r;         // per-field knots and penalty rank (F block structure)
};

// A_l = M + lambda_l * blockdiag(pen) (lower triangle); X_l = [F | rhs].
// One wave per row (four rows per workgroup), lanes along the columns.
__global__ __launch_bounds__(256) void k_init(Sys s, const double* __restrict__ M, int ldm,
                                              const double* __restrict__ rhs,
                                              const double* __restrict__ pen,
                                              const double* __restrict__ fk,
                                              const double* __restrict__ lambdas,
                                              const int* __restrict__ base, size_t m_stride,
                                              size_t rhs_stride) {
  const int l = blockIdx.y;
  const double lam = lambdas[l];
  if (base) {
    M += size_t(base[l]) * m_stride;
    rhs += size_t(base[l]) * rhs_stride;
  }
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (blockIdx.x == 0 && threadIdx.x == 0) s.fail[l] = 0;
  // the lambda's partial sums of squares start at zero
  if (blockIdx.x == 0)
    for (int e = threadIdx.x; e < s.ncolblk * kMaxPanels; e += blockDim.x)
      s.sq[size_t(l) * s.ncolblk * kMaxPanels + e] = 0.0;
  if (i >= s.p) return;
  const int bi = i / s.k, ii = i - bi * s.k;  // field block of the row, row within it
  const int d0 = bi * s.k;                     // first column of the diagonal block
  double* A = s.A + size_t(l) * s.p * s.ld + size_t(i) * s.ld;
  const double* Mi = M + size_t(i) * ldm;
#pragma unroll 4
  for (int j = lane; j <= i; j += 64) {  // the factorisation reads the lower triangle only
    double v = Mi[j];
    if (j >= d0) v += lam * pen[ii * s.k + (j - d0)];
    A[j] = v;
  }
  double* X = s.X + size_t(l) * s.p * s.ldx + size_t(i) * s.ldx;
  const int f0 = bi * s.r;  // first F column of the row's field block
#pragma unroll 4
  for (int c = lane; c <= s.rcols; c += 64) {
    double v;
    if (c == s.rcols)
      v = rhs[i];
    else
      v = (c >= f0 && c < f0 + s.r) ? fk[ii * s.k + (c - f0)] : 0.0;
    X[c] = v;
  }
}

// Broadcast of lane l's double to the whole wave (l a compile-time constant in
// the unrolled loops below).
__device__ __forceinline__ double lane_bcast(double v, int l) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b), l));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), l));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// (lambda, tile) of this workgroup in a (tiles x lambda) grid.  Workgroups
// are dealt to the eight XCDs in turn (workgroup i to XCD i mod 8), each XCD
// with its own L2; numbering them so that every XCD takes one contiguous run
// of (lambda, tile) pairs keeps a lambda's panel columns in one or two L2s
// instead of all eight.
__device__ __forceinline__ void xcd_tile(int& l, int& x) {
  const int n = gridDim.x * gridDim.y;
  const int lin = blockIdx.y * gridDim.x + blockIdx.x;
  const int q = n >> 3, r = n & 7, xcd = lin & 7, loc = lin >> 3;
  const int v = xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
  l = v / gridDim.x;
  x = v - l * gridDim.x;
}

// (lambda, tile) of a trailing-update workgroup: the first gridDim.y
// workgroups take tile 0 of every lambda (the next diagonal block, which is
// factored after its update), so that those start first; the rest follow
// lambda by lambda.
__device__ __forceinline__ void diag_first_tile(int& l, int& x) {
  const int lin = blockIdx.y * gridDim.x + blockIdx.x;
  if (lin < int(gridDim.y)) {
    l = lin;
    x = 0;
    return;
  }
  const int v = lin - gridDim.y, per = gridDim.x - 1;
  l = v / per;
  x = 1 + (v - l * per);
}

// Column block of X tile j when only nx of the ncolblk blocks are launched:
// the leading nx - 1 (the ones whose field blocks reach the panel) and the
// last one, which holds the rhs column; all of them when nx == ncolblk.
__device__ __forceinline__ int x_block(int j, int nx, int ncolblk) {
  return (nx == ncolblk || j < nx - 1) ? j : ncolblk - 1;
}

constexpr int SB = 16;                                   // sub-block of the diagonal factor
constexpr int kDiagLds = (2 * NB * (NB + 1) + 3 * SB * SB) * 8;  // bytes of k_diag's LDS

typedef double d4v __attribute__((ext_vector_type(4)));

// acc += X(16 x 16) * Y(16 x 16) on v_mfma_f64_16x16x4f64, operands in LDS:
// X = x[xr + i][xc + k]; Y(k, j) = y[yr + k][yc + j], or y[yr + j][yc + k]
// when YT.  Result entry (r, c) = acc[e] with r = (lane >> 4) + 4e,
// c = lane & 15.
template <bool YT>
__device__ __forceinline__ void mfma16(const double (*x)[NB + 1], int xr, int xc,
                                       const double (*y)[NB + 1], int yr, int yc, d4v& acc,
                                       int lane) {
  const int i = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < SB; k0 += 4) {
    const double av = x[xr + i][xc + k0 + kq];
    const double bv = YT ? y[yr + i][yc + k0 + kq] : y[yr + k0 + kq][yc + i];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  }
}

// Factor and invert the diagonal block of panel pnl for grid point l, held
// in LDS: a = the block's lower triangle (zero above, identity padding past
// the order), iv = 0.  Writes the factor back to A and the inverse to Linv;
// flags the grid point when a pivot is not positive.  Called by all 256
// threads.  Blocked by 16: for each sub-panel wave 0 holds one row per lane
// (the 16 rows of the diagonal block and the rows below it) and runs the
// right-looking elimination of its 16 columns in registers, the pivot rows'
// entries exchanged by lane broadcasts, so the rows below come out as
// L_bs = A_bs L_ss^-T in the same pass; then all threads apply the trailing
// update A22 -= L21 L21^T on MFMA.  Afterwards the four 16x16 diagonal
// blocks are inverted side by side (one wave each), and the off-diagonal
// blocks of the inverse follow block row by block row:
// X_ij = -D_i sum_{k=j}^{i-1} L_ik X_kj.
__device__ void diag_factor(const Sys& s, int l, int pnl, double (*a)[NB + 1],
                            double (*iv)[NB + 1], double* tmp, int* bad_flag) {
  const int kb = pnl * NB;
  const int nb = min(NB, s.p - kb);
  int& bad = *bad_flag;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double* A = s.A + size_t(l) * s.p * s.ld;
  if (t == 0) bad = 0;
  __syncthreads();
  for (int s0 = 0; s0 < NB; s0 += SB) {
    if (wv == 0) {
      const int ri = lane;  // row s0 + ri of the block
      const bool live = s0 + ri < NB;
      double r[SB];
#pragma unroll
      for (int c = 0; c < SB; ++c) r[c] = live ? a[s0 + ri][s0 + c] : 0.0;
      bool ok = true;
#pragma unroll
      for (int j = 0; j < SB; ++j) {
        const double d = lane_bcast(r[j], j);
        ok = ok && d > 0.0;
        // 1/sqrt(d) from the hardware estimate and two Newton steps (a few
        // ulp, no correctly rounded sqrt and division on the pivot chain)
        double y = __builtin_amdgcn_rsq(d);
        y = y * fma(-0.5 * d * y, y, 1.5);
        y = y * fma(-0.5 * d * y, y, 1.5);
        // 1 / L_jj is the inverse's diagonal entry
        if (lane == 0) iv[s0 + j][s0 + j] = y;
        r[j] = ri == j ? d * y : (ri > j ? r[j] * y : r[j]);
#pragma unroll
        for (int c = j + 1; c < SB; ++c) r[c] = fma(-r[j], lane_bcast(r[j], c), r[c]);
      }
      if (live) {
#pragma unroll
        for (int c = 0; c < SB; ++c) a[s0 + ri][s0 + c] = (ri >= SB || c <= ri) ? r[c] : 0.0;
      }
      if (lane == 0 && !ok) bad = 1;
    }
    __syncthreads();
    if (bad) {
      if (t == 0) s.fail[l] = 1;
      return;
    }
    const int sb = s0 / SB;  // sub-panel index
    if (sb == NB / SB - 1) break;
    // trailing update A_ij -= L_is L_js^T for s < j <= i (whole 16x16 blocks;
    // the diagonal blocks' upper halves take values that are never read)
    {
      const int m = NB / SB - 1 - sb;  // blocks below the sub-panel
      const int ntr = m * (m + 1) / 2;
      for (int q = wv; q < ntr; q += 4) {
        int bi = 0, rr = q;
        while (rr > bi) {
          rr -= bi + 1;
          ++bi;
        }
        const int ib = sb + 1 + bi, jb = sb + 1 + rr;
        d4v acc = d4v{0.0, 0.0, 0.0, 0.0};
        mfma16<true>(a, SB * ib, s0, a, SB * jb, s0, acc, lane);
#pragma unroll
        for (int e = 0; e < 4; ++e) a[SB * ib + (lane >> 4) + 4 * e][SB * jb + (lane & 15)] -= acc[e];
      }
    }
    __syncthreads();
  }
  // the diagonal blocks' inverses, wave w inverting block w: lane ri forms
  // column ri by forward substitution, x_q = x_q / L_qq, then
  // x_i -= L_iq x_q for i > q
  {
    const int o = SB * wv, ri = lane & (SB - 1);
    double r[SB], x[SB], rinv[SB];
#pragma unroll
    for (int c = 0; c < SB; ++c) {
      r[c] = a[o + ri][o + c];
      rinv[c] = iv[o + c][o + c];
      x[c] = c == ri ? 1.0 : 0.0;
    }
#pragma unroll
    for (int q = 0; q < SB; ++q) {
      x[q] = x[q] * rinv[q];
#pragma unroll
      for (int i = q + 1; i < SB; ++i) x[i] = fma(-lane_bcast(r[q], i), x[q], x[i]);
    }
    if (lane < SB) {
#pragma unroll
      for (int c = 0; c < SB; ++c) iv[o + c][o + ri] = x[c];
    }
  }
  __syncthreads();
  // off-diagonal blocks of the inverse by levels d = i - j:
  // X_ij = -D_i sum_{k=j}^{i-1} L_ik X_kj, the blocks of one level in parallel
  // (one wave each), T in the wave's slice of tmp
  for (int d = 1; d < NB / SB; ++d) {
    if (wv < NB / SB - d) {
      const int jb = wv, ib = wv + d;
      d4v acc = d4v{0.0, 0.0, 0.0, 0.0};
      for (int kb2 = jb; kb2 < ib; ++kb2)
        mfma16<false>(a, SB * ib, SB * kb2, iv, SB * kb2, SB * jb, acc, lane);
      // T -> LDS (this wave's 16 x 16 slice of tmp), then X = -D_i T
      double* tp = tmp + wv * SB * SB;
#pragma unroll
      for (int e = 0; e < 4; ++e) tp[((lane >> 4) + 4 * e) * SB + (lane & 15)] = acc[e];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      d4v x = d4v{0.0, 0.0, 0.0, 0.0};
      const int i = lane & 15, kq = lane >> 4;
#pragma unroll
      for (int k0 = 0; k0 < SB; k0 += 4) {
        const double av = iv[SB * ib + i][SB * ib + k0 + kq];
        const double bv = tp[(k0 + kq) * SB + i];
        x = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, x, 0, 0, 0);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) iv[SB * ib + (lane >> 4) + 4 * e][SB * jb + (lane & 15)] = -x[e];
    }
    __syncthreads();
  }
  double* Li = s.Linv + (size_t(l) * s.npanel + pnl) * NB * NB;
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int e = t + 256 * u, i = e >> 6, j = e & 63;
    Li[e] = (i < nb && j < nb) ? iv[i][j] : 0.0;
    if (i < nb && j < nb) A[size_t(kb + i) * s.ld + kb + j] = a[i][j];
  }
}


// Diagonal block of the first panel (later panels are factored by k_update
// as soon as their block is updated).
__global__ __launch_bounds__(256) void k_diag(Sys s, int pnl) {
  const int l = blockIdx.x;
  if (s.fail[l]) return;
  const int kb = pnl * NB;
  const int nb = min(NB, s.p - kb);
  extern __shared__ double dsm[];
  double(*a)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(dsm);
  double(*iv)[NB + 1] = reinterpret_cast<double(*)[NB + 1]>(dsm + NB * (NB + 1));
  double* tmp = dsm + 2 * NB * (NB + 1);  // 3 x 16 x 16
  __shared__ int bad;
  const int t = threadIdx.x;
  const double* A = s.A + size_t(l) * s.p * s.ld;
  {
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u, i = e >> 6, j = e & 63;
      v[u] = i == j ? 1.0 : 0.0;
      if (i < nb && j < nb) v[u] = j <= i ? A[size_t(kb + i) * s.ld + kb + j] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u, i = e >> 6, j = e & 63;
      a[i][j] = v[u];
      iv[i][j] = 0.0;
    }
  }
  diag_factor(s, l, pnl, a, iv, tmp, &bad);
}

// 64x64 tile product on v_mfma_f64_16x16x4f64: wave (wy, wx) of the four owns
// the 32x32 quadrant, as 2x2 blocks of 16x16.
//   acc[u][v] += X[32wy+16u+., q] * Y(q, 32wx+16v+.),  q < kmax
// with Y(q, c) = y[q][c], or y[c][q] when YT.  Entries of x and y past kmax
// (up to the next multiple of 4) must be zero.
template <bool YT>
__device__ __forceinline__ void tile_mfma(const double (*x)[NB + 1], const double (*y)[NB + 1],
                                          int kmax, d4v (&acc)[2][2], int wy, int wx, int lane) {
  const int i = lane & 15, kq = lane >> 4;
  auto load = [&](int q0, double(&av)[2], double(&bv)[2]) {
    const int q = q0 + kq;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      av[u] = x[32 * wy + 16 * u + i][q];
      bv[u] = YT ? y[32 * wx + 16 * u + i][q] : y[q][32 * wx + 16 * u + i];
    }
  };
  auto step = [&](const double(&av)[2], const double(&bv)[2]) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
        acc[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[v], acc[u][v], 0, 0, 0);
  };
  if (kmax == NB) {
    // full depth: the operands of step q0 + 4 are read while step q0's
    // MFMAs issue (the same products in the same order)
    double av[2], bv[2];
    load(0, av, bv);
#pragma unroll
    for (int q0 = 0; q0 < NB; q0 += 4) {
      double an[2], bn[2];
      if (q0 + 4 < NB) load(q0 + 4, an, bn);
      step(av, bv);
      if (q0 + 4 < NB) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          av[u] = an[u];
          bv[u] = bn[u];
        }
      }
    }
    return;
  }
  for (int q0 = 0; q0 < kmax; q0 += 4) {
    double av[2], bv[2];
    load(q0, av, bv);
    step(av, bv);
  }
}

// Tile coordinates of accumulator entry (u, v, e): f64 16x16 result layout,
// lane l and register e -> row l/16 + 4e, column l%16.
__device__ __forceinline__ int acc_row(int wy, int u, int e, int lane) {
  return 32 * wy + 16 * u + (lane >> 4) + 4 * e;
}
__device__ __forceinline__ int acc_col(int wx, int v, int lane) { return 32 * wx + 16 * v + (lane & 15); }

// Blocks [0, nblk): panel solve of row block I below the diagonal block;
// blocks [nblk, nblk + ncolblk): forward solve of the panel rows of X for one
// 64-column block, with the sum of squares of the solved F entries.
__global__ __launch_bounds__(256) void k_panel(Sys s, int pnl, int nblk, int nx) {
  int l, bx;
  xcd_tile(l, bx);
  if (s.fail[l]) return;
  const int kb = pnl * NB;
  const int nb = min(NB, s.p - kb);
  __shared__ double xa[NB][NB + 1];  // operand
  __shared__ double li[NB][NB + 1];  // Linv
  __shared__ double red[256];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wy = wv >> 1, wx = wv & 1;
  const double* Lin = s.Linv + (size_t(l) * s.npanel + pnl) * NB * NB;
  double* A = s.A + size_t(l) * s.p * s.ld;
  d4v acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = d4v{0.0, 0.0, 0.0, 0.0};
  if (bx < nblk) {
    const int i0 = kb + nb + bx * NB;
    // out[r][c] = sum_q A[i0+r][kb+q] * Linv[c][q]
    {
      double va[kPer], vl[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, q = e & 63;
        va[u] = (i0 + r < s.p && q < nb) ? A[size_t(i0 + r) * s.ld + kb + q] : 0.0;
        vl[u] = Lin[e];
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, q = e & 63;
        xa[r][q] = va[u];
        li[r][q] = vl[u];
      }
    }
    __syncthreads();
    tile_mfma<true>(xa, li, nb, acc, wy, wx, lane);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = i0 + acc_row(wy, u, e, lane), c = acc_col(wx, v, lane);
          if (i < s.p && c < nb) A[size_t(i) * s.ld + kb + c] = acc[u][v][e];
        }
    return;
  }
  const int cb = x_block(bx - nblk, nx, s.ncolblk);
  const int c0 = cb * NB;
  // the panel rows of columns in this block are all zero if every column's
  // field block starts below the panel (rhs column excluded)
  double* X = s.X + size_t(l) * s.p * s.ldx;
  const bool dead = (c0 + NB - 1 < s.rcols) && ((c0 / s.r) * s.k >= kb + nb);
  double ssq = 0.0;
  if (!dead) {
    // out[r][c] = sum_q Linv[r][q] * X[kb+q][c0+c]
    {
      double va[kPer], vl[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, q = e & 63;
        vl[u] = Lin[e];
        va[u] = (r < nb && c0 + q <= s.rcols) ? X[size_t(kb + r) * s.ldx + c0 + q] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, q = e & 63;
        li[r][q] = vl[u];
        xa[r][q] = va[u];
      }
    }
    __syncthreads();
    tile_mfma<false>(li, xa, nb, acc, wy, wx, lane);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = acc_row(wy, u, e, lane), c = c0 + acc_col(wx, v, lane);
          const double x = acc[u][v][e];
          if (r < nb && c <= s.rcols) {
            X[size_t(kb + r) * s.ldx + c] = x;
            if (c < s.rcols) ssq = fma(x, x, ssq);
          }
        }
  }
  red[t] = ssq;
  __syncthreads();
  for (int sft = 128; sft > 0; sft >>= 1) {
    if (t < sft) red[t] += red[t + sft];
    __syncthreads();
  }
  if (t == 0) s.sq[(size_t(l) * s.ncolblk + cb) * kMaxPanels + pnl] = red[0];
}

// Trailing updates by groups of panels p .. p+npan-1, so that each trailing
// tile is read and written once per group:
//   kColumn  the column block c = p+npan only: A[I, c] -= sum_panels L_I L_c^T
//            (I >= c) and X[c] -= sum_panels L_c X_panel; the workgroup of tile
//            (c, c) then factors that diagonal block;
//   kTrail   every tile at or below (c, c): A[I,J] -= sum_panels L_I L_J^T
//            (K = 64 npan, in 64-deep steps) and X[I] -= sum_panels L_I X_panel;
//            the workgroup of tile (c, c) factors it.
// Within a group the panels are solved one after another, each after the
// kColumn launch that brings its column block up to date.  Blocks [0, ntri)
// are A tiles, the rest X tiles (row block x 64-column block).
enum : int { kColumn = 0, kTrail = 1 };

template <int MODE>
__global__ __launch_bounds__(256) void k_update(Sys s, int pnl, int npan, int nx) {
  int l, bx;
  diag_first_tile(l, bx);
  if (s.fail[l]) return;
  const int kb = pnl * NB;                        // first column of panel p (full: 64 wide)
  const int r0 = kb + npan * NB;                  // first updated row / column block
  const int nrb = s.npanel - (pnl + npan);        // row blocks updated
  const int ntri = MODE == kTrail ? nrb * (nrb + 1) / 2 : nrb;
  __shared__ double xa[NB][NB + 1], yb[NB][NB + 1];
  __shared__ double dtmp[3 * SB * SB];
  __shared__ int dbad;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, wy = wv >> 1, wx = wv & 1;
  double* A = s.A + size_t(l) * s.p * s.ld;
  d4v acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v) acc[u][v] = d4v{0.0, 0.0, 0.0, 0.0};
  if (bx < ntri) {
    int ti = bx, tj = 0;
    if (MODE == kTrail) {
      int q = bx;
      ti = 0;
      while (q > ti) {
        q -= ti + 1;
        ++ti;
      }
      tj = q;
    }
    const int i0 = r0 + ti * NB, j0 = r0 + tj * NB;
    // A[I,J] -= sum over the panels of L_I L_J^T
    for (int h = 0; h < npan; ++h) {
      const int cb = kb + h * NB, w = min(NB, s.p - cb);
      double va[kPer], vb[kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, c = e & 63;
        va[u] = (i0 + r < s.p && c < w) ? A[size_t(i0 + r) * s.ld + cb + c] : 0.0;
        vb[u] = (j0 + r < s.p && c < w) ? A[size_t(j0 + r) * s.ld + cb + c] : 0.0;
      }
      if (h > 0) __syncthreads();  // every wave is done with the previous half
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = t + 256 * u, r = e >> 6, c = e & 63;
        xa[r][c] = va[u];
        yb[r][c] = vb[u];
      }
      __syncthreads();
      tile_mfma<true>(xa, yb, w, acc, wy, wx, lane);
    }
    // the tile's current values (loaded after the products, which keeps
    // them out of the registers live across the MFMA loop)
    double old[2][2][4];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = i0 + acc_row(wy, u, e, lane), j = j0 + acc_col(wx, v, lane);
          old[u][v][e] = (i < s.p && j < s.p && j <= i) ? A[size_t(i) * s.ld + j] : 0.0;
        }
    const bool next_diag = ti == 0 && tj == 0;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int i = i0 + acc_row(wy, u, e, lane), j = j0 + acc_col(wx, v, lane);
          if (i < s.p && j < s.p && j <= i) A[size_t(i) * s.ld + j] = old[u][v][e] - acc[u][v][e];
        }
    if (!next_diag) return;
    // this tile is the next diagonal block to factor: factor it now, from the
    // updated values still in registers
    __syncthreads();  // every wave is done reading xa / yb
    const int nb2 = min(NB, s.p - r0);
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = acc_row(wy, u, e, lane), c = acc_col(wx, v, lane);
          double x = r == c ? 1.0 : 0.0;
          if (r < nb2 && c < nb2) x = c <= r ? old[u][v][e] - acc[u][v][e] : 0.0;
          xa[r][c] = x;
          yb[r][c] = 0.0;
        }
    diag_factor(s, l, r0 / NB, xa, yb, dtmp, &dbad);
    return;
  }
  const int e2 = bx - ntri;
  const int bi = e2 / nx, cb = x_block(e2 - bi * nx, nx, s.ncolblk);
  const int i0 = r0 + bi * NB, c0 = cb * NB;
  // the panels' rows of these columns are all zero when every column's field
  // block starts below them (rhs column excluded)
  if ((c0 + NB - 1 < s.rcols) && ((c0 / s.r) * s.k >= min(r0, s.p))) return;
  double* X = s.X + size_t(l) * s.p * s.ldx;
  for (int h = 0; h < npan; ++h) {
    const int cbk = kb + h * NB, w = min(NB, s.p - cbk);
    double va[kPer], vb[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u, r = e >> 6, c = e & 63;
      va[u] = (i0 + r < s.p && c < w) ? A[size_t(i0 + r) * s.ld + cbk + c] : 0.0;
      vb[u] = (r < w && c0 + c <= s.rcols) ? X[size_t(cbk + r) * s.ldx + c0 + c] : 0.0;
    }
    if (h > 0) __syncthreads();
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = t + 256 * u, r = e >> 6, c = e & 63;
      xa[r][c] = va[u];
      yb[r][c] = vb[u];
    }
    __syncthreads();
    tile_mfma<false>(xa, yb, w, acc, wy, wx, lane);
  }
  double old[2][2][4];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + acc_row(wy, u, e, lane), c = c0 + acc_col(wx, v, lane);
        old[u][v][e] = (i < s.p && c <= s.rcols) ? X[size_t(i) * s.ldx + c] : 0.0;
      }
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int i = i0 + acc_row(wy, u, e, lane), c = c0 + acc_col(wx, v, lane);
        if (i < s.p && c <= s.rcols) X[size_t(i) * s.ldx + c] = old[u][v][e] - acc[u][v][e];
      }
}

// theta = L^-T y, one workgroup per lambda, panel by panel from the last:
// y_p -= sum_{rows below} L[i][p-cols] theta_i, then theta_p = Linv_p^T y_p.
// Also the deterministic sum of the ||L^-1 F||^2 partials.
constexpr int kBsWaves = 16;  // waves of k_backsolve
constexpr int kBsBatch = 32;  // rows per wave with loads in flight together

__global__ __launch_bounds__(64 * kBsWaves) void k_backsolve(Sys s, double* theta, double* sqsum) {
  const int l = blockIdx.x;
  if (s.fail[l]) return;
  extern __shared__ double th[];  // P
  __shared__ double yp[NB];
  __shared__ double part[kBsWaves][NB];
  const int t = threadIdx.x;
  const double* A = s.A + size_t(l) * s.p * s.ld;
  const double* X = s.X + size_t(l) * s.p * s.ldx;
  for (int i = t; i < s.p; i += 64 * kBsWaves) th[i] = 0.0;
  __syncthreads();
  const int c = t & 63, g = t >> 6;
  for (int pnl = s.npanel - 1; pnl >= 0; --pnl) {
    const int kb = pnl * NB, nb = min(NB, s.p - kb);
    // rows below the panel contribute sum_i L[i][kb+c] theta_i for each c;
    // wave g takes rows kb+nb+g, +16, ... in batches whose loads overlap
    double acc = 0.0;
    if (c < nb) {
      for (int i0 = kb + nb + g; i0 < s.p; i0 += kBsWaves * kBsBatch) {
        double v[kBsBatch];
#pragma unroll
        for (int u = 0; u < kBsBatch; ++u) {
          const int i = i0 + kBsWaves * u;
          v[u] = i < s.p ? A[size_t(i) * s.ld + kb + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBsBatch; ++u) {
          const int i = i0 + kBsWaves * u;
          if (i < s.p) acc = fma(v[u], th[i], acc);
        }
      }
    }
    part[g][c] = acc;
    __syncthreads();
    if (t < nb) {
      double sum = 0.0;
      for (int q = 0; q < kBsWaves; ++q) sum += part[q][t];
      yp[t] = X[size_t(kb + t) * s.ldx + s.rcols] - sum;
    }
    __syncthreads();
    // theta_p = Linv^T y_p: wave g sums q in [4g, 4g+4)
    const double* Lin = s.Linv + (size_t(l) * s.npanel + pnl) * NB * NB;
    double v = 0.0;
    if (c < nb) {
#pragma unroll
      for (int u = 0; u < NB / kBsWaves; ++u) {
        const int q = (NB / kBsWaves) * g + u;
        if (q >= c && q < nb) v = fma(Lin[q * NB + c], yp[q], v);
      }
    }
    part[g][c] = v;
    __syncthreads();
    if (t < nb) {
      double sum = 0.0;
      for (int q = 0; q < kBsWaves; ++q) sum += part[q][t];
      th[kb + t] = sum;
    }
    __syncthreads();
  }
  for (int i = t; i < s.p; i += 64 * kBsWaves) theta[size_t(l) * s.p + i] = th[i];
  if (t == 0) {
    double tot = 0.0;
    for (int pnl = 0; pnl < s.npanel; ++pnl)
      for (int b = 0; b < s.ncolblk; ++b) tot += s.sq[(size_t(l) * s.ncolblk + b) * kMaxPanels + pnl];
    sqsum[l] = tot;
  }
}

}  // namespace

size_t gcv_scratch_bytes(int p, int k, int nl) {
  const int ld = p + 1;
  const int rcols = p;  // upper bound of rank * (nc + 1)
  const int ldx = rcols + 2;
  const int npanel = (p + NB - 1) / NB;
  const int ncolblk = (rcols + 1 + NB - 1) / NB;
  (void)k;
  return Arena::round(sizeof(double) * size_t(nl) * p * ld) +
         Arena::round(sizeof(double) * size_t(nl) * p * ldx) +
         Arena::round(sizeof(double) * size_t(nl) * npanel * NB * NB) +
         Arena::round(sizeof(double) * size_t(nl) * ncolblk * kMaxPanels);
}

namespace {

// The launch sequence of one grid solve on stream st.
void gcv_enqueue(const Sys& s, const GcvProblem& g, hipStream_t st) {
  hipLaunchKernelGGL(k_init, dim3((s.p + 3) / 4, g.nl), dim3(256), 0, st, s, g.M, g.ldm, g.rhs, g.pen, g.fk,
                     g.lambdas, g.base, g.m_stride, g.rhs_stride);
  hipLaunchKernelGGL(k_diag, dim3(g.nl), dim3(256), kDiagLds, st, s, 0);
  const auto rows_below = [&](int pnl) {
    const int kb = pnl * NB;
    return (s.p - kb - std::min(NB, s.p - kb) + NB - 1) / NB;
  };
  // X column blocks worth launching when the rows above thr are the panel's:
  // a block whose columns all belong to field blocks starting at or below thr
  // (rhs column excluded) has all-zero panel rows, and those blocks form a
  // run before the last block; the kernels keep the same test
  const auto x_blocks = [&](int thr) {
    int cbl = 0;
    while (cbl < s.ncolblk) {
      const int c0 = cbl * NB;
      if (c0 + NB - 1 < s.rcols && (c0 / s.r) * s.k >= thr) break;
      ++cbl;
    }
    return cbl == s.ncolblk ? s.ncolblk : cbl + 1;
  };
  for (int pnl = 0; pnl < s.npanel; pnl += kGroup) {
    const int last = std::min(pnl + kGroup, s.npanel);  // panels [pnl, last) form the group
    for (int q = pnl; q < last; ++q) {
      if (q > pnl) {
        // bring column block q up to date with the group's earlier panels and
        // factor its diagonal block
        const int nxc = x_blocks(std::min(q * NB, s.p));
        hipLaunchKernelGGL(k_update<kColumn>, dim3(s.npanel - q + nxc, g.nl), dim3(256), 0,
                           st, s, pnl, q - pnl, nxc);
      }
      const int nblk = rows_below(q);
      const int nxp = x_blocks(q * NB + std::min(NB, s.p - q * NB));
      hipLaunchKernelGGL(k_panel, dim3(nblk + nxp, g.nl), dim3(256), 0, st, s, q, nblk, nxp);
    }
    const int nrb = s.npanel - last;
    if (nrb == 0) break;
    // the whole group into everything below it
    const int nxt = x_blocks(std::min(last * NB, s.p));
    hipLaunchKernelGGL(k_update<kTrail>, dim3(nrb * (nrb + 1) / 2 + nrb * nxt, g.nl),
                       dim3(256), 0, st, s, pnl, last - pnl, nxt);
  }
  hipLaunchKernelGGL(k_backsolve, dim3(g.nl), dim3(64 * kBsWaves), sizeof(double) * s.p, st, s,
                     g.theta, g.sq);
}

// Every argument of the sequence: a replay is valid only for the same key.
struct GraphKey {
  Sys s;
  const double *M, *rhs, *pen, *fk, *lambdas;
  const int* base;
  double *theta, *sqo;
  size_t m_stride, rhs_stride;
  int ldm, dev;
};

struct GraphSlot {
  GraphKey key;
  hipGraphExec_t exec;
};

constexpr size_t kMaxGraphs = 16;
std::mutex g_graph_mu;
std::vector<GraphSlot> g_graphs;
hipStream_t g_graph_stream = nullptr;
hipEvent_t g_ev_in = nullptr, g_ev_out = nullptr;
int g_graph_dev = -1;  // the device the graph stream and events belong to

bool graphs_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("DELTA_AMD_GRAPHS");
    return !(e && std::atoi(e) == 0);
  }();
  return on;
}

// Replays (capturing first if needed) on the library's own graph stream,
// ordered after the work already on st and before the work that follows it.
// False when the graph cannot be used (capture refused): launch directly.
bool gcv_replay(const Sys& s, const GcvProblem& g, hipStream_t st) {
  GraphKey key;
  std::memset(&key, 0, sizeof(key));  // padding bytes take part in the comparison
  key.s = s;
  key.M = g.M;
  key.rhs = g.rhs;
  key.pen = g.pen;
  key.fk = g.fk;
  key.lambdas = g.lambdas;
  key.base = g.base;
  key.theta = g.theta;
  key.sqo = g.sq;
  key.m_stride = g.m_stride;
  key.rhs_stride = g.rhs_stride;
  key.ldm = g.ldm;
  DG_HIP(hipGetDevice(&key.dev));
  std::lock_guard<std::mutex> lk(g_graph_mu);
  if (!g_graph_stream) {
    DG_HIP(hipStreamCreateWithFlags(&g_graph_stream, hipStreamNonBlocking));
    DG_HIP(hipEventCreateWithFlags(&g_ev_in, hipEventDisableTiming));
    DG_HIP(hipEventCreateWithFlags(&g_ev_out, hipEventDisableTiming));
    g_graph_dev = key.dev;
  }
  if (key.dev != g_graph_dev) return false;  // graphs on the first device used only
  hipGraphExec_t exec = nullptr;
  for (const GraphSlot& gs : g_graphs)
    if (std::memcmp(&gs.key, &key, sizeof(key)) == 0) {
      exec = gs.exec;
      break;
    }
  if (!exec) {
    if (g_graphs.size() >= kMaxGraphs) return false;  // keep the ones in use; launch directly
    if (hipStreamBeginCapture(g_graph_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    gcv_enqueue(s, g, g_graph_stream);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(g_graph_stream, &graph);
    if (ec != hipSuccess || !graph) {
      (void)hipGetLastError();
      return false;
    }
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    g_graphs.push_back(GraphSlot{key, exec});
  }
  DG_HIP(hipEventRecord(g_ev_in, st));
  DG_HIP(hipStreamWaitEvent(g_graph_stream, g_ev_in, 0));
  DG_HIP(hipGraphLaunch(exec, g_graph_stream));
  DG_HIP(hipEventRecord(g_ev_out, g_graph_stream));
  DG_HIP(hipStreamWaitEvent(st, g_ev_out, 0));
  return true;
}

}  // namespace

// Solves the grid; theta (nl x P), sq (nl) and fail (nl) are device outputs.
void gcv_solve(const GcvProblem& g, Arena& ar, hipStream_t st) {
  Sys s{};
  s.p = g.p;
  s.ld = g.p + 1;
  s.nl = g.nl;
  s.k = g.k;
  s.r = g.rank > 0 ? g.rank : 1;
  s.rcols = g.rank * (g.p / g.k);
  s.ldx = s.rcols + 2;
  s.npanel = (s.p + NB - 1) / NB;
  if (s.npanel > kMaxPanels) invalid("gcv: at most 1024 unknowns are supported");
  s.ncolblk = (s.rcols + 1 + NB - 1) / NB;
  s.A = ar.take<double>(size_t(g.nl) * s.p * s.ld);
  s.X = ar.take<double>(size_t(g.nl) * s.p * s.ldx);
  s.Linv = ar.take<double>(size_t(g.nl) * s.npanel * NB * NB);
  s.sq = ar.take<double>(size_t(g.nl) * s.ncolblk * kMaxPanels);
  s.fail = g.fail;
  set_max_lds(reinterpret_cast<const void*>(&k_diag), kDiagLds);
  if (!(g.graph && graphs_enabled() && gcv_replay(s, g, st))) gcv_enqueue(s, g, st);
  DG_LAUNCH_CHECK();
}

}  // namespace dg
