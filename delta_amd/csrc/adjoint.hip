// Differentiable constant-kernel model on the GPU: chi^2 of a spatially
// constant Gauss-Hermite kernel with a free scale beta and a sub-pixel shift,
// and its adjoint.
//
//   K(u)  = sum_n a_n g_nx(ux - dx) g_ny(uy - dy),  u in [-r, r]^2
//   D(x)  = S(x) - b - sum_u K(u) R(x - u)       (R zero outside the frame)
//   chi^2 = sum_x w(x) D(x)^2                    (pixels with w <= 0 excluded)
//
// g_n is the sampled Gauss-Hermite function of the reference's basis
// (src/basis.cpp:14-59, there at dx = dy = 0); the model is the reference's
// subtraction model (src/subtract.cpp:466-639) with one spatial term.  The
// shift, the beta derivative and the adjoint have no counterpart in the
// reference.  Parameters live on the device as
//   params = [a_0 .. a_{nc-1}, b, beta, dx, dy]          (nc + 4 doubles)
// and results as
//   out    = [chi^2, d/da_0 .. d/da_{nc-1}, d/db, d/dbeta, d/ddx, d/ddy].
//
// Adjoint.  With A(x) = w D (or, for a vector-Jacobian product of D, the
// upstream gradient G), the kernel-space gradient is the correlation
//   C(u) = sum_x A(x) R(x - u),
// and every parameter gradient is a contraction of C with the basis or its
// derivatives:  dchi^2/da_n = -2 <phi_n, C>,  dchi^2/dp = -2 <dK/dp, C>,
// dchi^2/db = -2 sum A, using the Hermite ladder identities
//   dg_n/dx    = (sqrt(n/2) g_{n-1} - sqrt((n+1)/2) g_{n+1}) / beta,
//   dg_n/dbeta = -(g_n / 2 + x dg_n/dx) / beta.
// Nothing image-sized is stored between the passes: each 64 x 64 tile
// recomputes its residual from an LDS copy of the reference patch and
// correlates it with the same patch before moving on.
//
// Kernels:
//   k_adj_setup     tables and the dense FP32 kernel (transposed) on device;
//   k_adj_tiles<R>  per tile: dense 2-D convolution (the model), residual,
//                   w D^2, then the lag correlation over (2R+1)^2 lags in
//                   register-blocked sliding windows; per-tile partials;
//   k_adj_reduce    per-lag FP64 sums over 16 groups of tiles, tile order;
//   k_adj_contract  one block per gradient entry: basis contractions.
#include "dg_kernels.hpp"

#include "../../include/deltagpu.h"

namespace dg {
namespace {

constexpr int kAdjTile = 64;     // output tile edge (one wave spans a row)
constexpr int kAdjThreads = 256;  // 4 waves: 16 tile rows each in the model pass
constexpr int kAdjMaxNmax = 10;   // g tables hold orders 0..n_max+1
constexpr int kAdjMaxKs = 31;     // radius <= 15

// Padded lag grid of the correlation: 2 x 8 blocks of UY x UX lags; the
// patch carries a guard of LY - KS rows and LX - KS columns on its low side
// so the padded lags read inside LDS (their sums are discarded).
template <int R>
struct AdjGeom {
  static constexpr int KS = 2 * R + 1;
  static constexpr int UY = (KS + 1) / 2;
  static constexpr int UX = (KS + 7) / 8;
  static constexpr int LY = 2 * UY, LX = 8 * UX, L = LY * LX;
  static constexpr int GY = LY - KS, GX = LX - KS;
  static constexpr int PH = GY + kAdjTile + 2 * R;
  static constexpr int PW = GX + kAdjTile + 2 * R;
  static constexpr int NACC = UY * UX;
  static_assert(NACC <= 64, "one accumulator per lane after the reduce-scatter");
  static_assert(L % 64 == 0, "lag grid is a whole number of 64-wide columns");
};

// Gauss-Hermite tables of one axis in LDS: g[n][i] for n = 0..n_max+1 and,
// when wanted, dg/dx and dg/dbeta for n = 0..n_max, at x_i = i - r - shift.
struct GhTables {
  double g[kAdjMaxNmax + 2][kAdjMaxKs];
  double dx[kAdjMaxNmax + 1][kAdjMaxKs];
  double db[kAdjMaxNmax + 1][kAdjMaxKs];
};

__device__ void fill_tables(GhTables* tab, const double* __restrict__ params, int nc, int n_max,
                            int r, bool derivs) {
  const int ks = 2 * r + 1;
  const double beta = params[nc + 1];
  for (int t = threadIdx.x; t < 2 * ks; t += blockDim.x) {
    const int axis = t / ks, i = t - axis * ks;
    const double x = double(i - r) - params[nc + 2 + axis];
    const double u = x / beta;
    // [2^n n! sqrt(pi) beta]^(-1/2) by N_0 = (sqrt(pi) beta)^(-1/2),
    // N_n = N_{n-1} / sqrt(2n)
    double norm = 1.0 / sqrt(1.77245385090551602730 * beta);
    const double e = exp(-0.5 * u * u);
    GhTables& T = tab[axis];
    double hm = 1.0, hc = 2.0 * u;
    for (int n = 0; n <= n_max + 1; ++n) {
      double hn;
      if (n == 0) {
        hn = 1.0;
      } else {
        norm /= sqrt(2.0 * n);
        if (n == 1) {
          hn = hc;
        } else {
          hn = 2.0 * u * hc - 2.0 * (n - 1) * hm;
          hm = hc;
          hc = hn;
        }
      }
      T.g[n][i] = norm * hn * e;
    }
    if (derivs) {
      for (int n = 0; n <= n_max; ++n) {
        const double lo = n > 0 ? sqrt(0.5 * n) * T.g[n - 1][i] : 0.0;
        const double d = (lo - sqrt(0.5 * (n + 1)) * T.g[n + 1][i]) / beta;
        T.dx[n][i] = d;
        T.db[n][i] = -(0.5 * T.g[n][i] + x * d) / beta;
      }
    }
  }
}

// Dense kernel, transposed: kt[kx * KSI + ky] for the instantiated radius RI
// (zero outside the true radius r); 256 outputs per block.
__global__ __launch_bounds__(256) void k_adj_setup(const double* __restrict__ params, int n_max,
                                                   int r, int ri, float* __restrict__ kt) {
  __shared__ GhTables tab[2];
  const int nc = (n_max + 1) * (n_max + 2) / 2;
  fill_tables(tab, params, nc, n_max, r, false);
  __syncthreads();
  const int ksi = 2 * ri + 1, off = ri - r, ks = 2 * r + 1;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ksi * ksi) return;
  const int kx = idx / ksi, ky = idx - kx * ksi;
  const int ix = kx - off, iy = ky - off;
  double k = 0.0;
  if (ix >= 0 && ix < ks && iy >= 0 && iy < ks) {
    int c = 0;
    for (int tot = 0; tot <= n_max; ++tot)
      for (int ny = 0; ny <= tot; ++ny, ++c)
        k += params[c] * (tab[1].g[ny][iy] * tab[0].g[tot - ny][ix]);
  }
  kt[idx] = static_cast<float>(k);
}

// Sum over the 64 lanes of v[i] for every i, scattered so that lane l ends
// with the total of v[l] (butterfly reduce-scatter, 63 exchanges).
template <int M>
__device__ __forceinline__ void reduce_scatter_step(float* v, int lane) {
  const bool up = (lane & M) != 0;
#pragma unroll
  for (int t = 0; t < M; ++t) {
    const float send = up ? v[t] : v[t + M];
    const float keep = up ? v[t + M] : v[t];
    v[t] = keep + __shfl_xor(send, M, 64);
  }
  if constexpr (M > 1) reduce_scatter_step<M / 2>(v, lane);
}

// MODE_CHI2: A = w D from the recomputed model (optionally writing D);
// MODE_VJP: A = the given upstream gradient (no model pass).
enum { MODE_CHI2 = 0, MODE_VJP = 1 };

template <int R, int MODE>
__global__ __launch_bounds__(kAdjThreads) void k_adj_tiles(
    int w, int h, const float* __restrict__ sci, const float* __restrict__ ref,
    const float* __restrict__ weight, const float* __restrict__ upstream,
    const double* __restrict__ params, int nc, const float* __restrict__ kt, int want_corr,
    float* __restrict__ diff, float* __restrict__ part, double* __restrict__ sc) {
  using G = AdjGeom<R>;
  constexpr int KS = G::KS, UY = G::UY, UX = G::UX;
  __shared__ float patch[G::PH][G::PW];
  __shared__ float acoef[kAdjTile][kAdjTile];
  __shared__ double red[2][kAdjThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int x0 = blockIdx.x * kAdjTile, y0 = blockIdx.y * kAdjTile;
  const size_t tile = size_t(blockIdx.y) * gridDim.x + blockIdx.x;

  // reference patch: frame rows y0-R-GY .. y0+T+R-1, zero outside the frame
  for (int py = wv; py < G::PH; py += kAdjThreads / 64) {
    const int fy = y0 - R - G::GY + py;
    for (int px = lane; px < G::PW; px += 64) {
      const int fx = x0 - R - G::GX + px;
      float v = 0.0f;
      if (py >= G::GY && px >= G::GX && fy >= 0 && fy < h && fx >= 0 && fx < w)
        v = ref[size_t(fy) * w + fx];
      patch[py][px] = v;
    }
  }
  double chi = 0.0, suma = 0.0;
  const int ly0 = wv * 16;
  if constexpr (MODE == MODE_CHI2) {
    __syncthreads();
    // model: m(ly, lx) = sum_{ky,kx} K[ky][kx] patch[ly+2R-ky][lx+2R-kx]
    float m[16];
#pragma unroll
    for (int o = 0; o < 16; ++o) m[o] = 0.0f;
    for (int kx = 0; kx < KS; ++kx) {
      float col[16 + 2 * R];
      const int pc = G::GX + lane + 2 * R - kx;
#pragma unroll
      for (int j = 0; j < 16 + 2 * R; ++j) col[j] = patch[G::GY + ly0 + j][pc];
      const float* kcol = kt + kx * KS;
#pragma unroll
      for (int ky = 0; ky < KS; ++ky) {
        const float kv = kcol[ky];
#pragma unroll
        for (int o = 0; o < 16; ++o) m[o] = fmaf(kv, col[o + 2 * R - ky], m[o]);
      }
    }
    const float b = static_cast<float>(params[nc]);
    const int fx = x0 + lane;
#pragma unroll
    for (int o = 0; o < 16; ++o) {
      const int fy = y0 + ly0 + o;
      float a = 0.0f;
      if (fx < w && fy < h) {
        const size_t p = size_t(fy) * w + fx;
        const float d = (sci[p] - b) - m[o];
        if (diff) diff[p] = d;
        const float wt = weight ? weight[p] : 1.0f;
        if (wt > 0.0f) {
          const double wd = double(wt) * double(d);
          chi += wd * double(d);
          suma += wd;
          a = wt * d;
        }
      }
      acoef[ly0 + o][lane] = a;
    }
  } else {
    const int fx = x0 + lane;
#pragma unroll
    for (int o = 0; o < 16; ++o) {
      const int fy = y0 + ly0 + o;
      float a = 0.0f;
      if (fx < w && fy < h) a = upstream[size_t(fy) * w + fx];
      suma += double(a);
      acoef[ly0 + o][lane] = a;
    }
  }
  // tile sums of chi^2 and of A, in a fixed order
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    chi += __shfl_xor(chi, m, 64);
    suma += __shfl_xor(suma, m, 64);
  }
  if (lane == 0) {
    red[0][wv] = chi;
    red[1][wv] = suma;
  }
  __syncthreads();
  if (tid == 0) {
    double c = 0.0, s = 0.0;
    for (int i = 0; i < kAdjThreads / 64; ++i) {
      c += red[0][i];
      s += red[1][i];
    }
    sc[2 * tile] = c;
    sc[2 * tile + 1] = s;
  }
  if (!want_corr) return;

  // correlation C(ky, kx) = sum_{ly,lx} A[ly][lx] patch[ly+2R-ky][lx+2R-kx],
  // lane = lx; blocks of UY x UX lags slide down the tile rows with the
  // patch rows of the block kept in a circular register window
  for (int blk = wv; blk < 16; blk += kAdjThreads / 64) {
    const int ky0 = (blk >> 3) * UY, kx0 = (blk & 7) * UX;
    const int pc = G::GX + lane + 2 * R - kx0;
    float win[UY][UX];
    float acc[UY][UX];
#pragma unroll
    for (int i = 0; i < UY; ++i)
#pragma unroll
      for (int j = 0; j < UX; ++j) acc[i][j] = 0.0f;
    // rows of lags i = 1..UY-1 at ly = 0 (row of lag i: ly + 2R - ky0 - i)
#pragma unroll
    for (int i = 1; i < UY; ++i) {
      const int slot = ((2 * R - i) % UY + UY) % UY;
#pragma unroll
      for (int j = 0; j < UX; ++j) win[slot][j] = patch[G::GY + 2 * R - ky0 - i][pc - j];
    }
    for (int c = 0; c < kAdjTile; c += UY) {
#pragma unroll
      for (int s = 0; s < UY; ++s) {
        const int ly = c + s;
        if ((kAdjTile % UY) != 0 && ly >= kAdjTile) break;
        const int slot0 = (s + 2 * R) % UY;
#pragma unroll
        for (int j = 0; j < UX; ++j) win[slot0][j] = patch[G::GY + ly + 2 * R - ky0][pc - j];
        const float a = acoef[ly][lane];
#pragma unroll
        for (int i = 0; i < UY; ++i) {
          const int sl = ((s + 2 * R - i) % UY + UY) % UY;
#pragma unroll
          for (int j = 0; j < UX; ++j) acc[i][j] = fmaf(a, win[sl][j], acc[i][j]);
        }
      }
    }
    float v[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) v[i] = i < G::NACC ? acc[i / UX][i % UX] : 0.0f;
    reduce_scatter_step<32>(v, lane);
    const float tot = v[0];
    if (lane < G::NACC) {
      const int i = lane / UX, j = lane - (lane / UX) * UX;
      part[tile * G::L + (ky0 + i) * G::LX + kx0 + j] = tot;
    }
  }
}

// First level of the per-lag sums: block (bx, g) adds tiles
// [g * chunk, (g + 1) * chunk) of columns bx*64 .. bx*64+63 of part[ntile][L]
// in FP64 and tile order -> cg[g][L]; the extra column block bx = L/64 does
// the same for the tile (chi^2, sum A) pairs -> sg[g][2].
constexpr int kAdjGroups = 16;

__global__ __launch_bounds__(256) void k_adj_reduce(const float* __restrict__ part,
                                                    const double* __restrict__ sc, int ntile,
                                                    int L, double* __restrict__ cg,
                                                    double* __restrict__ sg) {
  __shared__ double red[4][64];
  const int tid = threadIdx.x, col = tid & 63, sub = tid >> 6;
  const int g = blockIdx.y;
  const int chunk = (ntile + kAdjGroups - 1) / kAdjGroups;
  const int t0 = g * chunk, t1 = min(ntile, t0 + chunk);
  const bool last = blockIdx.x == gridDim.x - 1;
  double s = 0.0;
  if (!last) {
    const int c = blockIdx.x * 64 + col;
    for (int t = t0 + sub; t < t1; t += 4) s += double(part[size_t(t) * L + c]);
  } else if (col < 2) {
    for (int t = t0 + sub; t < t1; t += 4) s += sc[2 * size_t(t) + col];
  }
  red[sub][col] = s;
  __syncthreads();
  if (sub == 0) {
    const double tot = ((red[0][col] + red[1][col]) + red[2][col]) + red[3][col];
    if (!last)
      cg[size_t(g) * L + blockIdx.x * 64 + col] = tot;
    else if (col < 2)
      sg[2 * g + col] = tot;
  }
}

__device__ double block_sum256(double v, double* red) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return ((red[0] + red[1]) + red[2]) + red[3];
}

// Gradient from the group sums cg[kAdjGroups][L] of the lag correlation on
// the padded lag grid (lx columns) of the instantiated radius ri.  Block c <
// nc: out[1 + c] = factor <phi_c, C>; blocks nc, nc+1, nc+2: out[2 + nc + p]
// = factor <dK/dp, C> for p = beta, dx, dy; block nc + 3: out[0] = chi^2
// (when with_chi2) and out[1 + nc] = factor sum A.  factor = -2 for chi^2,
// -1 for a vector-Jacobian product of D.
__global__ __launch_bounds__(256) void k_adj_contract(const double* __restrict__ params,
                                                      int n_max, int r, int ri, int lx, int L,
                                                      const double* __restrict__ cg,
                                                      const double* __restrict__ sg,
                                                      double factor, int with_chi2,
                                                      double* __restrict__ out) {
  __shared__ GhTables tab[2];
  __shared__ double red[4];
  const int nc = (n_max + 1) * (n_max + 2) / 2;
  const int job = blockIdx.x;
  if (job == nc + 3) {
    if (threadIdx.x == 0) {
      double c = 0.0, a = 0.0;
      for (int g = 0; g < kAdjGroups; ++g) {
        c += sg[2 * g];
        a += sg[2 * g + 1];
      }
      out[0] = with_chi2 ? c : 0.0;
      out[1 + nc] = factor * a;
    }
    return;
  }
  fill_tables(tab, params, nc, n_max, r, job >= nc);
  __syncthreads();
  const int ks = 2 * r + 1, off = ri - r;
  int cx = 0, cy = 0;
  if (job < nc) {
    int c = 0;
    for (int tot = 0; tot <= n_max; ++tot)
      for (int ny = 0; ny <= tot; ++ny, ++c)
        if (c == job) {
          cx = tot - ny;
          cy = ny;
        }
  }
  double s = 0.0;
  for (int u = threadIdx.x; u < ks * ks; u += blockDim.x) {
    const int iy = u / ks, ix = u - iy * ks;
    const size_t lag = size_t(iy + off) * lx + ix + off;
    double cu = 0.0;
    for (int g = 0; g < kAdjGroups; ++g) cu += cg[size_t(g) * L + lag];
    double kv = 0.0;
    if (job < nc) {
      kv = tab[1].g[cy][iy] * tab[0].g[cx][ix];
    } else {
      int c = 0;
      for (int tot = 0; tot <= n_max; ++tot)
        for (int ny = 0; ny <= tot; ++ny, ++c) {
          const int nx = tot - ny;
          const double a = params[c];
          if (job == nc)
            kv += a * (tab[1].db[ny][iy] * tab[0].g[nx][ix] +
                       tab[1].g[ny][iy] * tab[0].db[nx][ix]);
          else if (job == nc + 1)
            kv -= a * tab[1].g[ny][iy] * tab[0].dx[nx][ix];
          else
            kv -= a * tab[1].dx[ny][iy] * tab[0].g[nx][ix];
        }
    }
    s += kv * cu;
  }
  const double t = block_sum256(s, red);
  if (threadIdx.x == 0) out[job < nc ? 1 + job : 2 + job] = factor * t;
}

__global__ void k_adj_chi2_only(const double* __restrict__ sg, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double c = 0.0;
    for (int g = 0; g < kAdjGroups; ++g) c += sg[2 * g];
    out[0] = c;
  }
}

int instantiated_radius(int r) {
  if (r <= 7) return 7;
  if (r <= 11) return 11;
  return 15;
}

template <int R>
void launch_tiles(int mode, int w, int h, const float* sci, const float* ref, const float* weight,
                  const float* upstream, const double* params, int nc, const float* kt,
                  int want_corr, float* diff, float* part, double* sc, hipStream_t s) {
  const dim3 grid((w + kAdjTile - 1) / kAdjTile, (h + kAdjTile - 1) / kAdjTile);
  if (mode == MODE_CHI2)
    k_adj_tiles<R, MODE_CHI2><<<grid, kAdjThreads, 0, s>>>(w, h, sci, ref, weight, upstream,
                                                           params, nc, kt, want_corr, diff, part,
                                                           sc);
  else
    k_adj_tiles<R, MODE_VJP><<<grid, kAdjThreads, 0, s>>>(w, h, sci, ref, weight, upstream,
                                                          params, nc, kt, 1, diff, part, sc);
  DG_LAUNCH_CHECK();
}

int lag_cols(int ri) {
  switch (ri) {
    case 7: return AdjGeom<7>::LX;
    case 11: return AdjGeom<11>::LX;
    default: return AdjGeom<15>::LX;
  }
}
int lag_count(int ri) {
  switch (ri) {
    case 7: return AdjGeom<7>::L;
    case 11: return AdjGeom<11>::L;
    default: return AdjGeom<15>::L;
  }
}

}  // namespace

void kernel_model_check(int w, int h, int n_max, int radius) {
  if (w <= 0 || h <= 0) invalid("kernel_chi2: empty frame");
  if (n_max < 0 || n_max > kAdjMaxNmax) invalid("kernel_chi2: n_max must be in 0..10");
  if (radius < 1 || radius > 15) invalid("kernel_chi2: radius must be in 1..15");
}

size_t kernel_model_scratch_bytes(int w, int h, int radius) {
  const int ri = instantiated_radius(radius);
  const size_t ntile = size_t((w + kAdjTile - 1) / kAdjTile) * ((h + kAdjTile - 1) / kAdjTile);
  const int L = lag_count(ri), ksi = 2 * ri + 1;
  return Arena::round(ntile * L * sizeof(float)) + Arena::round(ntile * 2 * sizeof(double)) +
         Arena::round(size_t(kAdjGroups) * L * sizeof(double)) +
         Arena::round(kAdjGroups * 2 * sizeof(double)) +
         Arena::round(size_t(ksi) * ksi * sizeof(float));
}

// One pass of the constant-kernel model.  mode 0: chi^2 with weights (out[0])
// and, when want_grad, its gradient (out[1..nc+4]); D to `diff` if non-null.
// mode 1: vector-Jacobian product of D with `upstream` into out[1..nc+4].
void run_kernel_model(int mode, int w, int h, const float* sci, const float* ref,
                      const float* weight, const float* upstream, const double* params,
                      int n_max, int radius, int want_grad, float* diff, double* out, Arena& ar,
                      hipStream_t s) {
  kernel_model_check(w, h, n_max, radius);
  const int nc = (n_max + 1) * (n_max + 2) / 2;
  const int ri = instantiated_radius(radius);
  const int gx = (w + kAdjTile - 1) / kAdjTile, gy = (h + kAdjTile - 1) / kAdjTile;
  const size_t ntile = size_t(gx) * gy;
  const int L = lag_count(ri), ksi = 2 * ri + 1;
  float* part = ar.take<float>(ntile * L);
  double* sc = ar.take<double>(ntile * 2);
  double* cg = ar.take<double>(size_t(kAdjGroups) * L);
  double* sg = ar.take<double>(kAdjGroups * 2);
  float* kt = ar.take<float>(size_t(ksi) * ksi);
  const bool corr = mode == MODE_VJP || want_grad;
  {
    ProfScope p("kernel_setup", s);
    if (mode == MODE_CHI2) {
      k_adj_setup<<<(ksi * ksi + 255) / 256, 256, 0, s>>>(params, n_max, radius, ri, kt);
      DG_LAUNCH_CHECK();
    }
  }
  {
    ProfScope p("kernel_tiles", s);
    switch (ri) {
      case 7:
        launch_tiles<7>(mode, w, h, sci, ref, weight, upstream, params, nc, kt, corr, diff, part,
                        sc, s);
        break;
      case 11:
        launch_tiles<11>(mode, w, h, sci, ref, weight, upstream, params, nc, kt, corr, diff,
                         part, sc, s);
        break;
      default:
        launch_tiles<15>(mode, w, h, sci, ref, weight, upstream, params, nc, kt, corr, diff,
                         part, sc, s);
        break;
    }
  }
  if (!out) return;
  {
    ProfScope p("kernel_reduce", s);
    const dim3 grid(corr ? L / 64 + 1 : 1, kAdjGroups);
    k_adj_reduce<<<grid, 256, 0, s>>>(part, sc, int(ntile), L, cg, sg);
    DG_LAUNCH_CHECK();
  }
  {
    ProfScope p("kernel_contract", s);
    if (corr) {
      k_adj_contract<<<nc + 4, 256, 0, s>>>(params, n_max, radius, ri, lag_cols(ri), L, cg, sg,
                                            mode == MODE_CHI2 ? -2.0 : -1.0, mode == MODE_CHI2,
                                            out);
    } else {
      k_adj_chi2_only<<<1, 64, 0, s>>>(sg, out);
    }
    DG_LAUNCH_CHECK();
  }
}

}  // namespace dg
