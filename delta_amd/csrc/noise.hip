// Noise decorrelation (whitening) of the difference image and the whitened
// score PSF (reference src/noise.cpp:27-533, bindings.cpp:766-874).
//
// Per n x n block (n a power of two, 8..512) the reference does
// r2c -> multiply by Phi -> c2r, unnormalised, in single precision.  Here one
// workgroup of eight waves does a block as three passes over a per-workgroup
// spectrum scratch S[n][n/2+1] (complex float):
//   rows     two real rows per complex FFT (z = a + ib), split into the two
//            half spectra by Hermitian symmetry;
//   columns  n/2+1 complex column FFTs, staged through LDS in chunks of
//            columns; the filter (or |K|^2) is applied between the forward and
//            the inverse column transform;
//   rows     the two half spectra recombined (z = A + iB, imaginary parts of
//            the DC and Nyquist bins dropped as PocketFFT's c2r does), one
//            inverse complex FFT, real part = row a, imaginary part = row b.
// Every 1-D transform is a wave-level Stockham autosort FFT in LDS (radix 4,
// one radix-2 stage when log2 n is odd), twiddles from a table rounded from
// double.  The same kernel computes the cached kernel power |Khat|^2 of each
// kernel cell (noise.cpp:65-84, 322-343) and the whitened PSF (:485-533).
//
// The overlap-add follows the reference's blend exactly (noise.cpp:346-459):
// one launch per block colour (ix % 3, iy % 3) in its colour order (blocks of
// one colour are disjoint), each pixel accumulating
// acc = float(acc + (hann_y / n^2 * hann_x) f) in double from the block's raw
// inverse transform f, then acc / wsum with the separable window sum; the
// variance blend (a per-block white level times the same weights) comes from
// per-row sums over the covering blocks.  For n = 256 the block kernel runs
// each 256-point FFT on a group of 16 lanes as 16 x 16 in registers
// (k_block_r16); n = 64, 128 keep each wave's half spectra in registers
// (k_block_reg); other sizes stage them through a per-block spectrum scratch
// (k_block_fft).
#include <cmath>

#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kNoiseWaves = 8;
constexpr int kMaxBlock = 512;
constexpr int kChunkComplex = 8192;  // LDS column chunk: n x cw complex floats

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Ordering of LDS traffic between the lanes of one wave.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stockham FFT of n = 2^lg points in LDS buffer a (natural order) using b as
// the second buffer; returns the buffer holding the result (natural order).
// Unnormalised; inv selects exp(+2 pi i jk/n).
__device__ float2* wave_fft(float2* a, float2* b, int n, int lg, const float2* tw, bool inv,
                            int lane) {
  int ns = 1;
  if (lg & 1) {
    for (int j = lane; j < n / 2; j += 64) {
      const float2 v0 = a[j], v1 = a[j + n / 2];
      b[2 * j] = make_float2(v0.x + v1.x, v0.y + v1.y);
      b[2 * j + 1] = make_float2(v0.x - v1.x, v0.y - v1.y);
    }
    float2* t = a;
    a = b;
    b = t;
    ns = 2;
    wave_lds_sync();
  }
  const int q = n / 4;
  for (; ns < n; ns *= 4) {
    const int tstep = n / (4 * ns);
    for (int j = lane; j < q; j += 64) {
      float2 v0 = a[j], v1 = a[j + q], v2 = a[j + 2 * q], v3 = a[j + 3 * q];
      const int jm = j & (ns - 1);
      if (ns > 1) {
        float2 w1 = tw[jm * tstep], w2 = tw[2 * jm * tstep], w3 = tw[3 * jm * tstep];
        if (inv) {
          w1.y = -w1.y;
          w2.y = -w2.y;
          w3.y = -w3.y;
        }
        v1 = cmul(v1, w1);
        v2 = cmul(v2, w2);
        v3 = cmul(v3, w3);
      }
      const float2 s02 = make_float2(v0.x + v2.x, v0.y + v2.y);
      const float2 d02 = make_float2(v0.x - v2.x, v0.y - v2.y);
      const float2 s13 = make_float2(v1.x + v3.x, v1.y + v3.y);
      const float2 d13 = make_float2(v1.x - v3.x, v1.y - v3.y);
      // -i d13 (forward) or +i d13 (inverse)
      const float2 m = inv ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
      const int d = (j - jm) * 4 + jm;
      b[d] = make_float2(s02.x + s13.x, s02.y + s13.y);
      b[d + ns] = make_float2(d02.x + m.x, d02.y + m.y);
      b[d + 2 * ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
      b[d + 3 * ns] = make_float2(d02.x - m.x, d02.y - m.y);
    }
    float2* t = a;
    a = b;
    b = t;
    wave_lds_sync();
  }
  return a;
}

__device__ __forceinline__ int reflect_clamp(int s, int extent) {
  if (s < 0) s = -s - 1;
  if (s >= extent) s = 2 * extent - s - 1;
  return min(max(s, 0), extent - 1);
}

// Source index of grid cell g under the zero-phase embedding g = (i - r) mod n
// of i in [0, ks): the largest such i (the last assigned), or -1.
__device__ __forceinline__ int embed_src(int g, int r, int ks, int n) {
  const int base = (g + r) % n;
  if (base >= ks) return -1;
  return base + n * ((ks - 1 - base) / n);
}

enum : int { kModeDecorrelate = 0, kModePower = 1, kModeWhiten = 2 };

struct BlockJob {
  int bx, by, cell, id, plane;
};

struct NoiseArgs {
  int n, lg, w, h, cw;
  const float2* tw;           // n twiddles exp(-2 pi i k / n)
  const double* hwin;         // n Hann samples
  float2* scratch;            // per workgroup n x (n/2+1)
  // decorrelate
  const float* diff;
  float* planes;              // raw inverse-transform rows (single-block whitening)
  size_t plane_stride;        // w * h
  const BlockJob* jobs;       // this launch's blocks
  // overlap-add into the output (frame decorrelation): the blocks of one
  // launch are disjoint; the first class covering a pixel stores, later ones add
  float* out;                 // w x h accumulator, or nullptr for the raw-plane output
  double inv_n2;              // 1 / n^2 (exact: n is a power of two)
  const unsigned char* fx;    // w: smallest ix % 3 of the blocks covering x
  const unsigned char* fy;    // h: smallest iy % 3 of the blocks covering y
  int nox, xcls, ycls;        // block columns; this launch's colour (ix % 3, iy % 3)
  const float* pwt;           // per cell (n/2+1) x n: |Khat|^2 as float, column-major
  float* pwt_out;             // the same, written by the 256-point power pass
  const double* vs;           // per block id
  const double* vr;
  const double* power;        // per cell n x (n/2+1)
  const double* sumk2;        // per cell
  // kernel power / whiten
  const float* kern;          // per cell ks x ks (power), or the whitening kernel cell 0
  int ks;
  double* power_out;          // per cell n x (n/2+1)
  const float* psf;           // whiten source, ps x ps
  int ps;
  float* psf_out;             // ps x ps, unnormalised
};

// One pixel of a block's raw inverse transform f blended into the
// accumulator as the reference does (noise.cpp:450-459):
// acc = float(acc + (hann(j) / n^2 * hann(i)) f) in double, the launches
// running in its colour order; the first colour covering a pixel starts from
// zero (a plain store).
__device__ __forceinline__ float blend_value(double wy_sig, double wx, float f, float prev) {
  return static_cast<float>(double(prev) + wy_sig * wx * double(f));
}

__device__ __forceinline__ void blend_px(const NoiseArgs& a, int i, int j, int x, int y, float v) {
  const size_t o = size_t(y) * a.w + x;
  const bool first = a.fx[x] == a.xcls && a.fy[y] == a.ycls;
  a.out[o] = blend_value(a.hwin[j] * a.inv_n2, a.hwin[i], v, first ? 0.0f : a.out[o]);
}

template <int MODE>
__global__ __launch_bounds__(64 * kNoiseWaves) void k_block_fft(NoiseArgs a) {
  extern __shared__ float2 nz_lds[];
  const int n = a.n, nh = n / 2 + 1;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  float2* tw = nz_lds;                                  // n
  float2* chunk = tw + n;                               // cw x (n + 1), column-contiguous
  float2* wbuf = chunk + size_t(a.cw) * (n + 1) + size_t(wv) * 2 * n;  // this wave: 2n
  for (int i = tid; i < n; i += 64 * kNoiseWaves) tw[i] = a.tw[i];
  float2* S = a.scratch + size_t(blockIdx.x) * n * nh;
  int bx = 0, by = 0, cell = blockIdx.x, id = 0;
  if (MODE == kModeDecorrelate) {
    const BlockJob j = a.jobs[blockIdx.x];
    bx = j.bx;
    by = j.by;
    cell = j.cell;
    id = j.id;
  }
  __syncthreads();

  // ---- pass 1: forward row transforms, two real rows per complex FFT
  for (int rp = wv; rp < n / 2; rp += kNoiseWaves) {
    const int ra = 2 * rp, rb = ra + 1;
    if (MODE == kModeDecorrelate) {
      const int sya = reflect_clamp(by + ra, a.h), syb = reflect_clamp(by + rb, a.h);
      for (int j = lane; j < n; j += 64) {
        const int sx = reflect_clamp(bx + j, a.w);
        wbuf[j] = make_float2(a.diff[size_t(sya) * a.w + sx], a.diff[size_t(syb) * a.w + sx]);
      }
    } else {
      // zero-phase embedding: grid cell g holds source index (g + r) mod n,
      // or, where ks > n folds several onto one cell, the last of them in
      // the reference's assignment order
      const int ks = MODE == kModePower ? a.ks : a.ps;
      const float* src = MODE == kModePower ? a.kern + size_t(cell) * ks * ks : a.psf;
      const int r = ks / 2;
      const int ga = embed_src(ra, r, ks, n), gb = embed_src(rb, r, ks, n);
      for (int j = lane; j < n; j += 64) {
        const int gx = embed_src(j, r, ks, n);
        const float va = (ga >= 0 && gx >= 0) ? src[ga * ks + gx] : 0.0f;
        const float vb = (gb >= 0 && gx >= 0) ? src[gb * ks + gx] : 0.0f;
        wbuf[j] = make_float2(va, vb);
      }
    }
    wave_lds_sync();
    const float2* z = wave_fft(wbuf, wbuf + n, n, a.lg, tw, false, lane);
    for (int k = lane; k < nh; k += 64) {
      const float2 zk = z[k], zm = z[(n - k) & (n - 1)];
      S[size_t(ra) * nh + k] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
      S[size_t(rb) * nh + k] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
    wave_lds_sync();
  }
  __syncthreads();

  // ---- pass 2: column transforms through LDS chunks
  double vs = 0.0, vr = 0.0, cnum = 0.0;
  const double* pw = nullptr;
  if (MODE != kModePower) {
    const int c = MODE == kModeDecorrelate ? cell : 0;
    vs = a.vs[MODE == kModeDecorrelate ? id : 0];
    vr = a.vr[MODE == kModeDecorrelate ? id : 0];
    cnum = sqrt(fmax(vs + vr * a.sumk2[c], 0.0));
    pw = a.power + size_t(c) * n * nh;
  }
  for (int c0 = 0; c0 < nh; c0 += a.cw) {
    const int cols = min(a.cw, nh - c0);
    for (int e = tid; e < n * cols; e += 64 * kNoiseWaves) {
      const int ky = e / cols, cc = e - ky * cols;
      chunk[size_t(cc) * (n + 1) + ky] = S[size_t(ky) * nh + c0 + cc];
    }
    __syncthreads();
    for (int cc = wv; cc < cols; cc += kNoiseWaves) {
      float2* col = chunk + size_t(cc) * (n + 1);
      float2* f = wave_fft(col, wbuf, n, a.lg, tw, false, lane);
      const int kx = c0 + cc;
      if (MODE == kModePower) {
        double* po = a.power_out + size_t(cell) * n * nh;
        for (int ky = lane; ky < n; ky += 64) {
          const double re = f[ky].x, im = f[ky].y;
          po[size_t(ky) * nh + kx] = re * re + im * im;
        }
        wave_lds_sync();
        continue;
      }
      for (int ky = lane; ky < n; ky += 64) {
        const double den = fma(vr, pw[size_t(ky) * nh + kx], vs);
        const float phi = den > 0.0 ? static_cast<float>(cnum / sqrt(den)) : 0.0f;
        f[ky] = make_float2(f[ky].x * phi, f[ky].y * phi);
      }
      wave_lds_sync();
      float2* other = f == col ? wbuf : col;
      float2* g = wave_fft(f, other, n, a.lg, tw, true, lane);
      if (g != col) {
        for (int ky = lane; ky < n; ky += 64) col[ky] = g[ky];
        wave_lds_sync();
      }
    }
    __syncthreads();
    if (MODE != kModePower) {
      for (int e = tid; e < n * cols; e += 64 * kNoiseWaves) {
        const int ky = e / cols, cc = e - ky * cols;
        S[size_t(ky) * nh + c0 + cc] = chunk[size_t(cc) * (n + 1) + ky];
      }
    }
    __syncthreads();
  }
  if (MODE == kModePower) return;

  // ---- pass 3: inverse row transforms (two rows per complex FFT) and output
  const float inv_n2 = 1.0f / static_cast<float>(n * n);
  for (int rp = wv; rp < n / 2; rp += kNoiseWaves) {
    const int ra = 2 * rp, rb = ra + 1;
    for (int k = lane; k < n; k += 64) {
      const bool lo = k <= n / 2;
      const int kk = lo ? k : n - k;
      float2 A = S[size_t(ra) * nh + kk], B = S[size_t(rb) * nh + kk];
      if (kk == 0 || kk == n / 2) {
        A.y = 0.0f;
        B.y = 0.0f;
      }
      wbuf[k] = lo ? make_float2(A.x - B.y, A.y + B.x) : make_float2(A.x + B.y, B.x - A.y);
    }
    wave_lds_sync();
    const float2* z = wave_fft(wbuf, wbuf + n, n, a.lg, tw, true, lane);
    if (MODE == kModeDecorrelate) {
      const int ya = by + ra, yb = by + rb;
      if (a.out) {
        for (int j = lane; j < n; j += 64) {
          const int x = bx + j;
          if (x >= a.w) continue;
          if (ya < a.h) blend_px(a, j, ra, x, ya, z[j].x);
          if (yb < a.h) blend_px(a, j, rb, x, yb, z[j].y);
        }
      } else {
        float* plane = a.planes + size_t(a.jobs[blockIdx.x].plane) * a.plane_stride;
        for (int j = lane; j < n; j += 64) {
          const int x = bx + j;
          if (x >= a.w) continue;
          if (ya < a.h) plane[size_t(ya) * a.w + x] = z[j].x;
          if (yb < a.h) plane[size_t(yb) * a.w + x] = z[j].y;
        }
      }
    } else {
      // centre crop of the zero-phase grid: output (py, px) <- ((py - r) mod n, ...)
      const int ps = a.ps, r = ps / 2;
      const int pa = (ra + r) % n, pb = (rb + r) % n;  // output rows of ra, rb (if < ps)
      for (int j = lane; j < n; j += 64) {
        const int px = (j + r) % n;
        if (px >= ps) continue;
        if (pa < ps) a.psf_out[pa * ps + px] = z[j].x * inv_n2;
        if (pb < ps) a.psf_out[pb * ps + px] = z[j].y * inv_n2;
      }
    }
    wave_lds_sync();
  }
}

// Register-resident decorrelation block for n in {64, 128}: wave w owns
// the row pairs w, w+8, ...; after the row transforms each lane keeps the half
// spectra of its rows at kx = lane + 64 m in registers (the Nyquist column in
// LDS).  The column transforms go through LDS 32 columns at a time (the
// Nyquist column with the last chunk); no global scratch.  Output: the raw
// inverse-transform rows blended into the accumulator.
template <int N>
__global__ __launch_bounds__(64 * kNoiseWaves) void k_block_reg(NoiseArgs a) {
  constexpr int HALF = N / 2, NH = HALF + 1;
  constexpr int NP = N / 16;                  // row pairs per wave
  constexpr int SL = (HALF + 63) / 64;        // register slots per row
  constexpr int CW = HALF < 32 ? HALF : 32;   // columns per LDS chunk
  constexpr int NCH = HALF / CW;
  constexpr int LG = N == 256 ? 8 : (N == 128 ? 7 : 6);
  extern __shared__ float2 rz_lds[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  float2* tw = rz_lds;                    // N
  float2* nyq = tw + N;                   // N: column kx = N/2
  float2* col = nyq + N;                  // CW x (N + 1)
  float2* wb = col + CW * (N + 1) + wv * 4 * N;  // this wave: 4N
  for (int i = tid; i < N; i += 64 * kNoiseWaves) tw[i] = a.tw[i];
  const BlockJob job = a.jobs[blockIdx.x];
  __syncthreads();
  float2 xa[NP][SL], xb[NP][SL];

  // ---- rows: two real rows per complex FFT, half spectra into registers
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int ra = 2 * (wv + 8 * j), rb = ra + 1;
    const int sya = reflect_clamp(job.by + ra, a.h), syb = reflect_clamp(job.by + rb, a.h);
    for (int c = lane; c < N; c += 64) {
      const int sx = reflect_clamp(job.bx + c, a.w);
      wb[c] = make_float2(a.diff[size_t(sya) * a.w + sx], a.diff[size_t(syb) * a.w + sx]);
    }
    wave_lds_sync();
    const float2* z = wave_fft(wb, wb + N, N, LG, tw, false, lane);
#pragma unroll
    for (int m = 0; m < SL; ++m) {
      const int k = lane + 64 * m;
      if (k < HALF) {
        const float2 zk = z[k], zm = z[(N - k) & (N - 1)];
        xa[j][m] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        xb[j][m] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
      }
    }
    if (lane == 0) {
      const float2 zk = z[HALF];
      nyq[ra] = make_float2(0.5f * (zk.x + zk.x), 0.5f * (zk.y - zk.y));
      nyq[rb] = make_float2(0.5f * (zk.y + zk.y), -0.5f * (zk.x - zk.x));
    }
    wave_lds_sync();
  }
  __syncthreads();

  // ---- columns through LDS chunks: forward, Phi, inverse
  const double vs = a.vs[job.id], vr = a.vr[job.id];
  const double cnum = sqrt(fmax(vs + vr * a.sumk2[job.cell], 0.0));
  const double* pw = a.power + size_t(job.cell) * N * NH;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int c0 = ch * CW;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int ra = 2 * (wv + 8 * j);
#pragma unroll
      for (int m = 0; m < SL; ++m) {
        const int k = lane + 64 * m;
        if (k >= c0 && k < c0 + CW) {
          col[(k - c0) * (N + 1) + ra] = xa[j][m];
          col[(k - c0) * (N + 1) + ra + 1] = xb[j][m];
        }
      }
    }
    __syncthreads();
    const int ncol = CW + (ch == NCH - 1 ? 1 : 0);
    for (int cc = wv; cc < ncol; cc += kNoiseWaves) {
      float2* c = cc < CW ? col + cc * (N + 1) : nyq;
      const int kx = cc < CW ? c0 + cc : HALF;
      float2* f = wave_fft(c, wb, N, LG, tw, false, lane);
      for (int ky = lane; ky < N; ky += 64) {
        const double den = fma(vr, pw[size_t(ky) * NH + kx], vs);
        const float phi = den > 0.0 ? static_cast<float>(cnum / sqrt(den)) : 0.0f;
        f[ky] = make_float2(f[ky].x * phi, f[ky].y * phi);
      }
      wave_lds_sync();
      float2* g = wave_fft(f, f == c ? wb : c, N, LG, tw, true, lane);
      if (g != c) {
        for (int ky = lane; ky < N; ky += 64) c[ky] = g[ky];
        wave_lds_sync();
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      const int ra = 2 * (wv + 8 * j);
#pragma unroll
      for (int m = 0; m < SL; ++m) {
        const int k = lane + 64 * m;
        if (k >= c0 && k < c0 + CW) {
          xa[j][m] = col[(k - c0) * (N + 1) + ra];
          xb[j][m] = col[(k - c0) * (N + 1) + ra + 1];
        }
      }
    }
    __syncthreads();
  }

  // ---- rows: recombine, inverse FFT, weighted rows into the output
  float2* sa = wb;            // Xa[0..HALF]
  float2* sb = wb + NH;       // Xb[0..HALF]
  float2* zb = wb + 2 * N;    // N, with wb + 3N as the second FFT buffer
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int ra = 2 * (wv + 8 * j), rb = ra + 1;
#pragma unroll
    for (int m = 0; m < SL; ++m) {
      const int k = lane + 64 * m;
      if (k < HALF) {
        sa[k] = xa[j][m];
        sb[k] = xb[j][m];
      }
    }
    if (lane == 0) {
      sa[HALF] = nyq[ra];
      sb[HALF] = nyq[rb];
    }
    wave_lds_sync();
    for (int k = lane; k < N; k += 64) {
      const bool lo = k <= HALF;
      const int kk = lo ? k : N - k;
      float2 A = sa[kk], B = sb[kk];
      if (kk == 0 || kk == HALF) {
        A.y = 0.0f;
        B.y = 0.0f;
      }
      zb[k] = lo ? make_float2(A.x - B.y, A.y + B.x) : make_float2(A.x + B.y, B.x - A.y);
    }
    wave_lds_sync();
    const float2* z = wave_fft(zb, wb + 3 * N, N, LG, tw, true, lane);
    const int ya = job.by + ra, yb = job.by + rb;
    for (int c = lane; c < N; c += 64) {
      const int x = job.bx + c;
      if (x >= a.w) continue;
      if (ya < a.h) blend_px(a, c, ra, x, ya, z[c].x);
      if (yb < a.h) blend_px(a, c, rb, x, yb, z[c].y);
    }
    wave_lds_sync();
  }
}

// ------------------------------------------------------------------ 256-point blocks
// Each 256-point FFT runs on a group of 16 lanes (four per wave) that holds it
// as x[t + 16 m] (lane t, slot m).  With k = k1 + 16 k2,
//   X[k1 + 16 k2] = sum_t w16^(t k2) [w256^(t k1) sum_m x[t + 16 m] w16^(m k1)],
// i.e. a 16-point DFT in registers, a twiddle, a 16 x 16 transpose through a
// padded LDS tile, and a second 16-point DFT; the result lands in the same
// layout (lane t holds X[t + 16 k2] in slot k2).  Twiddles w256 come from a
// table rounded from double.
constexpr int kR16Tr = 17;        // padded row of the 16 x 16 transpose tile
constexpr int kR16Region = 273;   // complex per group region (odd stride: bank spread)

// Whether slot m can hold the mirror Z[256 - kx] of a column kx in [c0, c0 + 32)
// for some lane: k = t + 16 m with k = 0 (kx = 0) or k > 128 (kx = 256 - k).
__host__ __device__ constexpr bool mirror_slot(int m, int c0) {
  return (m == 0 && c0 == 0) || (m >= 8 && 256 - 16 * m >= c0 && 241 - 16 * m < c0 + 32);
}

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// radix-4 DFT in place: a_k <- sum_q a_q w4^(q k), w4 = -i (forward) or +i
template <bool INV>
__device__ __forceinline__ void r4(float2& a0, float2& a1, float2& a2, float2& a3) {
  const float2 s02 = c_add(a0, a2), d02 = c_sub(a0, a2);
  const float2 s13 = c_add(a1, a3), d13 = c_sub(a1, a3);
  const float2 m = INV ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
  a0 = c_add(s02, s13);
  a1 = c_add(d02, m);
  a2 = c_sub(s02, s13);
  a3 = c_sub(d02, m);
}

// x <- x * w16^E (forward w16 = exp(-2 pi i / 16); the inverse conjugates)
template <bool INV, int E>
__device__ __forceinline__ void tw16(float2& x) {
  constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f;
  constexpr float R = 0.70710678118654752f;
  if (E == 4) {
    x = INV ? make_float2(-x.y, x.x) : make_float2(x.y, -x.x);
  } else if (E == 2) {
    x = INV ? make_float2(R * (x.x - x.y), R * (x.x + x.y))
            : make_float2(R * (x.x + x.y), R * (x.y - x.x));
  } else if (E == 6) {
    x = INV ? make_float2(-R * (x.x + x.y), R * (x.x - x.y))
            : make_float2(R * (x.y - x.x), -R * (x.x + x.y));
  } else {
    const float c = E == 1 ? C1 : (E == 3 ? S1 : -C1);
    const float sf = E == 1 ? -S1 : (E == 3 ? -C1 : S1);
    const float s = INV ? -sf : sf;
    x = make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
  }
}

// 16-point DFT in registers, natural order in and out, as radix 4 x 4
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int j = 0; j < 4; ++j) r4<INV>(v[j], v[4 + j], v[8 + j], v[12 + j]);
  // v[4 k1 + j] = Y[j][k1]; twiddle by w16^(j k1)
  tw16<INV, 1>(v[5]);
  tw16<INV, 2>(v[9]);
  tw16<INV, 3>(v[13]);
  tw16<INV, 2>(v[6]);
  tw16<INV, 4>(v[10]);
  tw16<INV, 6>(v[14]);
  tw16<INV, 3>(v[7]);
  tw16<INV, 6>(v[11]);
  tw16<INV, 9>(v[15]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) r4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
  // v[4 k1 + k2] = X[k1 + 4 k2]
  float2 o[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = o[i];
}

// 256-point FFT of the group's x[t + 16 m] (slot m of lane t), in place;
// tr is the group's LDS tile (16 x kR16Tr complex), tw the 256 twiddles.
template <bool INV>
__device__ __forceinline__ void fft256_g16(float2 (&v)[16], float2* tr, const float2* tw, int t) {
  dft16<INV>(v);
  // an opaque copy of the table pointer: each transform reads its fifteen
  // twiddles again instead of the compiler keeping them live in registers
  asm volatile("" : "+v"(tw));
#pragma unroll
  for (int k = 1; k < 16; ++k) {
    float2 w = tw[t * k];
    if (INV) w.y = -w.y;
    v[k] = cmul(v[k], w);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) tr[k * kR16Tr + t] = v[k];
  wave_lds_sync();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = tr[t * kR16Tr + k];
  wave_lds_sync();
  dft16<INV>(v);
}

// Whitens one 256 x 256 block per workgroup (eight waves, 32 groups of 16
// lanes) and overlap-adds it into the output.
//   rows     group gi = 4 wave + g takes the row pairs p = 32 r + gi (rounds
//            r = 0..3) as complex rows z = row 2p + i row 2p+1; the spectra
//            Z_p[k] of rounds 0-2 stay in registers (lane t, slot m:
//            k = t + 16 m), those of round 3 in an LDS tile (row p - 96);
//   columns  in four chunks of 32 columns kx (and the Nyquist column alone):
//            the register owners stage Z_p[kx] and Z_p[256 - kx] in LDS,
//            group gi splits them (and the tile's) into the half spectra of
//            rows 2p and 2p+1, runs the forward column FFT, applies
//            Phi = cnum / sqrt(vs + vr |Khat|^2) and the inverse column FFT;
//            the filtered columns go back as Z'_p[k] (the imaginary parts of
//            the DC and Nyquist bins dropped, as a c2r transform does);
//   rows     inverse row FFTs: real part row 2p, imaginary part row 2p+1.
constexpr int kR16Tile = 257;     // padded row of the round-3 spectrum tile

template <int MODE>
__global__ __launch_bounds__(512, 1) void k_block_r16(NoiseArgs a) {
  constexpr int N = 256;
  constexpr bool kPower = MODE == kModePower;
  extern __shared__ float2 r16_lds[];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int g = lane >> 4, t = lane & 15, gi = wv * 4 + g;
  float2* tw = r16_lds;                          // 256
  float2* reg = tw + N;                          // 32 x kR16Region
  float2* tile = reg + 32 * kR16Region;          // 32 x kR16Tile: round-3 spectra
  double* hw = reinterpret_cast<double*>(tile + 32 * kR16Tile);  // 256 Hann samples
  float2* mine = reg + gi * kR16Region;
  for (int i = tid; i < N; i += 512) {
    tw[i] = a.tw[i];
    if (!kPower) hw[i] = a.hwin[i];
  }
  const BlockJob job = kPower ? BlockJob{0, 0, int(blockIdx.x), 0, 0} : a.jobs[blockIdx.x];
  __syncthreads();
  float2 z[3][16];

  // ---- forward row transforms (the next round's rows load during this one's)
  auto load_rows = [&](int r, float2 (&v)[16]) {
    const int p = 32 * r + gi;
    if (kPower) {
      // zero-phase embedding of the cell's kernel: grid cell g holds source
      // index (g + r) mod n (the last of several when ks > n)
      const int ks = a.ks, rk = ks / 2;
      const float* src = a.kern + size_t(job.cell) * ks * ks;
      const int ga = embed_src(2 * p, rk, ks, N), gb = embed_src(2 * p + 1, rk, ks, N);
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int gx = embed_src(t + 16 * m, rk, ks, N);
        v[m] = make_float2((ga >= 0 && gx >= 0) ? src[ga * ks + gx] : 0.0f,
                           (gb >= 0 && gx >= 0) ? src[gb * ks + gx] : 0.0f);
      }
      return;
    }
    const int sya = reflect_clamp(job.by + 2 * p, a.h), syb = reflect_clamp(job.by + 2 * p + 1, a.h);
    const float* ra = a.diff + size_t(sya) * a.w;
    const float* rb = a.diff + size_t(syb) * a.w;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int sx = reflect_clamp(job.bx + t + 16 * m, a.w);
      v[m] = make_float2(ra[sx], rb[sx]);
    }
  };
  float2 nxt[16];
  load_rows(0, nxt);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = nxt[m];
    if (r < 3) load_rows(r + 1, nxt);
    fft256_g16<false>(v, mine, tw, t);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      if (r < 3)
        z[r < 3 ? r : 0][m] = v[m];
      else
        tile[gi * kR16Tile + t + 16 * m] = v[m];
    }
    __builtin_amdgcn_sched_barrier(0);
  }

  double vs = 0.0, vr = 0.0, cnum = 0.0;
  if (!kPower) {
    vs = a.vs[job.id];
    vr = a.vr[job.id];
    cnum = sqrt(fmax(vs + vr * a.sumk2[job.cell], 0.0));
  }
  const float* pw = a.pwt + size_t(job.cell) * (N / 2 + 1) * N;
  float* pw_out = a.pwt_out + size_t(job.cell) * (N / 2 + 1) * N;

  // ---- column chunks: kx in [c0, c0 + 32), then the Nyquist column
#pragma unroll
  for (int ch = 0; ch < 5; ++ch) {
    const int c0 = 32 * ch;
    const int ncols = ch < 4 ? 32 : 1;
    float pwc[16];
    if (!kPower && gi < ncols) {
      const float* pc = pw + size_t(c0 + gi) * N;
#pragma unroll
      for (int m = 0; m < 16; ++m) pwc[m] = pc[t + 16 * m];
    }
    __syncthreads();
    // stage Z_p[kx] at [kx - c0][p] and Z_p[256 - kx] at [kx - c0][128 + p] (p < 96)
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int p = 32 * r + gi;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int j = t + 16 * m;
        if (ch < 4) {
          if (16 * m + 15 >= c0 && 16 * m < c0 + 32 && j >= c0 && j < c0 + 32)
            reg[(j - c0) * kR16Region + p] = z[r][m];
          const int km = (256 - j) & 255;
          if (mirror_slot(m, c0) && (j == 0 || j > 128) && km >= c0 && km < c0 + 32)
            reg[(km - c0) * kR16Region + 128 + p] = z[r][m];
        } else if (m == 8 && t == 0) {
          reg[p] = z[r][m];
          reg[128 + p] = z[r][m];
        }
      }
    }
    __syncthreads();
    if (gi < ncols) {
      const int kx = c0 + gi, kxm = (256 - kx) & 255;
      if (!kPower) {
        // Phi = float(cnum / sqrt(vs + vr |Khat|^2)) as the reference rounds
        // it: the float reciprocal square root refined by one Newton step in
        // double; formed before the column is loaded, so that its double
        // temporaries and the transforms' registers are not live together
        // (four at a time: scheduled further ahead, the double temporaries spill)
#pragma unroll
        for (int m0 = 0; m0 < 16; m0 += 4) {
#pragma unroll
          for (int m = m0; m < m0 + 4; ++m) {
            const double den = fma(vr, double(pwc[m]), vs);
            float phi = 0.0f;
            if (den > 0.0) {
              const float df = static_cast<float>(den);
              double rs;
              if (df >= 1.17549435e-38f && df <= 3.40282347e38f) {
                rs = double(rsqrtf(df));
                rs = fma(0.5 * rs, fma(-den * rs, rs, 1.0), rs);
              } else {
                rs = 1.0 / sqrt(den);
              }
              phi = static_cast<float>(cnum * rs);
            }
            pwc[m] = phi;
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      float2 f[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int p = (t + 16 * m) >> 1;
        float2 zc, zm;
        if (m < 12) {
          zc = mine[p];
          zm = mine[128 + p];
        } else {
          zc = tile[(p - 96) * kR16Tile + kx];
          zm = tile[(p - 96) * kR16Tile + kxm];
        }
        f[m] = (t & 1) ? make_float2(0.5f * (zc.y + zm.y), -0.5f * (zc.x - zm.x))
                       : make_float2(0.5f * (zc.x + zm.x), 0.5f * (zc.y - zm.y));
      }
      wave_lds_sync();
      fft256_g16<false>(f, mine, tw, t);
      if (kPower) {
        // |Khat|^2 column-major: [kx][ky]
#pragma unroll
        for (int m = 0; m < 16; ++m)
          pw_out[size_t(kx) * N + t + 16 * m] = f[m].x * f[m].x + f[m].y * f[m].y;
      } else {
#pragma unroll
      for (int m = 0; m < 16; ++m) f[m] = make_float2(f[m].x * pwc[m], f[m].y * pwc[m]);
      fft256_g16<true>(f, mine, tw, t);
#pragma unroll
      for (int m = 0; m < 16; ++m) mine[t + 16 * m] = f[m];
      wave_lds_sync();
      // the tile's rows (p >= 96): Z'_p[kx] from the even lane of each pair,
      // Z'_p[256 - kx] from the odd one
#pragma unroll
      for (int m = 12; m < 16; ++m) {
        const int p = (t + 16 * m) >> 1;
        float2 A = mine[2 * p], B = mine[2 * p + 1];
        if (kx == 0 || kx == 128) {
          A.y = 0.0f;
          B.y = 0.0f;
        }
        if ((t & 1) == 0)
          tile[(p - 96) * kR16Tile + kx] = make_float2(A.x - B.y, A.y + B.x);
        else if (kx != 0 && kx != 128)
          tile[(p - 96) * kR16Tile + kxm] = make_float2(A.x + B.y, B.x - A.y);
      }
      }
    }
    __syncthreads();
    if (kPower) continue;  // uniform: every thread has passed this chunk's barriers
    // read back Z'_p[k] (p < 96): A' = column[2p], B' = column[2p + 1]
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int p = 32 * r + gi;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int j = t + 16 * m;
        if (ch < 4) {
          if (16 * m + 15 >= c0 && 16 * m < c0 + 32 && j >= c0 && j < c0 + 32) {
            const float2* col = reg + (j - c0) * kR16Region;
            float2 A = col[2 * p], B = col[2 * p + 1];
            if (j == 0) {
              A.y = 0.0f;
              B.y = 0.0f;
            }
            z[r][m] = make_float2(A.x - B.y, A.y + B.x);
          }
          const int km = 256 - j;
          if (mirror_slot(m, c0) && j > 128 && km >= c0 && km < c0 + 32) {
            const float2* col = reg + (km - c0) * kR16Region;
            const float2 A = col[2 * p], B = col[2 * p + 1];
            z[r][m] = make_float2(A.x + B.y, B.x - A.y);
          }
        } else if (m == 8 && t == 0) {
          z[r][m] = make_float2(reg[2 * p].x, reg[2 * p + 1].x);
        }
      }
    }
  }
  if (kPower) return;
  __syncthreads();

  // ---- inverse row transforms and the overlap-add into the accumulator
  unsigned firstx = 0;
#pragma unroll
  for (int m = 0; m < 16; ++m) {
    const int x = job.bx + t + 16 * m;
    if (x < a.w && a.fx[x] == a.xcls) firstx |= 1u << m;
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    __builtin_amdgcn_sched_barrier(0);
    const int p = 32 * r + gi;
    // the earlier colours' sums under this round's two rows, fetched before
    // the transform (pixels this colour covers first are not read)
    float prev[2][16];
    bool first[2];
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int y = min(job.by + 2 * p + h2, a.h - 1);
      first[h2] = a.fy[y] == a.ycls;
      const float* orow = a.out + size_t(y) * a.w;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int x = job.bx + t + 16 * m;
        prev[h2][m] = 0.0f;
        if (x < a.w && !(first[h2] && ((firstx >> m) & 1u))) prev[h2][m] = orow[x];
      }
    }
    float2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m)
      v[m] = r < 3 ? z[r < 3 ? r : 0][m] : tile[gi * kR16Tile + t + 16 * m];
    fft256_g16<true>(v, mine, tw, t);
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int j = 2 * p + h2, y = job.by + j;
      if (y >= a.h) continue;
      const double wy_sig = hw[j] * a.inv_n2;
      float* orow = a.out + size_t(y) * a.w;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int x = job.bx + t + 16 * m;
        if (x >= a.w) continue;
        orow[x] = blend_value(wy_sig, hw[t + 16 * m], h2 ? v[m].y : v[m].x, prev[h2][m]);
      }
    }
  }
}

constexpr size_t r16_lds_bytes() {
  return sizeof(float2) * (256 + 32 * size_t(kR16Region) + 32 * size_t(kR16Tile)) +
         sizeof(double) * 256;
}

// ------------------------------------------------------------------ block noise levels
// Median of the finite positive variance samples of a block on the stride
// max(1, n/64) sub-lattice (reference block_variance, noise.cpp:200-224): the
// sample of rank count/2 in ascending order, as nth_element returns it.  The
// bit patterns of positive floats order like the floats, so a radix select
// over four 8-bit digits (most significant first) finds it: per digit, a
// 256-bin histogram of the samples that share the prefix found so far, and a
// scan that picks the bin holding the remaining rank.
constexpr int kVarSamples = 4096;
constexpr int kNoiseThreads = 256;
constexpr int kKeysPerThread = kVarSamples / kNoiseThreads;
constexpr unsigned kNoKey = 0xffffffffu;

__global__ __launch_bounds__(kNoiseThreads) void k_block_noise(
    int n, int w, int h, const int* __restrict__ bxy, const float* __restrict__ var_s,
    const float* __restrict__ var_r, double* __restrict__ vs, double* __restrict__ vr) {
  __shared__ unsigned hist[256];
  __shared__ unsigned wtot[kNoiseThreads / 64];
  __shared__ int cnt;
  __shared__ unsigned s_digit, s_rank;
  const int b = blockIdx.x, which = blockIdx.y, tid = threadIdx.x;
  const float* var = which == 0 ? var_s : var_r;
  const int bx = bxy[2 * b], by = bxy[2 * b + 1];
  const int step = max(1, n / 64);
  const int m = (n + step - 1) / step;  // samples per axis
  if (tid == 0) cnt = 0;
  unsigned key[kKeysPerThread];
  int local = 0;
#pragma unroll
  for (int q = 0; q < kKeysPerThread; ++q) {
    const int e = tid + kNoiseThreads * q;
    key[q] = kNoKey;
    if (e < m * m) {
      const int y = by + (e / m) * step, x = bx + (e % m) * step;
      if (y >= 0 && y < h && x >= 0 && x < w) {
        const float v = var[size_t(y) * w + x];
        if (isfinite(v) && v > 0.0f) {
          key[q] = __float_as_uint(v);
          ++local;
        }
      }
    }
  }
  __syncthreads();
  atomicAdd(&cnt, local);
  __syncthreads();
  const int count = cnt;
  if (count == 0) {
    if (tid == 0) (which == 0 ? vs : vr)[b] = 0.0;
    return;
  }
  unsigned rank = unsigned(count / 2), prefix = 0;
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 24 - 8 * pass;
    hist[tid] = 0;
    __syncthreads();
    // Samples of one block are close in value, so most lanes of a wave share
    // a digit: a wave whose matching lanes all agree adds them in one atomic.
#pragma unroll
    for (int q = 0; q < kKeysPerThread; ++q) {
      const unsigned k = key[q];
      const bool match = k != kNoKey && (pass == 0 || (k >> (shift + 8)) == prefix);
      const unsigned d = (k >> shift) & 255u;
      const unsigned long long act = __ballot(match);
      if (act) {
        const int first = __ffsll(static_cast<long long>(act)) - 1;
        const unsigned d0 = __shfl(d, first);
        if (__all(!match || d == d0)) {
          if ((tid & 63) == first) atomicAdd(&hist[d0], unsigned(__popcll(act)));
        } else if (match) {
          atomicAdd(&hist[d], 1u);
        }
      }
    }
    __syncthreads();
    // inclusive prefix sums of the 256 bins: each wave scans its 64 bins
    // with lane shuffles, then adds the totals of the waves before it
    const unsigned hv = hist[tid];
    unsigned wsum = hv;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const unsigned o = __shfl_up(wsum, off);
      if ((tid & 63) >= off) wsum += o;
    }
    if ((tid & 63) == 63) wtot[tid >> 6] = wsum;
    __syncthreads();
    unsigned before = 0;
    for (int q = 0; q < (tid >> 6); ++q) before += wtot[q];
    const unsigned incl = before + wsum, excl = incl - hv;
    if (excl <= rank && rank < incl) {
      s_digit = unsigned(tid);
      s_rank = rank - excl;
    }
    __syncthreads();
    prefix = (prefix << 8) | s_digit;
    rank = s_rank;
    __syncthreads();
  }
  if (tid == 0) (which == 0 ? vs : vr)[b] = double(__uint_as_float(prefix));
}

// ------------------------------------------------------------------ local kernels
// K_cell = sum_n float(a_n(cell)) phi_n with fused multiply-adds in component
// order (noise.cpp:167-189), and sum K^2 in double in raster order.
// (the taps are kept in LDS too, so that the one thread forming sum K^2 in
// raster order reads them from there)
constexpr int kLocalPer = 4;        // taps per thread and pass
constexpr int kLocalMaxComp = 1024;  // components (n_max <= 43)
__global__ __launch_bounds__(256) void k_local_kernels(int ncell, int nc, int ks2,
                                                      const double* __restrict__ a,
                                                      const float* __restrict__ phi,
                                                      float* __restrict__ kern,
                                                      double* __restrict__ sumk2) {
  extern __shared__ float lk_taps[];
  __shared__ float af[kLocalMaxComp];
  const int c = blockIdx.x;
  for (int f = threadIdx.x; f < nc; f += 256) af[f] = static_cast<float>(a[size_t(f) * ncell + c]);
  __syncthreads();
  // up to kLocalPer taps per thread, the footprint loads of eight components
  // for them issued together; each tap sums the components in order
  for (int i0 = 0; i0 < ks2; i0 += 256 * kLocalPer) {
    float v[kLocalPer];
#pragma unroll
    for (int j = 0; j < kLocalPer; ++j) v[j] = 0.0f;
    for (int f0 = 0; f0 < nc; f0 += 8) {
      float pv[8][kLocalPer];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int j = 0; j < kLocalPer; ++j) {
          const int i = i0 + threadIdx.x + 256 * j;
          pv[u][j] = (f0 + u < nc && i < ks2) ? phi[size_t(f0 + u) * ks2 + i] : 0.0f;
        }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (f0 + u >= nc) break;
#pragma unroll
        for (int j = 0; j < kLocalPer; ++j) v[j] = fmaf(af[f0 + u], pv[u][j], v[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kLocalPer; ++j) {
      const int i = i0 + threadIdx.x + 256 * j;
      if (i < ks2) {
        kern[size_t(c) * ks2 + i] = v[j];
        lk_taps[i] = v[j];
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll 8
    for (int i = 0; i < ks2; ++i) {
      const double v = lk_taps[i];
      s = fma(v, v, s);
    }
    sumk2[c] = s;
  }
}

// per-block white-noise level num = cnum^2, cnum = sqrt(max(vs + vr sum K^2, 0))
__global__ void k_block_num(int nblk, const int* __restrict__ cell, const double* __restrict__ vs,
                            const double* __restrict__ vr, const double* __restrict__ sumk2,
                            double* __restrict__ num) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const double cn = sqrt(fmax(vs[b] + vr[b] * sumk2[cell[b]], 0.0));
  num[b] = cn * cn;
}

// ------------------------------------------------------------------ normalise
// acc / wsum with the separable window sum (noise.cpp:465-480), in place.
// Each thread takes four consecutive pixels of kNormRows rows (one 16-byte
// access per row when the rows are 16-byte aligned).
constexpr int kNormRows = 8;

__device__ __forceinline__ float norm_px(float a, double wsum) {
  return wsum > 0.0 ? static_cast<float>(double(a) / wsum) : 0.0f;
}

__global__ __launch_bounds__(256) void k_noise_normalise(int w, int h,
                                                         const double* __restrict__ wxt,
                                                         const double* __restrict__ wyt,
                                                         float* __restrict__ acc, int vec) {
  const int x0 = 4 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (x0 >= w) return;
  double wx[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) wx[j] = x0 + j < w ? wxt[x0 + j] : 0.0;
  for (int r = 0; r < kNormRows; ++r) {
    const int y = blockIdx.y * kNormRows + r;
    if (y >= h) break;
    const double wy = wyt[y];
    float* row = acc + size_t(y) * w;
    if (vec && x0 + 4 <= w) {
      float4 v = *reinterpret_cast<const float4*>(row + x0);
      v.x = norm_px(v.x, wy * wx[0]);
      v.y = norm_px(v.y, wy * wx[1]);
      v.z = norm_px(v.z, wy * wx[2]);
      v.w = norm_px(v.w, wy * wx[3]);
      *reinterpret_cast<float4*>(row + x0) = v;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (x0 + j < w) row[x0 + j] = norm_px(row[x0 + j], wy * wx[j]);
    }
  }
}

// ------------------------------------------------------------------ variance
// A covering block of a pixel row or column, its window weight already
// divided by the window sum of that row or column.
struct VarTab {
  float wgt;
  int blk;
};

struct VarArgs {
  int w, h, nox;
  const VarTab* xt;     // w x 3
  const VarTab* yt;     // h x 3
  const int* xlo;       // per tile column: the first block column covering it
  const double* num;    // per block (iy * nox + ix): vs + vr sum K^2
  float* var;
};

// The whitened difference's variance: the window-weighted blend of the
// covering blocks' white levels over the window sum (noise.cpp:461-482;
// the weight sum is separable).  A workgroup takes kVarRows rows by 256
// columns: per row it first forms, for every block column covering the tile,
// S = sum over the row's covering blocks of (row weight) x (white level), so
// that a pixel adds at most three (column weight) x S terms.
constexpr int kVarRows = 64;

__global__ __launch_bounds__(256) void k_noise_variance(VarArgs a, int nbx) {
  extern __shared__ float var_lds[];  // kVarRows x nbx
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y0 = blockIdx.y * kVarRows;
  const int rows = min(kVarRows, a.h - y0);
  const int ix0 = a.xlo[blockIdx.x];
  for (int e = threadIdx.x; e < rows * nbx; e += blockDim.x) {
    const int r = e / nbx, ix = ix0 + e - r * nbx;
    float acc = 0.0f;
    if (ix < a.nox) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const VarTab t = a.yt[3 * (y0 + r) + c];
        if (t.blk >= 0) acc = fmaf(t.wgt, float(a.num[size_t(t.blk) * a.nox + ix]), acc);
      }
    }
    var_lds[e] = acc;
  }
  __syncthreads();
  if (x >= a.w) return;
  VarTab tx[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    tx[c] = a.xt[3 * x + c];
    tx[c].blk = tx[c].blk >= 0 ? tx[c].blk - ix0 : -1;
  }
  for (int r = 0; r < rows; ++r) {
    float v = 0.0f;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (tx[c].blk >= 0) v = fmaf(tx[c].wgt, var_lds[r * nbx + tx[c].blk], v);
    a.var[size_t(y0 + r) * a.w + x] = v;
  }
}

// unit-sum normalisation of the whitened PSF (sum in double, raster order)
__global__ void k_psf_unit_sum(int ps, float* psf) {
  __shared__ float inv;
  __shared__ int pos;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < ps * ps; ++i) s += double(psf[i]);
    pos = s > 0.0;
    inv = static_cast<float>(1.0 / s);
  }
  __syncthreads();
  if (!pos) return;
  for (int i = threadIdx.x; i < ps * ps; i += blockDim.x) psf[i] *= inv;
}

// ------------------------------------------------------------------ whitened PSF, direct
// The whitened PSF of whiten_score_psf is the circular convolution, on the
// n x n block grid, of the zero-phase PSF with phi = ifft(Phi) / n^2, cropped
// to the PSF's own ps x ps footprint; only phi(d) for |dx|, |dy| < ps enters.
// Those offsets come from direct DFT sums in double instead of three
// 2-D transforms of the whole block:
//   R[i][kx]   = sum_j K[s_i][s_j] w^(kx g_j)            (kernel rows)
//   Khat       = sum_i R[i][kx] w^(ky g_i),  Phi = float(cnum / sqrt(vs + vr |Khat|^2))
//   C/S[ky][d] = sum_kx c_kx Phi[ky][kx] cos/sin(2 pi kx d / n), c = 1 at kx = 0, n/2, else 2
//   phi(dy,dx) = sum_ky cos(2 pi ky dy / n) C - sin(2 pi ky dy / n) S, over n^2
//   out(p)     = sum_q psf(q) phi(p - q)
// with w = exp(-2 pi i / n) and (g, s) the grid cells of the kernel's
// zero-phase embedding and the source taps they hold.  Phi is real and even
// (|Khat(-k)| = |Khat(k)| for a real kernel), which the half-spectrum weights
// c use.
constexpr int kWhitenDirectMax = 48;  // PSF sides whitened by the direct sums (LDS)

struct WhitenArgs {
  int n, nh, ks, nl, ps;
  const float* kern;        // ks x ks (cell 0)
  const int* lg;            // nl grid cells ...
  const int* ls;            // ... and their source taps, one axis
  const double* tc;         // cos(2 pi m / n), m < n
  const double* ts;         // sin(2 pi m / n)
  const double* vs;
  const double* vr;
  const double* sumk2;
  double2* rows;            // nl x nh
  double* phi_hat;          // n x nh
  double* cs;               // n x ps x 2 (C, S)
  double* phi;              // (2 ps - 1)^2
  const float* psf;
  float* out;
};

__global__ __launch_bounds__(256) void k_wh_rows(WhitenArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.nl * a.nh) return;
  const int i = e / a.nh, kx = e - i * a.nh, m = a.n - 1;
  const float* kr = a.kern + size_t(a.ls[i]) * a.ks;
  double re = 0.0, im = 0.0;
  for (int j = 0; j < a.nl; ++j) {
    const int t = (kx * a.lg[j]) & m;
    const double v = static_cast<double>(kr[a.ls[j]]);
    re = fma(v, a.tc[t], re);
    im = fma(-v, a.ts[t], im);
  }
  a.rows[e] = make_double2(re, im);
}

__global__ __launch_bounds__(256) void k_wh_filter(WhitenArgs a) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.n * a.nh) return;
  const int ky = e / a.nh, kx = e - ky * a.nh, m = a.n - 1;
  double re = 0.0, im = 0.0;
  for (int i = 0; i < a.nl; ++i) {
    const int t = (ky * a.lg[i]) & m;
    const double2 r = a.rows[size_t(i) * a.nh + kx];
    const double c = a.tc[t], sn = -a.ts[t];
    re = fma(r.x, c, fma(-r.y, sn, re));
    im = fma(r.x, sn, fma(r.y, c, im));
  }
  const double vs = *a.vs, vr = *a.vr;
  const double cnum = sqrt(fmax(vs + vr * *a.sumk2, 0.0));
  const double den = fma(vr, re * re + im * im, vs);
  a.phi_hat[e] = den > 0.0 ? static_cast<double>(static_cast<float>(cnum / sqrt(den))) : 0.0;
}

// one wave per (ky, d): lanes over kx
__global__ __launch_bounds__(256) void k_wh_cs(WhitenArgs a) {
  const int lane = threadIdx.x & 63;
  const int e = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (e >= a.n * a.ps) return;
  const int ky = e / a.ps, d = e - ky * a.ps, m = a.n - 1;
  const double* ph = a.phi_hat + size_t(ky) * a.nh;
  double c = 0.0, sn = 0.0;
  for (int kx = lane; kx < a.nh; kx += 64) {
    const double wgt = (kx == 0 || 2 * kx == a.n) ? ph[kx] : 2.0 * ph[kx];
    const int t = (kx * d) & m;
    c = fma(wgt, a.tc[t], c);
    sn = fma(wgt, a.ts[t], sn);
  }
  for (int off = 32; off > 0; off >>= 1) {
    c += __shfl_xor(c, off);
    sn += __shfl_xor(sn, off);
  }
  if (lane == 0) {
    a.cs[2 * size_t(e)] = c;
    a.cs[2 * size_t(e) + 1] = sn;
  }
}

// one wave per offset (dy, dx): lanes over ky
__global__ __launch_bounds__(256) void k_wh_phi(WhitenArgs a) {
  const int lane = threadIdx.x & 63;
  const int span = 2 * a.ps - 1;
  const int e = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (e >= span * span) return;
  const int dy = e / span - (a.ps - 1), dx = e % span - (a.ps - 1);
  const int ax = dx < 0 ? -dx : dx, m = a.n - 1;
  const double sg = dx < 0 ? -1.0 : 1.0;
  double v = 0.0;
  for (int ky = lane; ky < a.n; ky += 64) {
    const int t = (ky * dy) & m;
    const double* q = a.cs + 2 * (size_t(ky) * a.ps + ax);
    v = fma(a.tc[t], q[0], fma(-a.ts[t] * sg, q[1], v));
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (lane == 0) a.phi[e] = v / (static_cast<double>(a.n) * a.n);
}

// out(p) = sum_q psf(q) phi(p - q) (one wave per output, lanes over q, phi
// and the PSF staged in LDS), then the unit-sum normalisation of
// k_psf_unit_sum: the sum in double in raster order, each entry scaled by
// the float reciprocal.  One workgroup.
constexpr int kWhApplyThreads = 1024;
constexpr size_t wh_apply_lds(int ps) {
  return sizeof(double) * size_t(2 * ps - 1) * (2 * ps - 1) + 2 * sizeof(float) * size_t(ps) * ps;
}
__global__ __launch_bounds__(kWhApplyThreads) void k_wh_apply(WhitenArgs a) {
  extern __shared__ double wh_lds[];
  __shared__ float inv;
  __shared__ int pos;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ps = a.ps, span = 2 * ps - 1, np = ps * ps;
  double* phi = wh_lds;
  float* psf = reinterpret_cast<float*>(phi + span * span);
  float* vals = psf + np;
  for (int i = threadIdx.x; i < span * span; i += kWhApplyThreads) phi[i] = a.phi[i];
  for (int i = threadIdx.x; i < np; i += kWhApplyThreads) psf[i] = a.psf[i];
  __syncthreads();
  for (int e = wv; e < np; e += kWhApplyThreads / 64) {
    const int py = e / ps, px = e - py * ps;
    double v = 0.0;
    for (int q = lane; q < np; q += 64) {
      const int qy = q / ps, qx = q - qy * ps;
      v = fma(static_cast<double>(psf[q]), phi[(py - qy + ps - 1) * span + (px - qx + ps - 1)], v);
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) vals[e] = static_cast<float>(v);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sm = 0.0;
    for (int i = 0; i < np; ++i) sm += double(vals[i]);
    pos = sm > 0.0;
    inv = static_cast<float>(1.0 / sm);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < np; i += kWhApplyThreads) a.out[i] = pos ? vals[i] * inv : vals[i];
}

// ------------------------------------------------------------------ host helpers
std::vector<int> block_origins(int extent, int blk, int strd) {
  std::vector<int> o;
  if (extent <= blk) return {0};
  for (int p = 0; p + blk <= extent; p += strd) o.push_back(p);
  if (o.back() + blk < extent) o.push_back(extent - blk);
  return o;
}

std::vector<double> hann(int n) {
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) w[i] = 0.5 * (1.0 - std::cos(2.0 * M_PI * (i + 0.5) / n));
  return w;
}

std::vector<float2> twiddles(int n) {
  std::vector<float2> t(n);
  for (int k = 0; k < n; ++k) {
    const double ang = -2.0 * M_PI * k / n;
    t[k] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
  }
  return t;
}

int log2_exact(int n) {
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  return lg;
}

struct FftPlan {
  int n, lg, nh, cw;
  size_t lds;
};

FftPlan fft_plan(int n) {
  FftPlan p;
  p.n = n;
  p.lg = log2_exact(n);
  p.nh = n / 2 + 1;
  p.cw = std::min(kChunkComplex / n, p.nh);
  p.lds = sizeof(float2) * (size_t(n) + size_t(p.cw) * (n + 1) + size_t(kNoiseWaves) * 2 * n);
  return p;
}

constexpr size_t reg_lds_bytes(int n) {
  const int cw = n / 2 < 32 ? n / 2 : 32;
  return sizeof(float2) * (2 * size_t(n) + size_t(cw) * (n + 1) + size_t(kNoiseWaves) * 4 * n);
}

template <int N>
void launch_block_reg(int grid, const NoiseArgs& a, hipStream_t s) {
  constexpr size_t lds = reg_lds_bytes(N);
  set_max_lds(reinterpret_cast<const void*>(&k_block_reg<N>), int(lds));
  hipLaunchKernelGGL((k_block_reg<N>), dim3(grid), dim3(64 * kNoiseWaves), lds, s, a);
}

template <int MODE>
void launch_block_fft(const FftPlan& p, int grid, const NoiseArgs& a, hipStream_t s) {
  set_max_lds(reinterpret_cast<const void*>(&k_block_fft<MODE>), int(p.lds));
  hipLaunchKernelGGL((k_block_fft<MODE>), dim3(grid), dim3(64 * kNoiseWaves), p.lds, s, a);
}

// Local kernels and their power spectra at npts points (host xy pairs).
struct CellKernels {
  float* kern;     // npts x ks^2
  double* sumk2;   // npts
  double* power;   // npts x n x (n/2+1)
  int ks;
};

// With pwt (n = 256 only) the power goes there as float [cell][kx][ky] from
// the 256-point kernel; otherwise as double [cell][ky][kx] in ck.power.
CellKernels cell_kernels(const Tps& tps, const GhBasis& basis, const double* theta,
                         const std::vector<double>& pts, const FftPlan& p, const float2* d_tw,
                         Arena& ar, hipStream_t s, float* pwt = nullptr,
                         bool want_power = true) {
  const int npts = static_cast<int>(pts.size() / 2);
  const int nc = basis.nc, ks = basis.ks, ks2 = ks * ks;
  double* d_pts = ar.take<double>(pts.size());
  PinnedStage& stage = PinnedStage::local();
  stage.h2d(d_pts, pts.data(), pts.size(), s);
  double* d_a = ar.take<double>(size_t(nc + 1) * npts);
  eval_fields_points(tps, theta, nc + 1, d_pts, npts, d_a, ar, s, &stage);
  const std::vector<float> fp = basis.footprints_f32();
  float* d_phi = ar.take<float>(fp.size());
  stage.h2d(d_phi, fp.data(), fp.size(), s);
  CellKernels ck;
  ck.ks = ks;
  ck.kern = ar.take<float>(size_t(npts) * ks2);
  ck.sumk2 = ar.take<double>(npts);
  ck.power = nullptr;
  if (nc > kLocalMaxComp) invalid("decorrelate: at most 1024 basis components are supported");
  set_max_lds(reinterpret_cast<const void*>(&k_local_kernels), int(sizeof(float) * ks2));
  hipLaunchKernelGGL(k_local_kernels, dim3(npts), dim3(256), sizeof(float) * ks2, s, npts, nc, ks2,
                     static_cast<const double*>(d_a), static_cast<const float*>(d_phi), ck.kern,
                     ck.sumk2);
  if (!want_power) {
    DG_LAUNCH_CHECK();
    return ck;
  }
  NoiseArgs a{};
  a.n = p.n;
  a.lg = p.lg;
  a.cw = p.cw;
  a.tw = d_tw;
  a.kern = ck.kern;
  a.ks = ks;
  if (pwt) {
    if (p.n != 256) invalid("cell_kernels: float power only for 256-point blocks");
    a.pwt_out = pwt;
    set_max_lds(reinterpret_cast<const void*>(&k_block_r16<kModePower>), int(r16_lds_bytes()));
    hipLaunchKernelGGL(k_block_r16<kModePower>, dim3(npts), dim3(512), r16_lds_bytes(), s, a);
  } else {
    ck.power = ar.take<double>(size_t(npts) * p.n * p.nh);
    a.scratch = ar.take<float2>(size_t(npts) * p.n * p.nh);
    a.power_out = ck.power;
    launch_block_fft<kModePower>(p, npts, a, s);
  }
  DG_LAUNCH_CHECK();
  return ck;
}

size_t cell_kernels_bytes(const Tps& tps, const GhBasis& basis, int npts, const FftPlan& p) {
  const int nc = basis.nc, ks2 = basis.ks * basis.ks;
  return Arena::round(sizeof(double) * 2 * npts) + Arena::round(sizeof(double) * (nc + 1) * npts) +
         eval_fields_points_bytes(tps, nc + 1) + Arena::round(sizeof(float) * nc * ks2) +
         Arena::round(sizeof(float) * size_t(npts) * ks2) + Arena::round(sizeof(double) * npts) +
         Arena::round(sizeof(double) * size_t(npts) * p.n * p.nh) +
         Arena::round(sizeof(float2) * size_t(npts) * p.n * p.nh);
}

void check_block(int block) {
  if (block < 8 || (block & (block - 1)) != 0)
    runtime("decorrelate: block must be a power of two >= 8");
  if (block > kMaxBlock) invalid("decorrelate: block above 512 is not supported");
}

// out[y][x] = raw[src] / n^2 with src = (y, x), or the fft-shifted
// ((y + n/2) % n, (x + n/2) % n) (reference noise.cpp:119-161)
__global__ void k_block_scale(int n, int shift, const float* __restrict__ raw,
                              float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= n) return;
  const float inv_n2 = 1.0f / static_cast<float>(n * n);
  const int sy = shift ? (y + n / 2) % n : y, sx = shift ? (x + n / 2) % n : x;
  out[size_t(y) * n + x] = raw[size_t(sy) * n + sx] * inv_n2;
}

__global__ void k_unit_impulse(int n, float* __restrict__ img) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * n) img[i] = i == 0 ? 1.0f : 0.0f;
}

}  // namespace

void run_decorrelate(int w, int h, const float* diff, const double* knots, int n_knots,
                     const double* theta, int n_theta, double beta, int n_max, const float* var_s,
                     const float* var_r, int block, int radius, int kernel_cell_blocks,
                     float* out, float* out_var, hipStream_t s) {
  check_block(block);
  if (w <= 0 || h <= 0) invalid("decorrelate: empty frame");
  const Tps tps(knots, n_knots);
  const GhBasis basis(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
  if (n_theta != (basis.nc + 1) * tps.k) invalid("decorrelate: theta size does not match k*(nc+1)");
  const FftPlan p = fft_plan(block);
  const int stride = block / 2;
  const std::vector<int> oxs = block_origins(w, block, stride), oys = block_origins(h, block, stride);
  const int nox = static_cast<int>(oxs.size()), noy = static_cast<int>(oys.size());
  int kgroup = kernel_cell_blocks;
  if (kgroup <= 0) {
    const double spacing = tps.min_spacing();
    kgroup = 1;
    if (spacing > 0.0 && stride > 0)
      kgroup = std::max(1, static_cast<int>(std::lround(spacing / 4.0 / stride)));
  }
  const int ncellx = (nox + kgroup - 1) / kgroup, ncelly = (noy + kgroup - 1) / kgroup;
  const int ncell = ncellx * ncelly;
  std::vector<double> cpts(2 * size_t(ncell));
  for (int ci = 0; ci < ncell; ++ci) {
    const int cix = ci % ncellx, ciy = ci / ncellx;
    const int repx = std::min(cix * kgroup + kgroup / 2, nox - 1);
    const int repy = std::min(ciy * kgroup + kgroup / 2, noy - 1);
    cpts[2 * ci] = std::clamp(oxs[repx] + block / 2.0, 0.0, w - 1.0);
    cpts[2 * ci + 1] = std::clamp(oys[repy] + block / 2.0, 0.0, h - 1.0);
  }
  // Colours (ix % 3, iy % 3), as the reference orders its overlap-add
  // (noise.cpp:346-355): blocks of one colour are pairwise disjoint (origins
  // at least three strides apart, a flush last block included), so each
  // colour is one launch, and the launches run in the reference's colour
  // order (ix % 3) * 3 + (iy % 3).  A pixel's first covering colour stores,
  // the later ones accumulate.
  const int nblk = nox * noy;
  std::vector<BlockJob> jobs;
  std::vector<int> class_start, class_count, class_x, class_y;
  jobs.reserve(nblk);
  for (int cx = 0; cx < 3; ++cx)
    for (int cy = 0; cy < 3; ++cy) {
      const int start = static_cast<int>(jobs.size());
      for (int iy = cy; iy < noy; iy += 3)
        for (int ix = cx; ix < nox; ix += 3)
          jobs.push_back({oxs[ix], oys[iy], (iy / kgroup) * ncellx + (ix / kgroup),
                          iy * nox + ix, 0});
      if (static_cast<int>(jobs.size()) == start) continue;
      class_start.push_back(start);
      class_count.push_back(static_cast<int>(jobs.size()) - start);
      class_x.push_back(cx);
      class_y.push_back(cy);
    }
  std::vector<int> bxy(2 * size_t(nblk));
  for (int iy = 0; iy < noy; ++iy)
    for (int ix = 0; ix < nox; ++ix) {
      bxy[2 * (iy * nox + ix)] = oxs[ix];
      bxy[2 * (iy * nox + ix) + 1] = oys[iy];
    }
  // window sums (separable) and the covering blocks of every row and column
  const std::vector<double> hw = hann(block);
  std::vector<double> wxt(w, 0.0), wyt(h, 0.0);
  std::vector<int> xsel(3 * size_t(w), -1), ysel(3 * size_t(h), -1);
  std::vector<unsigned char> fx(w, 255), fy(h, 255);
  for (int ix = 0; ix < nox; ++ix)
    for (int j = 0; j < block; ++j) {
      const int x = oxs[ix] + j;
      if (x < 0 || x >= w) continue;
      wxt[x] += hw[j];
      xsel[3 * x + ix % 3] = ix;
      fx[x] = static_cast<unsigned char>(std::min(int(fx[x]), ix % 3));
    }
  for (int iy = 0; iy < noy; ++iy)
    for (int j = 0; j < block; ++j) {
      const int y = oys[iy] + j;
      if (y < 0 || y >= h) continue;
      wyt[y] += hw[j];
      ysel[3 * y + iy % 3] = iy;
      fy[y] = static_cast<unsigned char>(std::min(int(fy[y]), iy % 3));
    }
  const std::vector<float2> tw = twiddles(block);
  const bool r16 = block == 256;
  const bool reg = block == 64 || block == 128;
  const int nh = block / 2 + 1;

  // small uploads from pinned staging (the callers drain the stream before
  // returning, so the next call on this thread may reuse it)
  PinnedStage& stage = PinnedStage::local();
  stage.reset();
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(cell_kernels_bytes(tps, basis, ncell, p) + Arena::round(sizeof(float2) * block) +
             Arena::round(sizeof(double) * block) + Arena::round(sizeof(BlockJob) * nblk) +
             Arena::round(sizeof(int) * 3 * nblk) + 3 * Arena::round(sizeof(double) * nblk) +
             (r16 ? Arena::round(sizeof(float) * size_t(ncell) * nh * block) : 0) +
             ((reg || r16) ? 0 : Arena::round(sizeof(float2) * size_t(nblk) * p.n * p.nh)) +
             Arena::round(sizeof(double) * (w + h)) + Arena::round(w + h) +
             2 * Arena::round(sizeof(VarTab) * 3 * (w + h)) + Arena::round(sizeof(int) * ((w + 255) / 256)) +
             (1 << 16));
  float2* d_tw = ar.take<float2>(block);
  stage.h2d(d_tw, tw.data(), tw.size(), s);
  double* d_hw = ar.take<double>(block);
  stage.h2d(d_hw, hw.data(), hw.size(), s);
  ProfScope prof_all("decorrelate", s);
  float* d_pwt = r16 ? ar.take<float>(size_t(ncell) * nh * block) : nullptr;
  CellKernels ck = [&] {
    ProfScope pc("decorrelate_cells", s);
    return cell_kernels(tps, basis, theta, cpts, p, d_tw, ar, s, d_pwt);
  }();
  BlockJob* d_jobs = ar.take<BlockJob>(nblk);
  stage.h2d(d_jobs, jobs.data(), jobs.size(), s);
  int* d_bxy = ar.take<int>(2 * nblk);
  stage.h2d(d_bxy, bxy.data(), bxy.size(), s);
  double* d_vs = ar.take<double>(nblk);
  double* d_vr = ar.take<double>(nblk);
  double* d_num = ar.take<double>(nblk);
  double* d_wxt = ar.take<double>(w);
  double* d_wyt = ar.take<double>(h);
  stage.h2d(d_wxt, wxt.data(), w, s);
  stage.h2d(d_wyt, wyt.data(), h, s);
  unsigned char* d_fx = ar.take<unsigned char>(w + h);
  unsigned char* d_fy = d_fx + w;
  stage.h2d(d_fx, fx.data(), w, s);
  stage.h2d(d_fy, fy.data(), h, s);
  {
    ProfScope pn("decorrelate_noise", s);
    hipLaunchKernelGGL(k_block_noise, dim3(nblk, 2), dim3(kNoiseThreads), 0, s, block, w, h,
                       static_cast<const int*>(d_bxy), var_s, var_r, d_vs, d_vr);
  }
  float2* scr = (reg || r16) ? nullptr : ar.take<float2>(size_t(nblk) * p.n * p.nh);
  NoiseArgs a{};
  a.n = p.n;
  a.lg = p.lg;
  a.cw = p.cw;
  a.w = w;
  a.h = h;
  a.tw = d_tw;
  a.hwin = d_hw;
  a.scratch = scr;
  a.diff = diff;
  a.vs = d_vs;
  a.vr = d_vr;
  a.power = ck.power;
  a.sumk2 = ck.sumk2;
  a.pwt = d_pwt;
  a.out = out;
  a.inv_n2 = 1.0 / (double(block) * block);
  a.fx = d_fx;
  a.fy = d_fy;
  a.nox = nox;
  {
    ProfScope pb("decorrelate_blocks", s);
    if (r16)
      set_max_lds(reinterpret_cast<const void*>(&k_block_r16<kModeDecorrelate>), int(r16_lds_bytes()));
    for (size_t c = 0; c < class_start.size(); ++c) {
      a.jobs = d_jobs + class_start[c];
      a.xcls = class_x[c];
      a.ycls = class_y[c];
      const int grid = class_count[c];
      if (r16) {
        hipLaunchKernelGGL(k_block_r16<kModeDecorrelate>, dim3(grid), dim3(512), r16_lds_bytes(), s, a);
      } else {
        switch (reg ? block : 0) {
          case 64: launch_block_reg<64>(grid, a, s); break;
          case 128: launch_block_reg<128>(grid, a, s); break;
          default: launch_block_fft<kModeDecorrelate>(p, grid, a, s);
        }
      }
      DG_LAUNCH_CHECK();
    }
  }
  {
    ProfScope pz("decorrelate_normalise", s);
    const int nvec = (reinterpret_cast<uintptr_t>(out) % 16 == 0 && w % 4 == 0) ? 1 : 0;
    hipLaunchKernelGGL(k_noise_normalise, dim3((w + 1023) / 1024, (h + kNormRows - 1) / kNormRows),
                       dim3(256), 0, s, w, h, static_cast<const double*>(d_wxt),
                       static_cast<const double*>(d_wyt), out, nvec);
  }
  DG_LAUNCH_CHECK();
  std::vector<int> bcell(nblk);
  for (int iy = 0; iy < noy; ++iy)
    for (int ix = 0; ix < nox; ++ix) bcell[iy * nox + ix] = (iy / kgroup) * ncellx + (ix / kgroup);
  int* d_bcell = ar.take<int>(nblk);
  stage.h2d(d_bcell, bcell.data(), nblk, s);
  hipLaunchKernelGGL(k_block_num, dim3((nblk + 255) / 256), dim3(256), 0, s, nblk,
                     static_cast<const int*>(d_bcell), static_cast<const double*>(d_vs),
                     static_cast<const double*>(d_vr), static_cast<const double*>(ck.sumk2), d_num);
  // covering blocks of every column and row with weights over the window sum
  std::vector<VarTab> xtab(3 * size_t(w), VarTab{0.0f, -1}), ytab(3 * size_t(h), VarTab{0.0f, -1});
  for (int x = 0; x < w; ++x)
    for (int c = 0; c < 3; ++c) {
      const int ix = xsel[3 * x + c];
      if (ix >= 0 && wxt[x] > 0.0) xtab[3 * x + c] = VarTab{float(hw[x - oxs[ix]] / wxt[x]), ix};
    }
  for (int y = 0; y < h; ++y)
    for (int c = 0; c < 3; ++c) {
      const int iy = ysel[3 * y + c];
      if (iy >= 0 && wyt[y] > 0.0) ytab[3 * y + c] = VarTab{float(hw[y - oys[iy]] / wyt[y]), iy};
    }
  // per 256-column tile: its first and last covering block columns
  const int ntx = (w + 255) / 256;
  std::vector<int> xlo(ntx);
  int nbx = 1;
  for (int tx = 0; tx < ntx; ++tx) {
    int lo = nox, hi = -1;
    for (int x = tx * 256; x < std::min(w, tx * 256 + 256); ++x)
      for (int c = 0; c < 3; ++c) {
        const int ix = xsel[3 * x + c];
        if (ix >= 0) {
          lo = std::min(lo, ix);
          hi = std::max(hi, ix);
        }
      }
    xlo[tx] = lo <= hi ? lo : 0;
    nbx = std::max(nbx, hi - lo + 1);
  }
  VarTab* d_xtab = ar.take<VarTab>(xtab.size());
  VarTab* d_ytab = ar.take<VarTab>(ytab.size());
  int* d_xlo = ar.take<int>(ntx);
  stage.h2d(d_xtab, xtab.data(), xtab.size(), s);
  stage.h2d(d_ytab, ytab.data(), ytab.size(), s);
  stage.h2d(d_xlo, xlo.data(), xlo.size(), s);
  VarArgs na{};
  na.w = w;
  na.h = h;
  na.nox = nox;
  na.xt = d_xtab;
  na.yt = d_ytab;
  na.xlo = d_xlo;
  na.num = d_num;
  na.var = out_var;
  {
    ProfScope pz("decorrelate_variance", s);
    const size_t lds = sizeof(float) * kVarRows * size_t(nbx);
    set_max_lds(reinterpret_cast<const void*>(&k_noise_variance), int(lds));
    hipLaunchKernelGGL(k_noise_variance, dim3(ntx, (h + kVarRows - 1) / kVarRows), dim3(256), lds, s,
                       na, nbx);
  }
  DG_LAUNCH_CHECK();
}

void run_whiten_score_psf(const float* psf, int ps, const double* knots, int n_knots,
                          const double* theta, int n_theta, double beta, int n_max,
                          const float* var_s, const float* var_r, int w, int h, int block,
                          int radius, float* out, hipStream_t s) {
  if (block < ps) runtime("whiten_score_psf: block must be >= psf side");
  check_block(block);
  if (w <= 0 || h <= 0) invalid("whiten_score_psf: empty frame");
  const Tps tps(knots, n_knots);
  const GhBasis basis(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
  if (n_theta != (basis.nc + 1) * tps.k)
    invalid("whiten_score_psf: theta size does not match k*(nc+1)");
  const FftPlan p = fft_plan(block);
  const std::vector<float2> tw = twiddles(block);
  const bool direct = ps <= kWhitenDirectMax;
  const int n = p.n, nh = p.nh, ks = basis.ks, span = 2 * ps - 1;
  // small uploads from pinned staging (the callers drain the stream)
  PinnedStage& stage = PinnedStage::local();
  stage.reset();
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(cell_kernels_bytes(tps, basis, 1, p) + Arena::round(sizeof(float2) * block) +
             Arena::round(sizeof(int) * 2) + 2 * Arena::round(sizeof(double)) +
             Arena::round(sizeof(float2) * p.n * p.nh) + (1 << 16) +
             (direct ? 2 * Arena::round(sizeof(int) * n) + 2 * Arena::round(sizeof(double) * n) +
                           Arena::round(sizeof(double2) * size_t(n) * nh) +
                           Arena::round(sizeof(double) * size_t(n) * nh) +
                           Arena::round(sizeof(double) * 2 * size_t(n) * ps) +
                           Arena::round(sizeof(double) * size_t(span) * span)
                     : 0));
  float2* d_tw = ar.take<float2>(block);
  stage.h2d(d_tw, tw.data(), tw.size(), s);
  ProfScope prof("whiten_psf", s);
  const std::vector<double> centre{0.5 * (w - 1), 0.5 * (h - 1)};
  CellKernels ck = cell_kernels(tps, basis, theta, centre, p, d_tw, ar, s, nullptr, !direct);
  const int bxy[2] = {std::max(0, w / 2 - block / 2), std::max(0, h / 2 - block / 2)};
  int* d_bxy = ar.take<int>(2);
  stage.h2d(d_bxy, bxy, 2, s);
  double* d_vs = ar.take<double>(1);
  double* d_vr = ar.take<double>(1);
  hipLaunchKernelGGL(k_block_noise, dim3(1, 2), dim3(256), 0, s, block, w, h,
                     static_cast<const int*>(d_bxy), var_s, var_r, d_vs, d_vr);
  if (direct) {
    // the kernel's zero-phase embedding along one axis: grid cell g holds
    // source tap (g + r) mod n, or the last of several when ks > n
    std::vector<int> lg, ls;
    for (int g = 0; g < n; ++g) {
      const int base = (g + ks / 2) % n;
      if (base >= ks) continue;
      lg.push_back(g);
      ls.push_back(base + n * ((ks - 1 - base) / n));
    }
    std::vector<double> tc(n), ts(n);
    for (int m = 0; m < n; ++m) {
      tc[m] = std::cos(2.0 * M_PI * m / n);
      ts[m] = std::sin(2.0 * M_PI * m / n);
    }
    WhitenArgs wa{};
    wa.n = n;
    wa.nh = nh;
    wa.ks = ks;
    wa.nl = static_cast<int>(lg.size());
    wa.ps = ps;
    wa.kern = ck.kern;
    int* d_lg = ar.take<int>(n);
    int* d_ls = ar.take<int>(n);
    double* d_tc = ar.take<double>(n);
    double* d_ts = ar.take<double>(n);
    stage.h2d(d_lg, lg.data(), lg.size(), s);
    stage.h2d(d_ls, ls.data(), ls.size(), s);
    stage.h2d(d_tc, tc.data(), tc.size(), s);
    stage.h2d(d_ts, ts.data(), ts.size(), s);
    wa.lg = d_lg;
    wa.ls = d_ls;
    wa.tc = d_tc;
    wa.ts = d_ts;
    wa.vs = d_vs;
    wa.vr = d_vr;
    wa.sumk2 = ck.sumk2;
    wa.rows = ar.take<double2>(size_t(n) * nh);
    wa.phi_hat = ar.take<double>(size_t(n) * nh);
    wa.cs = ar.take<double>(2 * size_t(n) * ps);
    wa.phi = ar.take<double>(size_t(span) * span);
    wa.psf = psf;
    wa.out = out;
    const auto blocks = [](size_t e) { return dim3(static_cast<unsigned>((e + 255) / 256)); };
    hipLaunchKernelGGL(k_wh_rows, blocks(size_t(wa.nl) * nh), dim3(256), 0, s, wa);
    hipLaunchKernelGGL(k_wh_filter, blocks(size_t(n) * nh), dim3(256), 0, s, wa);
    hipLaunchKernelGGL(k_wh_cs, blocks(64 * size_t(n) * ps), dim3(256), 0, s, wa);
    hipLaunchKernelGGL(k_wh_phi, blocks(64 * size_t(span) * span), dim3(256), 0, s, wa);
    set_max_lds(reinterpret_cast<const void*>(&k_wh_apply), int(wh_apply_lds(ps)));
    hipLaunchKernelGGL(k_wh_apply, dim3(1), dim3(kWhApplyThreads), wh_apply_lds(ps), s, wa);
    DG_LAUNCH_CHECK();
    return;
  }
  NoiseArgs a{};
  a.n = p.n;
  a.lg = p.lg;
  a.cw = p.cw;
  a.tw = d_tw;
  a.scratch = ar.take<float2>(size_t(p.n) * p.nh);
  a.vs = d_vs;
  a.vr = d_vr;
  a.power = ck.power;
  a.sumk2 = ck.sumk2;
  a.psf = psf;
  a.ps = ps;
  a.psf_out = out;
  launch_block_fft<kModeWhiten>(p, 1, a, s);
  hipLaunchKernelGGL(k_psf_unit_sum, dim3(1), dim3(256), 0, s, ps, out);
  DG_LAUNCH_CHECK();
}

// One n x n block whitened with a constant kernel and noise levels: the
// filter Phi from the kernel's power (noise.cpp:65-117), then
// r2c * Phi -> c2r / n^2 (apply_filter_fft, :119-137).  With img == nullptr
// the block is a unit impulse at the origin, whose spectrum is exactly one, so
// the result is ifft(Phi) / n^2, fft-shifted when `shift` (the QA kernel image
// of decorrelation_kernel_image, :139-161).  Host inputs, host output.
void run_decorrelate_block(int n, const float* img, const float* kernel, int ks, double vs,
                           double vr, bool shift, float* out) {
  if (n < 8 || (n & (n - 1)) != 0 || n > kMaxBlock)
    invalid("decorrelate_block: n must be a power of two in [8, 512] on the GPU");
  if (ks < 1) invalid("decorrelate_block: kernel must be non-empty");
  double sumk2 = 0.0;
  for (int i = 0; i < ks * ks; ++i) sumk2 = std::fma(double(kernel[i]), double(kernel[i]), sumk2);
  const FftPlan p = fft_plan(n);
  const std::vector<float2> tw = twiddles(n);
  const size_t nn = size_t(n) * n, spec = size_t(n) * p.nh;
  OwnedStream st;
  hipStream_t s = st.get();
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(Arena::round(sizeof(float2) * n) + Arena::round(sizeof(float) * ks * ks) +
             3 * Arena::round(sizeof(double)) + Arena::round(sizeof(double) * spec) +
             Arena::round(sizeof(float2) * spec) + 3 * Arena::round(sizeof(float) * nn) +
             Arena::round(sizeof(BlockJob)) + (1 << 12));
  float2* d_tw = ar.take<float2>(n);
  float* d_k = ar.take<float>(size_t(ks) * ks);
  double* d_sk = ar.take<double>(1);
  double* d_vs = ar.take<double>(1);
  double* d_vr = ar.take<double>(1);
  double* d_pow = ar.take<double>(spec);
  float2* scr = ar.take<float2>(spec);
  float* d_img = ar.take<float>(nn);
  float* d_raw = ar.take<float>(nn);
  float* d_out = ar.take<float>(nn);
  BlockJob* d_job = ar.take<BlockJob>(1);
  const BlockJob job{0, 0, 0, 0, 0};
  h2d(d_tw, tw.data(), tw.size(), s);
  h2d(d_k, kernel, size_t(ks) * ks, s);
  h2d(d_sk, &sumk2, 1, s);
  h2d(d_vs, &vs, 1, s);
  h2d(d_vr, &vr, 1, s);
  h2d(d_job, &job, 1, s);
  if (img) {
    h2d(d_img, img, nn, s);
  } else {
    hipLaunchKernelGGL(k_unit_impulse, dim3(unsigned((nn + 255) / 256)), dim3(256), 0, s, n,
                       d_img);
  }
  NoiseArgs a{};
  a.n = p.n;
  a.lg = p.lg;
  a.cw = p.cw;
  a.w = n;
  a.h = n;
  a.tw = d_tw;
  a.scratch = scr;
  a.kern = d_k;
  a.ks = ks;
  a.power_out = d_pow;
  launch_block_fft<kModePower>(p, 1, a, s);
  a.diff = d_img;
  a.planes = d_raw;
  a.plane_stride = nn;
  a.jobs = d_job;
  a.vs = d_vs;
  a.vr = d_vr;
  a.power = d_pow;
  a.sumk2 = d_sk;
  launch_block_fft<kModeDecorrelate>(p, 1, a, s);
  hipLaunchKernelGGL(k_block_scale, dim3((n + 255) / 256, n), dim3(256), 0, s, n, shift ? 1 : 0,
                     static_cast<const float*>(d_raw), d_out);
  DG_LAUNCH_CHECK();
  d2h(out, d_out, nn, s);
  DG_HIP(hipStreamSynchronize(s));
}

}  // namespace dg
