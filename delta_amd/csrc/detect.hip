// Detection stage on the GPU: the match-filtered score image and the
// 8-connected labelling of candidate pixels.
//
//   k_matched_filter   src/noise.cpp:535-632.  Per output pixel the taps are
//                      applied ly-outer, lx-inner with fused multiply-adds and
//                      zero taps skipped, then score = acc / sqrt(v * sum psf^2)
//                      with IEEE division and square root: the reference's
//                      operation order, so the score is bit-identical.
//   labelling          src/catalog.cpp:61-261.  seed_mask, then union-find on
//                      the label image itself (1-based parent index, 0 =
//                      background): every link goes from the larger index to
//                      the smaller (atomicMin), so each component's root is its
//                      raster-first pixel.  Unions run first inside 32x32 tiles
//                      in LDS, then across tile edges on the global image.  The
//                      final label of a component is 1 + the number of roots
//                      before its root in raster order (an exclusive scan of
//                      root flags), which is exactly the reference's numbering.
#include <mutex>
#include <string>

#include "dg_kernels.hpp"

namespace dg {
namespace {

// ------------------------------------------------------------------ matched filter
constexpr int kMfW = 128, kMfH = 16;  // output tile; each thread owns 8 columns of one row
constexpr int kMfOut = 8;
constexpr int kMfMaxPs = 63;
// The window is stored skewed: column x of a row at x + x / 8, and rows
// kMfLd(ps) apart with kMfLd = 16 mod 64, so the 16 threads of a row, which
// start 8 columns apart, and the 4 rows of a wave fall on distinct banks.
__host__ __device__ constexpr int mf_pos(int x) { return x + x / 8; }
__host__ __device__ constexpr int mf_ld(int ps) {
  return ((mf_pos(kMfW + 2 * (ps / 2) - 1) + 1 + 47) / 64) * 64 + 16;
}

// PS > 0: the PSF side known at compile time (window row in registers, taps
// as uniform scalar loads); PS == 0: any side up to kMfMaxPs.
template <int PS>
__global__ __launch_bounds__(256) void k_matched_filter(int w, int h, const float* __restrict__ img,
                                                        const float* __restrict__ psf, int ps_rt,
                                                        const float* __restrict__ var,
                                                        float* __restrict__ out) {
  extern __shared__ float mf_lds[];
  const int ps = PS > 0 ? PS : ps_rt;
  const int r = ps / 2;
  const int ww = kMfW + 2 * r, wh = kMfH + 2 * r;
  const int ld = mf_ld(ps);
  float* win = mf_lds;   // wh rows of ld, skewed, zero padded at the frame edge
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * kMfW, y0 = blockIdx.y * kMfH;
  // sum psf^2 (one chain, the reference's order) first: its latency overlaps
  // the window loads
  float sumpsf2 = 0.0f;
  for (int i = 0; i < ps * ps; ++i) sumpsf2 = fmaf(psf[i], psf[i], sumpsf2);
  // 8 loads per thread in flight before any is stored
  for (int base = 0; base < wh * ww; base += 256 * 8) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = base + tid + 256 * j;
      const int iy = i / ww, ix = i - iy * ww;
      const int gy = y0 - r + iy, gx = x0 - r + ix;
      v[j] = (i >= wh * ww || gy < 0 || gy >= h || gx < 0 || gx >= w) ? 0.0f
                                                                   : img[size_t(gy) * w + gx];
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int i = base + tid + 256 * j;
      if (i < wh * ww) {
        const int iy = i / ww, ix = i - iy * ww;
        win[iy * ld + mf_pos(ix)] = v[j];
      }
    }
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  float acc[kMfOut];
#pragma unroll
  for (int j = 0; j < kMfOut; ++j) acc[j] = 0.0f;
  for (int ly = 0; ly < ps; ++ly) {
    // column kMfOut tx + j of the row sits at mf_pos(kMfOut tx) + mf_pos(j)
    const float* row = win + (ty + ly) * ld + mf_pos(kMfOut * tx);
    if constexpr (PS > 0) {
      float v[PS + kMfOut - 1];
#pragma unroll
      for (int j = 0; j < PS + kMfOut - 1; ++j) v[j] = row[mf_pos(j)];
#pragma unroll
      for (int lx = 0; lx < PS; ++lx) {
        const float p = psf[ly * PS + lx];
        if (p == 0.0f) continue;
#pragma unroll
        for (int j = 0; j < kMfOut; ++j) acc[j] = fmaf(p, v[lx + j], acc[j]);
      }
    } else {
      for (int lx = 0; lx < ps; ++lx) {
        const float p = psf[ly * ps + lx];
        if (p == 0.0f) continue;
#pragma unroll
        for (int j = 0; j < kMfOut; ++j) acc[j] = fmaf(p, row[mf_pos(lx + j)], acc[j]);
      }
    }
  }
  const int y = y0 + ty;
  if (y >= h) return;
#pragma unroll
  for (int j = 0; j < kMfOut; ++j) {
    const int x = x0 + kMfOut * tx + j;
    if (x >= w) continue;
    const float vv = var[size_t(y) * w + x];
    out[size_t(y) * w + x] = vv > 0.0f ? acc[j] / sqrtf(vv * sumpsf2) : 0.0f;
  }
}

template <int PS>
void launch_mf(int w, int h, const float* img, const float* psf, int ps, const float* var,
               float* out, hipStream_t s) {
  const int r = ps / 2;
  const size_t lds = sizeof(float) * size_t(kMfH + 2 * r) * mf_ld(ps);
  set_max_lds(reinterpret_cast<const void*>(&k_matched_filter<PS>), int(lds));
  hipLaunchKernelGGL((k_matched_filter<PS>), dim3((w + kMfW - 1) / kMfW, (h + kMfH - 1) / kMfH),
                     dim3(256), lds, s, w, h, img, psf, ps, var, out);
}

// ------------------------------------------------------------------ labelling
__device__ __forceinline__ int ld_parent(const int* p, int i) {
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root (0-based) of pixel x; parents are 1-based, a root points at itself
__device__ __forceinline__ int uf_root(const int* p, int x) {
  int q = ld_parent(p, x) - 1;
  while (q != x) {
    x = q;
    q = ld_parent(p, x) - 1;
  }
  return x;
}

// Link the sets of a and b, always hanging the larger root under the smaller.
// Parents only decrease, so a stale read still names an ancestor; the loop
// ends when atomicMin finds the expected root unchanged.
__device__ void uf_unite(int* p, int a, int b) {
  for (;;) {
    a = uf_root(p, a);
    b = uf_root(p, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(p + b, a + 1);
    if (old == b + 1) return;
    b = old - 1;
  }
}

// Tile-local pass: seeds of a 32x32 tile and their 8-connected union-find in
// LDS (same min-index linking), then every seed pixel gets the global index of
// its tile-local root (its tile component's raster-first pixel) + 1, or 0.
constexpr int kCclT = 32;

__device__ __forceinline__ int lds_root(const int* lab, int x) {
  int q = __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (q != x) {
    x = q;
    q = __hip_atomic_load(lab + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  return x;
}

__device__ void lds_unite(int* lab, int a, int b) {
  for (;;) {
    a = lds_root(lab, a);
    b = lds_root(lab, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicMin(lab + b, a);
    if (old == b) return;
    b = old;
  }
}

// Thread t owns the four consecutive pixels 4t .. 4t+3 of the tile (row t / 8),
// loaded and stored as one 16-byte vector where the frame's rows allow it.
__global__ __launch_bounds__(256) void k_ccl_local(int w, int h, const float* __restrict__ score,
                                                   const uint8_t* __restrict__ mask, float thr,
                                                   int polarity, int* __restrict__ p,
                                                   uint8_t* __restrict__ occ) {
  __shared__ int lab[kCclT * kCclT];
  const int x0 = blockIdx.x * kCclT, y0 = blockIdx.y * kCclT;
  const int l0 = 4 * threadIdx.x, x = x0 + (l0 & (kCclT - 1)), y = y0 + l0 / kCclT;
  const bool vec = (w & 3) == 0 && x + 3 < w && y < h;
  const size_t i0 = size_t(y) * w + x;
  float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  bool in[4] = {false, false, false, false};
  if (vec) {
    const float4 q = *reinterpret_cast<const float4*>(score + i0);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
    in[0] = in[1] = in[2] = in[3] = true;
  } else if (y < h) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (x + c < w) {
        v[c] = score[i0 + c];
        in[c] = true;
      }
  }
  int any = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    bool pass = in[c] && isfinite(v[c]) && (polarity > 0 ? v[c] >= thr : v[c] <= -thr);
    if (pass && mask) pass = mask[i0 + c] == 0;
    lab[l0 + c] = pass ? l0 + c : -1;
    any |= pass ? 1 : 0;
  }
  int out[4] = {0, 0, 0, 0};
  // the tile's seed occupancy: the later passes skip tiles without seeds
  const bool occupied = __syncthreads_or(any) != 0;
  if (threadIdx.x == 0) occ[blockIdx.y * gridDim.x + blockIdx.x] = occupied ? 1 : 0;
  if (occupied) {
    for (int l = threadIdx.x; l < kCclT * kCclT; l += 256) {
      if (lab[l] < 0) continue;
      const int lx = l & (kCclT - 1), ly = l / kCclT;
      if (lx > 0 && lab[l - 1] >= 0) lds_unite(lab, l, l - 1);
      if (ly > 0) {
        if (lx > 0 && lab[l - kCclT - 1] >= 0) lds_unite(lab, l, l - kCclT - 1);
        if (lab[l - kCclT] >= 0) lds_unite(lab, l, l - kCclT);
        if (lx + 1 < kCclT && lab[l - kCclT + 1] >= 0) lds_unite(lab, l, l - kCclT + 1);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (lab[l0 + c] >= 0) {
        const int r = lds_root(lab, l0 + c);
        out[c] = (y0 + r / kCclT) * w + x0 + (r & (kCclT - 1)) + 1;
      }
  }
  if (vec) {
    *reinterpret_cast<int4*>(p + i0) = make_int4(out[0], out[1], out[2], out[3]);
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (in[c]) p[i0 + c] = out[c];
  }
}

// Border pass: unions across tile edges only (first row: NW, N, NE; first
// column: W, NW; last column: NE), on the global min-index union-find.
__global__ __launch_bounds__(128) void k_ccl_border(int w, int h, int* p,
                                                    const uint8_t* __restrict__ occ) {
  if (!occ[blockIdx.y * gridDim.x + blockIdx.x]) return;
  const int x0 = blockIdx.x * kCclT, y0 = blockIdx.y * kCclT;
  // 3 * kCclT border positions: first row, first column, last column
  for (int e = threadIdx.x; e < 3 * kCclT; e += 128) {
    int x, y;
    if (e < kCclT) {
      x = x0 + e;
      y = y0;
    } else if (e < 2 * kCclT) {
      x = x0;
      y = y0 + e - kCclT;
    } else {
      x = x0 + kCclT - 1;
      y = y0 + e - 2 * kCclT;
    }
    if (x >= w || y >= h) continue;
    const int i = y * w + x;
    if (p[i] == 0) continue;
    const bool first_row = y == y0, first_col = x == x0, last_col = x == x0 + kCclT - 1;
    if (first_col && x > 0 && p[i - 1] != 0) uf_unite(p, i, i - 1);
    if (y > 0) {
      if ((first_row || first_col) && x > 0 && p[i - w - 1] != 0) uf_unite(p, i, i - w - 1);
      if (first_row && p[i - w] != 0) uf_unite(p, i, i - w);
      if ((first_row || last_col) && x + 1 < w && p[i - w + 1] != 0) uf_unite(p, i, i - w + 1);
    }
  }
}

// Path compression, tile by tile (tiles without seeds hold only zeros).
__global__ __launch_bounds__(256) void k_compress(int w, int h, int* p,
                                                  const uint8_t* __restrict__ occ) {
  if (!occ[blockIdx.y * gridDim.x + blockIdx.x]) return;
  const int x0 = blockIdx.x * kCclT, y0 = blockIdx.y * kCclT;
  for (int l = threadIdx.x; l < kCclT * kCclT; l += 256) {
    const int x = x0 + (l & (kCclT - 1)), y = y0 + l / kCclT;
    if (x >= w || y >= h) continue;
    const int i = y * w + x;
    const int v = p[i];
    if (v == 0 || v == i + 1) continue;  // background, or already a root
    p[i] = uf_root(p, v - 1) + 1;
  }
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off);
    if (lane >= off) v += u;
  }
  return v;
}

// roots (p[i] == i + 1) in the 32-pixel segment of row y in tile column tx
__device__ __forceinline__ int segment_roots(const int* p, int w, int y, int tx) {
  int c = 0;
  const int x1 = min(w, (tx + 1) * kCclT);
  for (int x = tx * kCclT; x < x1; ++x) {
    const int i = y * w + x;
    c += p[i] == i + 1 ? 1 : 0;
  }
  return c;
}

// Root count of every row, one wave per row over the row's occupied tiles.
constexpr int kRowWaves = 4;

__global__ __launch_bounds__(64 * kRowWaves) void k_count_roots(int w, int h, int ntx,
                                                                const int* __restrict__ p,
                                                                const uint8_t* __restrict__ occ,
                                                                int* __restrict__ row_count) {
  const int lane = threadIdx.x & 63, y = blockIdx.x * kRowWaves + (threadIdx.x >> 6);
  if (y >= h) return;
  const uint8_t* orow = occ + size_t(y / kCclT) * ntx;
  int c = 0;
  for (int tx = lane; tx < ntx; tx += 64)
    if (orow[tx]) c += segment_roots(p, w, y, tx);
  c = wave_sum(c);
  if (lane == 0) row_count[y] = c;
}

// one workgroup: exclusive scan of the row counts; total in off[n]
__device__ int wg_exclusive_scan(int v, int* tmp, int& total) {
  // 256-thread exclusive scan through LDS (Hillis-Steele on 256 values)
  const int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int add = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += add;
    __syncthreads();
  }
  total = tmp[255];
  const int incl = tmp[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_rows(int n, const int* __restrict__ cnt,
                                                   int* __restrict__ off) {
  __shared__ int tmp[256];
  int carry = 0;
  for (int b = 0; b < n; b += 256 * 16) {
    int v[16], s = 0;
    const int i0 = b + threadIdx.x * 16;
    for (int j = 0; j < 16; ++j) {
      v[j] = i0 + j < n ? cnt[i0 + j] : 0;
      s += v[j];
    }
    int total;
    int run = carry + wg_exclusive_scan(s, tmp, total);
    for (int j = 0; j < 16; ++j) {
      if (i0 + j < n) off[i0 + j] = run;
      run += v[j];
    }
    carry += total;
  }
  if (threadIdx.x == 0) off[n] = carry;
}

// Roots take the marker -(rank + 1), rank = the number of roots before them in
// raster order: the row's offset, then the row's occupied segments in x
// order (a wave-wide scan of their counts), then x order within a segment.
__global__ __launch_bounds__(64 * kRowWaves) void k_rank_roots(int w, int h, int ntx,
                                                               int* __restrict__ p,
                                                               const uint8_t* __restrict__ occ,
                                                               const int* __restrict__ row_off) {
  const int lane = threadIdx.x & 63, y = blockIdx.x * kRowWaves + (threadIdx.x >> 6);
  if (y >= h) return;
  const uint8_t* orow = occ + size_t(y / kCclT) * ntx;
  int run = row_off[y];
  for (int t0 = 0; t0 < ntx; t0 += 64) {
    const int tx = t0 + lane;
    const bool live = tx < ntx && orow[tx];
    const int c = live ? segment_roots(p, w, y, tx) : 0;
    const int incl = wave_incl_scan(c, lane);
    int rank = run + incl - c;
    if (c > 0) {
      const int x1 = min(w, (tx + 1) * kCclT);
      for (int x = tx * kCclT; x < x1; ++x) {
        const int i = y * w + x;
        if (p[i] == i + 1) p[i] = -(++rank);
      }
    }
    run += __shfl(incl, 63);
  }
}

// Every seed pixel takes its root's rank + 1.  Roots hold -(rank+1) and flip
// to rank+1 here; a pixel reading its root's entry decodes either sign.
// Non-root entries are read by nobody but their own thread.
__global__ __launch_bounds__(256) void k_finalize(int w, int h, int* p,
                                                  const uint8_t* __restrict__ occ) {
  if (!occ[blockIdx.y * gridDim.x + blockIdx.x]) return;
  const int x0 = blockIdx.x * kCclT, y0 = blockIdx.y * kCclT;
  for (int l = threadIdx.x; l < kCclT * kCclT; l += 256) {
    const int x = x0 + (l & (kCclT - 1)), y = y0 + l / kCclT;
    if (x >= w || y >= h) continue;
    const int i = y * w + x;
    const int v = p[i];
    if (v < 0) {
      p[i] = -v;
    } else if (v > 0) {
      const int rv = ld_parent(p, v - 1);
      p[i] = rv < 0 ? -rv : rv;
    }
  }
}

}  // namespace

void run_matched_filter(int w, int h, const float* img, const float* psf, int ps, const float* var,
                        float* out, hipStream_t s) {
  if (ps <= 0 || ps > kMfMaxPs) invalid("matched_filter: psf side must be 1..63");
  ProfScope prof("matched_filter", s);
  switch (ps) {
#define DG_MF(P) \
  case P: launch_mf<P>(w, h, img, psf, ps, var, out, s); break;
    DG_MF(3) DG_MF(5) DG_MF(7) DG_MF(9) DG_MF(11) DG_MF(13) DG_MF(15) DG_MF(17) DG_MF(19)
    DG_MF(21) DG_MF(23) DG_MF(25) DG_MF(27) DG_MF(29) DG_MF(31) DG_MF(33) DG_MF(35)
#undef DG_MF
    default: launch_mf<0>(w, h, img, psf, ps, var, out, s);
  }
  DG_LAUNCH_CHECK();
}

// Scratch of one labelling: the seed occupancy of every 32 x 32 tile (one
// byte each), the root count of every row and their exclusive offsets (total
// at [h]), each rounded to 256 bytes.  All of it is rewritten by every call,
// so nothing needs clearing.
struct LabelScratch {
  uint8_t* occ;
  int* cnt;
  int* off;
};

size_t label_scratch_bytes(int w, int h) {
  const size_t ntiles = size_t((w + kCclT - 1) / kCclT) * ((h + kCclT - 1) / kCclT);
  return Arena::round(ntiles) + Arena::round(sizeof(int) * size_t(h)) +
         Arena::round(sizeof(int) * (size_t(h) + 1));
}

namespace {
LabelScratch carve_label_scratch(char* base, int w, int h) {
  const size_t ntiles = size_t((w + kCclT - 1) / kCclT) * ((h + kCclT - 1) / kCclT);
  char* cnt = base + Arena::round(ntiles);
  char* off = cnt + Arena::round(sizeof(int) * size_t(h));
  return {reinterpret_cast<uint8_t*>(base), reinterpret_cast<int*>(cnt),
          reinterpret_cast<int*>(off)};
}

// labels: device int32 w*h (also the union-find parents); returns the count.
int label_with(int w, int h, const float* score, const uint8_t* mask, double threshold,
               int polarity, int* labels, const LabelScratch& sc, hipStream_t s) {
  ProfScope prof("label", s);
  const int ntx = (w + kCclT - 1) / kCclT;
  const dim3 tiles(ntx, (h + kCclT - 1) / kCclT);
  const int rows = (h + kRowWaves - 1) / kRowWaves;
  hipLaunchKernelGGL(k_ccl_local, tiles, dim3(256), 0, s, w, h, score, mask,
                     static_cast<float>(threshold), polarity, labels, sc.occ);
  hipLaunchKernelGGL(k_ccl_border, tiles, dim3(128), 0, s, w, h, labels,
                     static_cast<const uint8_t*>(sc.occ));
  hipLaunchKernelGGL(k_compress, tiles, dim3(256), 0, s, w, h, labels,
                     static_cast<const uint8_t*>(sc.occ));
  hipLaunchKernelGGL(k_count_roots, dim3(rows), dim3(64 * kRowWaves), 0, s, w, h, ntx,
                     static_cast<const int*>(labels), static_cast<const uint8_t*>(sc.occ), sc.cnt);
  hipLaunchKernelGGL(k_scan_rows, dim3(1), dim3(256), 0, s, h, static_cast<const int*>(sc.cnt),
                     sc.off);
  hipLaunchKernelGGL(k_rank_roots, dim3(rows), dim3(64 * kRowWaves), 0, s, w, h, ntx, labels,
                     static_cast<const uint8_t*>(sc.occ), static_cast<const int*>(sc.off));
  hipLaunchKernelGGL(k_finalize, tiles, dim3(256), 0, s, w, h, labels,
                     static_cast<const uint8_t*>(sc.occ));
  DG_LAUNCH_CHECK();
  int total = 0;
  d2h(&total, sc.off + h, 1, s);
  DG_HIP(hipStreamSynchronize(s));
  return total;
}

void check_label_frame(int w, int h) {
  if (size_t(w) * h >= size_t(1) << 31) invalid("label: frame too large");
}

// ------------------------------------------------ allocation-free labelling
// The scratch of the allocation-free entry point: a file-scope device array,
// one copy per GPU at that GPU's own address, held until the process ends.
// Host code must never pass `scratch` itself to a kernel (its host-side name
// is not a device address): the address comes from hipGetSymbolAddress, once
// per device with that device current, and is kept in g_noalloc_ptr.
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;   // 1 MiB per GPU
constexpr size_t kNoAllocMaxPixels = size_t(1) << 28;      // 16384 x 16384
// a 16384 x 16384 frame needs 256 KiB of tile bytes and 128 KiB of row counts
static_assert(size_t(512) * 512 + 2 * 4 * 16385 + 512 <= kNoAllocScratchBytes,
              "a 16384 x 16384 frame's labelling scratch must fit the static array");
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

void* noalloc_scratch() {
  int dev = 0;
  DG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) runtime("label_noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      runtime(std::string("label_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_noalloc_ptr[dev] = p;
  }
  return g_noalloc_ptr[dev];
}
}  // namespace

int run_label(int w, int h, const float* score, const uint8_t* mask, double threshold,
              int polarity, int* labels, Arena& ar, hipStream_t s) {
  check_label_frame(w, h);
  if (size_t(w) * h == 0) return 0;
  char* base = ar.take<char>(label_scratch_bytes(w, h));
  return label_with(w, h, score, mask, threshold, polarity, labels,
                    carve_label_scratch(base, w, h), s);
}

// run_label on the device's static scratch: no device memory is allocated or
// freed.  The frame is checked against the array before anything else.
int run_label_noalloc(int w, int h, const float* score, const uint8_t* mask, double threshold,
                      int polarity, int* labels, hipStream_t s) {
  const size_t n = size_t(w) * h;
  if (n > kNoAllocMaxPixels)
    invalid("label_noalloc: a frame of " + std::to_string(w) + " x " + std::to_string(h) +
            " exceeds the allocation-free limit of 2^28 pixels (16384 x 16384)");
  if (w > 0 && h > 0 && label_scratch_bytes(w, h) > kNoAllocScratchBytes)
    invalid("label_noalloc: a frame of " + std::to_string(w) + " x " + std::to_string(h) +
            " needs " + std::to_string(label_scratch_bytes(w, h)) +
            " bytes of scratch (a byte per 32 x 32 tile, 8 per row), more than the 1 MiB "
            "static array");
  if (n == 0) return 0;
  char* base = static_cast<char*>(noalloc_scratch());
  return label_with(w, h, score, mask, threshold, polarity, labels,
                    carve_label_scratch(base, w, h), s);
}

}  // namespace dg
