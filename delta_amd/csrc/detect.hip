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
  float* win = mf_lds;   // wh x ww, zero padded at the frame edge
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * kMfW, y0 = blockIdx.y * kMfH;
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
      if (i < wh * ww) win[i] = v[j];
    }
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  float acc[kMfOut];
#pragma unroll
  for (int j = 0; j < kMfOut; ++j) acc[j] = 0.0f;
  for (int ly = 0; ly < ps; ++ly) {
    const float* row = win + (ty + ly) * ww + kMfOut * tx;
    if constexpr (PS > 0) {
      float v[PS + kMfOut - 1];
#pragma unroll
      for (int j = 0; j < PS + kMfOut - 1; ++j) v[j] = row[j];
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
        for (int j = 0; j < kMfOut; ++j) acc[j] = fmaf(p, row[lx + j], acc[j]);
      }
    }
  }
  float sumpsf2 = 0.0f;
  for (int i = 0; i < ps * ps; ++i) sumpsf2 = fmaf(psf[i], psf[i], sumpsf2);
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
  const size_t lds = sizeof(float) * size_t(kMfH + 2 * r) * (kMfW + 2 * r);
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

__global__ __launch_bounds__(256) void k_ccl_local(int w, int h, const float* __restrict__ score,
                                                   const uint8_t* __restrict__ mask, float thr,
                                                   int polarity, int* __restrict__ p) {
  __shared__ int lab[kCclT * kCclT];
  const int x0 = blockIdx.x * kCclT, y0 = blockIdx.y * kCclT;
  for (int l = threadIdx.x; l < kCclT * kCclT; l += 256) {
    const int x = x0 + (l & (kCclT - 1)), y = y0 + l / kCclT;
    bool pass = false;
    if (x < w && y < h) {
      const size_t i = size_t(y) * w + x;
      const float v = score[i];
      pass = isfinite(v) && (polarity > 0 ? v >= thr : v <= -thr);
      if (mask) pass = pass && mask[i] == 0;
    }
    lab[l] = pass ? l : -1;
  }
  __syncthreads();
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
  for (int l = threadIdx.x; l < kCclT * kCclT; l += 256) {
    const int x = x0 + (l & (kCclT - 1)), y = y0 + l / kCclT;
    if (x >= w || y >= h) continue;
    int v = 0;
    if (lab[l] >= 0) {
      const int r = lds_root(lab, l);
      v = (y0 + r / kCclT) * w + x0 + (r & (kCclT - 1)) + 1;
    }
    p[size_t(y) * w + x] = v;
  }
}

// Border pass: unions across tile edges only (first row: NW, N, NE; first
// column: W, NW; last column: NE), on the global min-index union-find.
__global__ __launch_bounds__(128) void k_ccl_border(int w, int h, int* p) {
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

__global__ void k_compress(int n, int* p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = p[i];
  if (v == 0 || v == i + 1) return;  // background, or already a root
  p[i] = uf_root(p, v - 1) + 1;
}

constexpr int kScanTile = 4096;  // pixels per tile of the root scan (256 threads x 16)

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

__global__ __launch_bounds__(256) void k_count_roots(int n, const int* __restrict__ p,
                                                     int* __restrict__ tile_count) {
  __shared__ int tmp[256];
  const int base = blockIdx.x * kScanTile + threadIdx.x * 16;
  int c = 0;
  for (int j = 0; j < 16; ++j) {
    const int i = base + j;
    if (i < n && p[i] == i + 1) ++c;
  }
  int total;
  wg_exclusive_scan(c, tmp, total);
  if (threadIdx.x == 0) tile_count[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tile counts; total in tile_off[ntile]
__global__ __launch_bounds__(256) void k_scan_tiles(int ntile, const int* __restrict__ cnt,
                                                    int* __restrict__ off) {
  __shared__ int tmp[256];
  int carry = 0;
  for (int b = 0; b < ntile; b += 256 * 16) {
    int v[16], s = 0;
    const int i0 = b + threadIdx.x * 16;
    for (int j = 0; j < 16; ++j) {
      v[j] = i0 + j < ntile ? cnt[i0 + j] : 0;
      s += v[j];
    }
    int total;
    int run = carry + wg_exclusive_scan(s, tmp, total);
    for (int j = 0; j < 16; ++j) {
      if (i0 + j < ntile) off[i0 + j] = run;
      run += v[j];
    }
    carry += total;
  }
  if (threadIdx.x == 0) off[ntile] = carry;
}

// roots take the marker -(rank + 1)
__global__ __launch_bounds__(256) void k_rank_roots(int n, int* __restrict__ p,
                                                    const int* __restrict__ tile_off) {
  __shared__ int tmp[256];
  const int base = blockIdx.x * kScanTile + threadIdx.x * 16;
  int c = 0;
  for (int j = 0; j < 16; ++j) {
    const int i = base + j;
    if (i < n && p[i] == i + 1) ++c;
  }
  int total;
  int rank = tile_off[blockIdx.x] + wg_exclusive_scan(c, tmp, total);
  for (int j = 0; j < 16; ++j) {
    const int i = base + j;
    if (i < n && p[i] == i + 1) p[i] = -(++rank);
  }
}

// Every seed pixel takes its root's rank + 1.  Roots hold -(rank+1) and flip
// to rank+1 here; a pixel reading its root's entry decodes either sign.
// Non-root entries are read by nobody but their own thread.
__global__ void k_finalize(int n, int* p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = p[i];
  if (v < 0) {
    p[i] = -v;
  } else if (v > 0) {
    const int rv = ld_parent(p, v - 1);
    p[i] = rv < 0 ? -rv : rv;
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

// Scratch of one labelling: the root count of every scan tile and their
// exclusive offsets (total at [ntile]), each rounded to 256 bytes.  Both are
// fully rewritten by every call, so nothing needs clearing.
struct LabelScratch {
  int* cnt;
  int* off;
};

size_t label_scratch_bytes(int w, int h) {
  const size_t n = size_t(w) * h;
  const size_t ntile = (n + kScanTile - 1) / kScanTile;
  return 2 * Arena::round(sizeof(int) * (ntile + 1));
}

namespace {
LabelScratch carve_label_scratch(char* base, int w, int h) {
  const size_t ntile = (size_t(w) * h + kScanTile - 1) / kScanTile;
  return {reinterpret_cast<int*>(base),
          reinterpret_cast<int*>(base + Arena::round(sizeof(int) * (ntile + 1)))};
}

// labels: device int32 w*h (also the union-find parents); returns the count.
int label_with(int w, int h, const float* score, const uint8_t* mask, double threshold,
               int polarity, int* labels, const LabelScratch& sc, hipStream_t s) {
  const int n = w * h;
  const int ntile = (n + kScanTile - 1) / kScanTile;
  ProfScope prof("label", s);
  const dim3 tiles((w + kCclT - 1) / kCclT, (h + kCclT - 1) / kCclT);
  hipLaunchKernelGGL(k_ccl_local, tiles, dim3(256), 0, s, w, h, score, mask,
                     static_cast<float>(threshold), polarity, labels);
  hipLaunchKernelGGL(k_ccl_border, tiles, dim3(128), 0, s, w, h, labels);
  hipLaunchKernelGGL(k_compress, dim3((n + 255) / 256), dim3(256), 0, s, n, labels);
  hipLaunchKernelGGL(k_count_roots, dim3(ntile), dim3(256), 0, s, n,
                     static_cast<const int*>(labels), sc.cnt);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, ntile,
                     static_cast<const int*>(sc.cnt), sc.off);
  hipLaunchKernelGGL(k_rank_roots, dim3(ntile), dim3(256), 0, s, n, labels,
                     static_cast<const int*>(sc.off));
  hipLaunchKernelGGL(k_finalize, dim3((n + 255) / 256), dim3(256), 0, s, n, labels);
  DG_LAUNCH_CHECK();
  int total = 0;
  d2h(&total, sc.off + ntile, 1, s);
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
  static_assert(2 * ((sizeof(int) * ((kNoAllocMaxPixels + kScanTile - 1) / kScanTile + 1) +
                      255) & ~size_t(255)) <= kNoAllocScratchBytes,
                "the largest allowed frame's scratch must fit the static array");
  if (label_scratch_bytes(w, h) > kNoAllocScratchBytes)
    invalid("label_noalloc: scratch of this frame exceeds the static array");
  if (n == 0) return 0;
  char* base = static_cast<char*>(noalloc_scratch());
  return label_with(w, h, score, mask, threshold, polarity, labels,
                    carve_label_scratch(base, w, h), s);
}

}  // namespace dg
