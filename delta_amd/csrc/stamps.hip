// Stamp selection on device (reference src/detect.cpp:84-314): robust
// background, peak detection, brightest-first isolation, FWHM moments, and the
// science/reference cross-match or catalog validation.
//
// Detection is one full-frame pass (strict 3x3 maxima above threshold, raster
// index appended to a list), one wave per peak for the footprint test, a
// bitonic sort of 64-bit keys (~orderable(value) : raster index, so ascending
// order is value descending, raster index ascending), and a one-wave greedy
// isolation over the sorted keys that stops at max_stamps.  The threshold and
// every per-stamp statistic are computed in the reference's order and
// rounding (FP64, multiply-adds fused exactly where the reference's compiler
// fuses them), so the results are bit-identical to the reference's.
#include "../../include/deltagpu.h"
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr size_t kTargetSamples = size_t(1) << 18;  // src/detect.cpp:96
constexpr uint32_t kPeakCap0 = 1u << 18;           // first peak-list capacity
constexpr int kSortLocal = 2048;                   // keys per LDS bitonic block
constexpr int kKeptLds = 8192;                     // kept stamps mirrored in LDS
constexpr double kFwhmPerSigma = 2.354820045030949;

__device__ __forceinline__ uint32_t orderable(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ------------------------------------------------------------------ bitonic sort
__device__ __forceinline__ uint32_t pair_lo(uint32_t t, uint32_t j) {
  return ((t & ~(j - 1)) << 1) | (t & (j - 1));
}

// Every step (k, j) of the ascending bitonic network with k in [k_lo, k_hi]
// and j < chunk, on one chunk of `chunk` keys in LDS (chunk/2 threads).
__global__ __launch_bounds__(kSortLocal / 2) void k_bitonic_local(
    unsigned long long* __restrict__ keys, uint32_t chunk, uint32_t k_lo, uint32_t k_hi) {
  __shared__ unsigned long long sh[kSortLocal];
  const uint32_t t = threadIdx.x;
  const size_t base = size_t(blockIdx.x) * chunk;
  sh[t] = keys[base + t];
  sh[t + chunk / 2] = keys[base + t + chunk / 2];
  __syncthreads();
  for (uint32_t k = k_lo; k <= k_hi; k <<= 1) {
    for (uint32_t j = min(k, chunk) >> 1; j > 0; j >>= 1) {
      const uint32_t i = pair_lo(t, j);
      const bool asc = ((base + i) & k) == 0;
      const unsigned long long a = sh[i], b = sh[i + j];
      if ((a > b) == asc) {
        sh[i] = b;
        sh[i + j] = a;
      }
      __syncthreads();
    }
  }
  keys[base + t] = sh[t];
  keys[base + t + chunk / 2] = sh[t + chunk / 2];
}

__global__ __launch_bounds__(256) void k_bitonic_global(unsigned long long* __restrict__ keys,
                                                        uint32_t half, uint32_t k, uint32_t j) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half) return;
  const uint32_t i = pair_lo(t, j);
  const bool asc = (i & k) == 0;
  const unsigned long long a = keys[i], b = keys[i + j];
  if ((a > b) == asc) {
    keys[i] = b;
    keys[i + j] = a;
  }
}

}  // namespace

// Ascending sort of npad (a power of two >= 2) 64-bit keys in place.
void sort_keys_u64(unsigned long long* keys, uint32_t npad, hipStream_t s) {
  const uint32_t chunk = std::min<uint32_t>(npad, kSortLocal);
  const uint32_t nchunk = npad / chunk;
  hipLaunchKernelGGL(k_bitonic_local, dim3(nchunk), dim3(chunk / 2), 0, s, keys, chunk, 2u,
                     chunk);
  for (uint32_t k = 2 * chunk; k <= npad; k <<= 1) {
    for (uint32_t j = k >> 1; j >= chunk; j >>= 1)
      hipLaunchKernelGGL(k_bitonic_global, dim3((npad / 2 + 255) / 256), dim3(256), 0, s, keys,
                         npad / 2, k, j);
    hipLaunchKernelGGL(k_bitonic_local, dim3(nchunk), dim3(chunk / 2), 0, s, keys, chunk, k, k);
  }
}

namespace {

// ------------------------------------------------------------------ background
__global__ __launch_bounds__(256) void k_bg_gather(const float* __restrict__ img,
                                                   const uint8_t* __restrict__ mask,
                                                   size_t nsamp, size_t step,
                                                   float* __restrict__ sample) {
  for (size_t j = blockIdx.x * size_t(blockDim.x) + threadIdx.x; j < nsamp;
       j += size_t(gridDim.x) * blockDim.x) {
    const size_t i = j * step;
    const float v = img[i];
    const bool ok = isfinite(v) && !(mask && mask[i] != 0);
    sample[j] = ok ? v : __builtin_nanf("");
  }
}

// |v - (float)median| in float, as src/detect.cpp:110; NaN (invalid) stays NaN
// and an overflow to inf is dropped by the selection, as median_of drops it.
__global__ __launch_bounds__(256) void k_bg_deviation(float* __restrict__ sample, size_t nsamp,
                                                      const float* __restrict__ med) {
  const float m = *med;
  for (size_t j = blockIdx.x * size_t(blockDim.x) + threadIdx.x; j < nsamp;
       j += size_t(gridDim.x) * blockDim.x)
    sample[j] = fabsf(sample[j] - m);
}

// bg = {median, sigma, threshold}: src/detect.cpp:109-112 and :164-165, the
// threshold's product rounded before the add, as the reference's compiler
// leaves it (its conditional operand keeps the multiply-add apart).
__global__ void k_bg_final(const float* __restrict__ medf, const float* __restrict__ madf,
                           double thr_sigma, double* __restrict__ bg) {
  const double med = static_cast<double>(*medf);
  const double sigma = madf ? 1.4826 * static_cast<double>(*madf) : 0.0;
  bg[0] = med;
  bg[1] = sigma;
  double prod = thr_sigma * (sigma > 0.0 ? sigma : 1.0);
  asm volatile("" : "+v"(prod));  // the library builds with -ffp-contract=fast
  bg[2] = med + prod;
}

// ------------------------------------------------------------------ detection
// Strict 3x3 maxima above the threshold inside [border, w-border) x
// [border, h-border) (src/detect.cpp:187-202); the border is >= 1 so every
// neighbour is on the frame.
__global__ __launch_bounds__(256) void k_peaks(const float* __restrict__ img, int w, int border,
                                               int segs, const double* __restrict__ bg,
                                               uint32_t* __restrict__ list, uint32_t cap,
                                               uint32_t* __restrict__ count) {
  const int seg = blockIdx.x % segs;
  const int y = border + blockIdx.x / segs;
  const int x = border + seg * 256 + threadIdx.x;
  if (x >= w - border) return;
  const size_t i = size_t(y) * w + x;
  const float v = img[i];
  if (!(static_cast<double>(v) > bg[2])) return;
  const float* up = img + i - w;
  const float* dn = img + i + w;
  if (up[-1] > v || up[0] > v || up[1] > v || img[i - 1] > v || img[i + 1] > v ||
      dn[-1] > v || dn[0] > v || dn[1] > v)
    return;
  const uint32_t pos = atomicAdd(count, 1u);
  if (pos < cap) list[pos] = static_cast<uint32_t>(i);
}

// src/detect.cpp:55-75 over one wave: on-frame, finite, unmasked, below
// saturation (when > 0).  A negative radius checks the bounds only, as there.
__device__ __forceinline__ bool footprint_wave(const float* __restrict__ img,
                                               const uint8_t* __restrict__ mask, int w, int h,
                                               int x, int y, int r, double sat, int lane) {
  if (x - r < 0 || x + r >= w || y - r < 0 || y + r >= h) return false;
  if (r < 0) return true;
  const uint32_t d = 2u * r + 1u;
  const uint32_t cnt = d * d;
  bool bad = false;
  for (uint32_t e = lane; e < cnt; e += 64) {
    const int yy = y - r + static_cast<int>(e / d);
    const int xx = x - r + static_cast<int>(e % d);
    const size_t i = size_t(yy) * w + xx;
    const float v = img[i];
    bad |= !isfinite(v) || (mask && mask[i] != 0) || (sat > 0.0 && static_cast<double>(v) >= sat);
  }
  return __ballot(bad) == 0;
}

__global__ __launch_bounds__(256) void k_peak_footprint(
    const float* __restrict__ img, const uint8_t* __restrict__ mask, int w, int h, int r,
    double sat, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
    uint32_t cap, unsigned long long* __restrict__ keys, uint32_t* __restrict__ nkeys) {
  const uint32_t n = min(*count, cap);
  const int lane = threadIdx.x & 63;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t e = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; e < n; e += nw) {
    const uint32_t idx = list[e];
    const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
    if (footprint_wave(img, mask, w, h, x, y, r, sat, lane) && lane == 0) {
      const uint32_t pos = atomicAdd(nkeys, 1u);
      keys[pos] = (static_cast<unsigned long long>(~orderable(img[idx])) << 32) | idx;
    }
  }
}

__global__ __launch_bounds__(256) void k_pad_keys(unsigned long long* __restrict__ keys,
                                                  const uint32_t* __restrict__ nkeys,
                                                  uint32_t npad) {
  for (uint32_t i = *nkeys + blockIdx.x * blockDim.x + threadIdx.x; i < npad;
       i += gridDim.x * blockDim.x)
    keys[i] = ~0ull;
}

// ------------------------------------------------------------------ isolation
__device__ __forceinline__ unsigned long long pack_xy(int x, int y) {
  return static_cast<uint32_t>(x) | (static_cast<unsigned long long>(static_cast<uint32_t>(y)) << 32);
}
__device__ __forceinline__ int unpack_x(unsigned long long p) { return static_cast<int>(uint32_t(p)); }
__device__ __forceinline__ int unpack_y(unsigned long long p) {
  return static_cast<int>(uint32_t(p >> 32));
}

// Brightest-first greedy isolation (src/detect.cpp:223-240) in one wave: 64
// sorted candidates at a time are tested against every stamp kept so far, then
// resolved among themselves in order.  Kept positions live in LDS (the first
// kKeptLds) and in `kept_xy`, read back at device scope beyond that.
__global__ __launch_bounds__(64) void k_isolate(const unsigned long long* __restrict__ keys,
                                                const uint32_t* __restrict__ nkeys, int w,
                                                long long iso2, int max_stamps,
                                                uint32_t* __restrict__ kept_idx,
                                                unsigned long long* kept_xy,
                                                uint32_t* __restrict__ nkept) {
  __shared__ unsigned long long s_kept[kKeptLds];
  const int lane = threadIdx.x;
  const uint32_t n = *nkeys;
  int kept = 0;
  for (uint32_t base = 0; base < n && kept < max_stamps; base += 64) {
    const uint32_t c = base + lane;
    const bool valid = c < n;
    const uint32_t idx = valid ? static_cast<uint32_t>(keys[c]) : 0u;
    const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
    bool alive = valid;
    for (int k = 0; k < kept && alive; ++k) {
      const unsigned long long q =
          k < kKeptLds ? s_kept[k]
                       : __hip_atomic_load(&kept_xy[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const long long dx = x - unpack_x(q), dy = y - unpack_y(q);
      if (dx * dx + dy * dy < iso2) alive = false;
    }
    // conflicts with earlier candidates of this chunk
    unsigned long long conf = 0;
    for (int j = 0; j < 64; ++j) {
      const int xj = __shfl(x, j), yj = __shfl(y, j);
      const long long dx = x - xj, dy = y - yj;
      if (j < lane && dx * dx + dy * dy < iso2) conf |= 1ull << j;
    }
    const unsigned long long alive_mask = __ballot(alive);
    const uint32_t conf_lo = static_cast<uint32_t>(conf), conf_hi = static_cast<uint32_t>(conf >> 32);
    unsigned long long keep = 0;
    int total = kept;
    for (int i = 0; i < 64 && total < max_stamps; ++i) {
      if (!((alive_mask >> i) & 1ull)) continue;
      const unsigned long long ci =
          static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(conf_lo), i)) |
          (static_cast<unsigned long long>(
               static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(conf_hi), i)))
           << 32);
      if ((ci & keep) == 0) {
        keep |= 1ull << i;
        ++total;
      }
    }
    if ((keep >> lane) & 1ull) {
      const int pos = kept + __popcll(keep & ((1ull << lane) - 1ull));
      const unsigned long long p = pack_xy(x, y);
      kept_idx[pos] = idx;
      __hip_atomic_store(&kept_xy[pos], p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (pos < kKeptLds) s_kept[pos] = p;
    }
    kept = total;
    __syncthreads();
  }
  if (lane == 0) *nkept = static_cast<uint32_t>(kept);
}

// src/detect.cpp:115-157 in the reference's summation order; its compiler
// fuses `s += a * b` into fma(a, b, s), written out here, and no other
// product here feeds an add.
__device__ double fwhm_at(const float* __restrict__ img, int w, int h, int x, int y, int r,
                          double bg) {
  double sum = 0.0, sx = 0.0, sy = 0.0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    const float* row = img + size_t(yy) * w;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= w) continue;
      const double f = static_cast<double>(row[xx]) - bg;
      const double wt = f > 0.0 ? f : 0.0;
      sum += wt;
      sx = fma(wt, static_cast<double>(dx), sx);
      sy = fma(wt, static_cast<double>(dy), sy);
    }
  }
  if (sum <= 0.0) return __builtin_nan("");
  const double mx = sx / sum, my = sy / sum;
  double sxx = 0.0, syy = 0.0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    const float* row = img + size_t(yy) * w;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= w) continue;
      const double f = static_cast<double>(row[xx]) - bg;
      const double wt = f > 0.0 ? f : 0.0;
      const double ex = dx - mx, ey = dy - my;
      sxx = fma(wt * ex, ex, sxx);
      syy = fma(wt * ey, ey, syy);
    }
  }
  const double var = 0.5 * (sxx / sum + syy / sum);
  if (var <= 0.0) return __builtin_nan("");
  return kFwhmPerSigma * sqrt(var);
}

// x, y, flux, snr, fwhm of each kept stamp (src/detect.cpp:206-211).
__global__ __launch_bounds__(64) void k_stamp_stats(
    const float* __restrict__ img, int w, int h, int frad, const double* __restrict__ bg,
    const uint32_t* __restrict__ kept_idx, const uint32_t* __restrict__ nkept,
    int32_t* __restrict__ ox, int32_t* __restrict__ oy, double* __restrict__ flux,
    double* __restrict__ snr, double* __restrict__ fwhm) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *nkept) return;
  const uint32_t idx = kept_idx[i];
  const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
  const double med = bg[0], sigma = bg[1];
  const double f = static_cast<double>(img[idx]) - med;
  if (ox) ox[i] = x;
  if (oy) oy[i] = y;
  if (flux) flux[i] = f;
  if (snr) snr[i] = sigma > 0.0 ? f / sigma : f;
  if (fwhm) fwhm[i] = fwhm_at(img, w, h, x, y, frad, med);
}

// ------------------------------------------------------------------ selection
// For each reference stamp the first science stamp at the least squared
// distance <= tol2 (src/detect.cpp:268-279), or -1.
__global__ __launch_bounds__(256) void k_match(const int32_t* __restrict__ rx,
                                               const int32_t* __restrict__ ry, int nr,
                                               const int32_t* __restrict__ sx,
                                               const int32_t* __restrict__ sy, int ns,
                                               long long tol2, int* __restrict__ best) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  int b = -1;
  long long bd = tol2 + 1;
  for (int s = 0; s < ns; ++s) {
    const long long dx = rx[r] - sx[s], dy = ry[r] - sy[s];
    const long long d2 = dx * dx + dy * dy;
    if (d2 <= tol2 && d2 < bd) {
      bd = d2;
      b = s;
    }
  }
  best[r] = b;
}

// One wave per catalog entry: footprint passes in both images (:300-305).
__global__ __launch_bounds__(256) void k_catalog_check(
    const float* __restrict__ sci, const uint8_t* __restrict__ smask, int ws, int hs,
    const float* __restrict__ ref, const uint8_t* __restrict__ rmask, int wr, int hr,
    const int32_t* __restrict__ cx, const int32_t* __restrict__ cy, int n, int r, double sat,
    int* __restrict__ best) {
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int e = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; e < n; e += nw) {
    const int x = cx[e], y = cy[e];
    const bool ok = footprint_wave(sci, smask, ws, hs, x, y, r, sat, lane) &&
                    footprint_wave(ref, rmask, wr, hr, x, y, r, sat, lane);
    if (lane == 0) best[e] = ok ? e : -1;
  }
}

// Order-preserving compaction of the entries with best[i] >= 0 (one block):
// out_x/out_y from (px, py)[i], fwhm_sci from fs[best[i]] and fwhm_ref from
// fr[i] when given, else computed at (x, y) against each background median.
__global__ __launch_bounds__(1024) void k_compact_selection(
    const int* __restrict__ best, int n, const int32_t* __restrict__ px,
    const int32_t* __restrict__ py, const double* __restrict__ fs, const double* __restrict__ fr,
    int32_t* __restrict__ ox, int32_t* __restrict__ oy, double* __restrict__ ofs,
    double* __restrict__ ofr, int32_t* __restrict__ src, int* __restrict__ nout) {
  __shared__ int s_wave[16];
  __shared__ int s_base;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t == 0) s_base = 0;
  __syncthreads();
  for (int c0 = 0; c0 < n; c0 += 1024) {
    const int i = c0 + t;
    const bool take = i < n && best[i] >= 0;
    const unsigned long long bal = __ballot(take);
    if (lane == 0) s_wave[wv] = __popcll(bal);
    __syncthreads();
    int off = s_base;
    for (int k = 0; k < wv; ++k) off += s_wave[k];
    if (take) {
      const int pos = off + __popcll(bal & ((1ull << lane) - 1ull));
      ox[pos] = px[i];
      oy[pos] = py[i];
      if (fs) ofs[pos] = fs[best[i]];
      if (fr) ofr[pos] = fr[i];
      src[pos] = i;
    }
    __syncthreads();
    if (t == 0) {
      int tot = 0;
      for (int k = 0; k < 16; ++k) tot += s_wave[k];
      s_base += tot;
    }
    __syncthreads();
  }
  if (t == 0) *nout = s_base;
}

// FWHM of the catalog path at the compacted positions.
__global__ __launch_bounds__(64) void k_catalog_fwhm(
    const float* __restrict__ sci, int ws, int hs, const float* __restrict__ ref, int wr, int hr,
    const int32_t* __restrict__ x, const int32_t* __restrict__ y, const int* __restrict__ n,
    int frad, const double* __restrict__ bg_s, const double* __restrict__ bg_r,
    double* __restrict__ fs, double* __restrict__ fr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *n) return;
  fs[i] = fwhm_at(sci, ws, hs, x[i], y[i], frad, bg_s[0]);
  fr[i] = fwhm_at(ref, wr, hr, x[i], y[i], frad, bg_r[0]);
}

// Element m/2 of the finite values of a[0..n) in ascending order (the
// nth_element statistic of src/detect.cpp:28-40), NaN when none is finite:
// a value is it when fewer than m/2+1 values are below it and more than m/2
// are at or below it.  One block per array.
__global__ __launch_bounds__(1024) void k_upper_median(const double* __restrict__ a0,
                                                       const double* __restrict__ a1,
                                                       const int* __restrict__ np,
                                                       double* __restrict__ out) {
  const double* a = blockIdx.x == 0 ? a0 : a1;
  const int n = *np;
  __shared__ int s_m;
  __shared__ double s_med;
  if (threadIdx.x == 0) {
    s_m = 0;
    s_med = __builtin_nan("");
  }
  __syncthreads();
  int local = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) local += isfinite(a[i]) ? 1 : 0;
  atomicAdd(&s_m, local);
  __syncthreads();
  const int half = s_m / 2;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = a[i];
    if (!isfinite(v)) continue;
    int lt = 0, le = 0;
    for (int j = 0; j < n; ++j) {
      const double u = a[j];
      if (!isfinite(u)) continue;
      lt += u < v;
      le += u <= v;
    }
    if (lt <= half && half < le) s_med = v;  // equal values: any writer is right
  }
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = s_med;
}

int blocks_for(size_t n, int per) {
  return static_cast<int>(std::min<size_t>(2048, std::max<size_t>(1, (n + per - 1) / per)));
}

uint32_t next_pow2(uint32_t v) {
  uint32_t p = 2;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

// ------------------------------------------------------------------ host side
size_t background_scratch_bytes(size_t n) {
  const size_t step = n > kTargetSamples ? n / kTargetSamples : 1;
  const size_t nsamp = (n + step - 1) / step;
  return Arena::round(sizeof(float) * nsamp) + Arena::round(sizeof(float) * 2) +
         2 * select_scratch_bytes(nsamp, 1) + Arena::round(sizeof(double) * 4);
}

// bg (device, 3 doubles) = {median, sigma, threshold}; sigma (and the MAD
// pass) only when want_sigma.
void run_background(const float* img, const uint8_t* mask, size_t n, double thr_sigma,
                    bool want_sigma, double* bg, Arena& ar, hipStream_t s) {
  const size_t step = n > kTargetSamples ? n / kTargetSamples : 1;
  const size_t nsamp = (n + step - 1) / step;
  float* sample = ar.take<float>(nsamp);
  float* stat = ar.take<float>(2);
  ProfScope prof("background", s);
  hipLaunchKernelGGL(k_bg_gather, dim3(blocks_for(nsamp, 256)), dim3(256), 0, s, img, mask,
                     nsamp, step, sample);
  select_strided_median(sample, nullptr, nsamp, 1, stat, ar, s);
  if (want_sigma) {
    hipLaunchKernelGGL(k_bg_deviation, dim3(blocks_for(nsamp, 256)), dim3(256), 0, s, sample,
                       nsamp, static_cast<const float*>(stat));
    select_strided_median(sample, nullptr, nsamp, 1, stat + 1, ar, s);
  }
  hipLaunchKernelGGL(k_bg_final, dim3(1), dim3(1), 0, s, static_cast<const float*>(stat),
                     want_sigma ? static_cast<const float*>(stat + 1) : nullptr, thr_sigma, bg);
  DG_LAUNCH_CHECK();
}

namespace {
int resolved_fwhm_radius(const dg_detect_params& p) {
  return p.fwhm_radius > 0 ? p.fwhm_radius : std::max(5, p.stamp_radius / 2);
}

void check_frame(int w, int h, const char* what) {
  if (w <= 0 || h <= 0) invalid(std::string(what) + ": empty image");
  if (size_t(w) * size_t(h) >= (size_t(1) << 32))
    invalid(std::string(what) + ": frame above 2^32 pixels");
}
}  // namespace

// Detection into device arrays of capacity >= max_stamps; returns the count.
int run_detect_stamps(const float* img, const uint8_t* mask, int w, int h,
                      const dg_detect_params& p, int32_t* ox, int32_t* oy, double* flux,
                      double* snr, double* fwhm, hipStream_t s) {
  check_frame(w, h, "detect_stamps");
  const int border = p.border > 0 ? p.border : p.stamp_radius;
  if (border < 1)
    invalid("detect_stamps: border (or stamp_radius when border is 0) must be >= 1");
  const size_t n = size_t(w) * h;
  const int sw = w - 2 * border, sh = h - 2 * border;
  const int kcap = std::max(p.max_stamps, 0);

  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(background_scratch_bytes(n) + Arena::round(sizeof(uint32_t) * kPeakCap0) +
             Arena::round(sizeof(uint32_t) * 4) +
             Arena::round(sizeof(unsigned long long) * kPeakCap0) +
             Arena::round(sizeof(uint32_t) * (kcap + 1)) +
             Arena::round(sizeof(unsigned long long) * (kcap + 1)));
  double* bg = ar.take<double>(4);
  uint32_t* cnt = ar.take<uint32_t>(4);  // peaks, keys, kept
  uint32_t* kept_idx = ar.take<uint32_t>(kcap + 1);
  unsigned long long* kept_xy = ar.take<unsigned long long>(kcap + 1);
  uint32_t* list = ar.take<uint32_t>(kPeakCap0);
  unsigned long long* keys = ar.take<unsigned long long>(kPeakCap0);
  DG_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * 4, s));
  const size_t mark = ar.mark();
  run_background(img, mask, n, p.threshold_sigma, true, bg, ar, s);
  ar.rewind(mark);
  if (sw <= 0 || sh <= 0 || kcap == 0) {
    DG_HIP(hipStreamSynchronize(s));
    return 0;
  }
  const int segs = (sw + 255) / 256;
  ProfScope prof("detect", s);
  hipLaunchKernelGGL(k_peaks, dim3(segs * sh), dim3(256), 0, s, img, w, border, segs,
                     static_cast<const double*>(bg), list, kPeakCap0, cnt);
  DG_LAUNCH_CHECK();
  uint32_t npeaks = 0;
  d2h(&npeaks, cnt, 1, s);
  DG_HIP(hipStreamSynchronize(s));
  if (npeaks == 0) return 0;

  ArenaLease big;  // only when the first list overflowed
  uint32_t cap = kPeakCap0;
  if (npeaks > kPeakCap0) {
    cap = npeaks;
    const uint32_t npad = next_pow2(cap);
    big->reserve(Arena::round(sizeof(uint32_t) * cap) +
                 Arena::round(sizeof(unsigned long long) * npad));
    list = big->take<uint32_t>(cap);
    keys = big->take<unsigned long long>(npad);
    DG_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_peaks, dim3(segs * sh), dim3(256), 0, s, img, w, border, segs,
                       static_cast<const double*>(bg), list, cap, cnt);
  }
  const uint32_t npad = next_pow2(npeaks);
  hipLaunchKernelGGL(k_peak_footprint, dim3(blocks_for(npeaks, 4)), dim3(256), 0, s, img, mask, w,
                     h, p.stamp_radius, p.saturation, static_cast<const uint32_t*>(list),
                     static_cast<const uint32_t*>(cnt), cap, keys, cnt + 1);
  hipLaunchKernelGGL(k_pad_keys, dim3(blocks_for(npad, 256)), dim3(256), 0, s, keys,
                     static_cast<const uint32_t*>(cnt + 1), npad);
  sort_keys_u64(keys, npad, s);
  const int iso = p.isolation_radius > 0 ? p.isolation_radius : 2 * p.stamp_radius;
  hipLaunchKernelGGL(k_isolate, dim3(1), dim3(64), 0, s,
                     static_cast<const unsigned long long*>(keys),
                     static_cast<const uint32_t*>(cnt + 1), w,
                     static_cast<long long>(iso) * iso, kcap, kept_idx, kept_xy, cnt + 2);
  hipLaunchKernelGGL(k_stamp_stats, dim3((kcap + 63) / 64), dim3(64), 0, s, img, w, h,
                     resolved_fwhm_radius(p), static_cast<const double*>(bg),
                     static_cast<const uint32_t*>(kept_idx), static_cast<const uint32_t*>(cnt + 2),
                     ox, oy, flux, snr, fwhm);
  DG_LAUNCH_CHECK();
  uint32_t nk = 0;
  d2h(&nk, cnt + 2, 1, s);
  DG_HIP(hipStreamSynchronize(s));
  return static_cast<int>(nk);
}

// Background only: host (median, sigma).
void run_estimate_background(const float* img, const uint8_t* mask, int w, int h, double* median,
                             double* sigma, hipStream_t s) {
  check_frame(w, h, "estimate_background");
  const size_t n = size_t(w) * h;
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(background_scratch_bytes(n) + Arena::round(sizeof(double) * 4));
  double* bg = ar.take<double>(4);
  run_background(img, mask, n, 0.0, true, bg, ar, s);
  double hb[3];
  d2h(hb, bg, 3, s);
  DG_HIP(hipStreamSynchronize(s));
  *median = hb[0];
  *sigma = hb[1];
}

// Selection into device arrays of capacity >= max_stamps (detection) or
// n_catalog (catalog); host medians[2] and direction.
int run_select_stamps(int ws, int hs, const float* sci, const uint8_t* smask, int wr, int hr,
                      const float* ref, const uint8_t* rmask, const int32_t* cat_x,
                      const int32_t* cat_y, int n_cat, const dg_detect_params& p, int32_t* ox,
                      int32_t* oy, double* ofs, double* ofr, double* medians, int* direction,
                      hipStream_t s) {
  check_frame(ws, hs, "select_stamps");
  check_frame(wr, hr, "select_stamps");
  const bool catalog = cat_x && cat_y;
  const int nin = catalog ? std::max(n_cat, 0) : std::max(p.max_stamps, 0);
  const size_t ni = Arena::round(sizeof(int32_t) * (nin + 1));
  const size_t nd = Arena::round(sizeof(double) * (nin + 1));
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(Arena::round(sizeof(double) * 16) + Arena::round(sizeof(int) * 2) + 6 * ni +
             2 * nd +
             (catalog ? background_scratch_bytes(std::max(size_t(ws) * hs, size_t(wr) * hr))
                      : 0));
  double* misc = ar.take<double>(16);
  double *bgs = misc, *bgr = misc + 4, *med = misc + 8;
  int* nout = ar.take<int>(2);
  int* best = ar.take<int>(nin + 1);
  int32_t* src = ar.take<int32_t>(nin + 1);
  int32_t *sx = ar.take<int32_t>(nin + 1), *sy = ar.take<int32_t>(nin + 1);
  int32_t *rx = ar.take<int32_t>(nin + 1), *ry = ar.take<int32_t>(nin + 1);
  double *fs = ar.take<double>(nin + 1), *fr = ar.take<double>(nin + 1);
  DG_HIP(hipMemsetAsync(nout, 0, sizeof(int) * 2, s));
  if (catalog) {
    const size_t mark = ar.mark();
    run_background(sci, smask, size_t(ws) * hs, 0.0, false, bgs, ar, s);
    ar.rewind(mark);
    run_background(ref, rmask, size_t(wr) * hr, 0.0, false, bgr, ar, s);
    ar.rewind(mark);
    ProfScope prof("select_stamps", s);
    if (nin > 0) {
      hipLaunchKernelGGL(k_catalog_check, dim3(blocks_for(nin, 4)), dim3(256), 0, s, sci, smask,
                         ws, hs, ref, rmask, wr, hr, cat_x, cat_y, nin, p.stamp_radius,
                         p.saturation, best);
      hipLaunchKernelGGL(k_compact_selection, dim3(1), dim3(1024), 0, s,
                         static_cast<const int*>(best), nin, cat_x, cat_y, nullptr, nullptr, ox,
                         oy, nullptr, nullptr, src, nout);
      hipLaunchKernelGGL(k_catalog_fwhm, dim3((nin + 63) / 64), dim3(64), 0, s, sci, ws, hs, ref,
                         wr, hr, static_cast<const int32_t*>(ox), static_cast<const int32_t*>(oy),
                         static_cast<const int*>(nout), resolved_fwhm_radius(p),
                         static_cast<const double*>(bgs), static_cast<const double*>(bgr), ofs,
                         ofr);
    }
  } else {
    const int ns = run_detect_stamps(sci, smask, ws, hs, p, sx, sy, nullptr, nullptr, fs, s);
    const int nr = run_detect_stamps(ref, rmask, wr, hr, p, rx, ry, nullptr, nullptr, fr, s);
    ProfScope prof("select_stamps", s);
    if (nr > 0) {
      const long long tol2 = static_cast<long long>(p.match_radius) * p.match_radius;
      hipLaunchKernelGGL(k_match, dim3((nr + 255) / 256), dim3(256), 0, s,
                         static_cast<const int32_t*>(rx), static_cast<const int32_t*>(ry), nr,
                         static_cast<const int32_t*>(sx), static_cast<const int32_t*>(sy), ns,
                         tol2, best);
      hipLaunchKernelGGL(k_compact_selection, dim3(1), dim3(1024), 0, s,
                         static_cast<const int*>(best), nr, static_cast<const int32_t*>(rx),
                         static_cast<const int32_t*>(ry), static_cast<const double*>(fs),
                         static_cast<const double*>(fr), ox, oy, ofs, ofr, src, nout);
    }
  }
  hipLaunchKernelGGL(k_upper_median, dim3(2), dim3(1024), 0, s, static_cast<const double*>(ofs),
                     static_cast<const double*>(ofr), static_cast<const int*>(nout), med);
  DG_LAUNCH_CHECK();
  int n = 0;
  d2h(&n, nout, 1, s);
  d2h(medians, static_cast<const double*>(med), 2, s);
  DG_HIP(hipStreamSynchronize(s));
  *direction = medians[1] <= medians[0] ? 0 : 1;  // src/detect.cpp:243-246
  return n;
}

}  // namespace dg
