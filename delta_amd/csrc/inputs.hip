// Caller-side image arithmetic of the reference's Python layer, on device:
//
//   synth_variance   Var = max(data, 0) / gain + read_noise^2 in FP32
//                    (python/delta/_inputs.py:75-86; NaN stays NaN, as np.clip)
//   variance_rescale median(d^2 / v) / median(chi2_1) over the good pixels
//                    (v > 0, d and v finite, unmasked), in FP64
//                    (python/delta/pipeline.py:66-86, np.median: the mean of
//                    the two middle elements for an even count)
//   estimate_psf     median unit-sum stamp of the sources (pipeline.py:285-302)
//
// The median is an exact order statistic: z = d^2/v >= 0, so the bit pattern
// of the double is an order-preserving 64-bit key, selected by six
// most-significant-digit radix passes (11 bits, the last 9) that recompute z
// from the frame on every pass (9 bytes per pixel read, no key buffer).  Both
// middle ranks are carried through the same passes.
#include <algorithm>
#include <cstring>

#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kRsBins = 2048;

__global__ void k_synth_variance(size_t n, const float* __restrict__ data, float gain, float rn2,
                                 float* __restrict__ out) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    float v = data[i];
    v = v < 0.0f ? 0.0f : v;
    out[i] = v / gain + rn2;
  }
}

struct ZArgs {
  size_t n;
  const float* d;
  const float* v;
  const uint8_t* m;
};

__device__ __forceinline__ bool z_key(const ZArgs& a, size_t i, uint64_t& key) {
  const float v = a.v[i], d = a.d[i];
  if (!(v > 0.0f) || !isfinite(d) || !isfinite(v)) return false;
  if (a.m && a.m[i] != 0) return false;
  const double dd = d;
  const double z = (dd * dd) / double(v);
  key = static_cast<uint64_t>(__double_as_longlong(z));
  return true;
}

// Histograms of the digit at `shift` (width `bits`) of the keys whose bits
// above the digit equal prefix0 (h0) / prefix1 (h1); pass 0 also counts.
__global__ __launch_bounds__(256) void k_z_hist(ZArgs a, int shift, int bits, uint64_t hi_mask,
                                                uint64_t prefix0, uint64_t prefix1,
                                                unsigned* __restrict__ h0,
                                                unsigned* __restrict__ h1,
                                                unsigned long long* __restrict__ count) {
  __shared__ unsigned s0[kRsBins], s1[kRsBins];
  for (int i = threadIdx.x; i < kRsBins; i += blockDim.x) {
    s0[i] = 0;
    s1[i] = 0;
  }
  __syncthreads();
  const uint64_t dmask = (uint64_t(1) << bits) - 1;
  unsigned long long local = 0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n;
       i += size_t(gridDim.x) * blockDim.x) {
    uint64_t key;
    if (!z_key(a, i, key)) continue;
    ++local;
    const unsigned dgt = unsigned((key >> shift) & dmask);
    const uint64_t hi = key & hi_mask;
    if (hi == prefix0) atomicAdd(&s0[dgt], 1u);
    if (hi == prefix1) atomicAdd(&s1[dgt], 1u);
  }
  if (count) {
    for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(count, local);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kRsBins; i += blockDim.x) {
    if (s0[i]) atomicAdd(&h0[i], s0[i]);
    if (s1[i]) atomicAdd(&h1[i], s1[i]);
  }
}

__global__ void k_zero_u32(unsigned* p, int n, unsigned long long* c) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
  if (c && threadIdx.x == 0) *c = 0;
}

// numpy's pairwise sum of a[0..n) (sequential, one thread): blocks of at
// most 128 summed with eight interleaved accumulators, longer runs split in
// halves rounded down to a multiple of eight.  D bounds the split depth
// (n <= 128 * 2^D).
__device__ double np_pairwise_leaf(const double* a, int n) {
  if (n < 8) {
    double res = 0.0;
    for (int i = 0; i < n; ++i) res += a[i];
    return res;
  }
  double r[8];
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i = 8;
  for (; i < n - (n % 8); i += 8)
    for (int j = 0; j < 8; ++j) r[j] += a[i + j];
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += a[i];
  return res;
}

template <int D>
__device__ double np_pairwise_sum(const double* a, int n) {
  if (n <= 128) return np_pairwise_leaf(a, n);
  int n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise_sum<D - 1>(a, n2) + np_pairwise_sum<D - 1>(a + n2, n - n2);
}

template <>
__device__ double np_pairwise_sum<0>(const double* a, int n) {
  return np_pairwise_leaf(a, n);
}

constexpr int kPsfMaxPix = 1024;  // (2r+1)^2 for r <= 15

// One workgroup per source: the cutout minus its median, over its sum; ok[s]
// = 1 when the cutout lies inside the frame and the sum is positive.
__global__ __launch_bounds__(256) void k_psf_stamps(int w, int h, const float* __restrict__ img,
                                                    const int* __restrict__ xs,
                                                    const int* __restrict__ ys, int r,
                                                    double* __restrict__ norm,
                                                    int* __restrict__ ok) {
  __shared__ double v[kPsfMaxPix];
  __shared__ double med;
  __shared__ int any_nan;
  const int s = blockIdx.x, ks = 2 * r + 1, P = ks * ks;
  const int x = xs[s], y = ys[s];
  if (x - r < 0 || x + r >= w || y - r < 0 || y + r >= h) {
    if (threadIdx.x == 0) ok[s] = 0;
    return;
  }
  if (threadIdx.x == 0) any_nan = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const int cy = i / ks, cx = i - cy * ks;
    const double d = img[size_t(y - r + cy) * w + (x - r + cx)];
    v[i] = d;
    if (d != d) atomicOr(&any_nan, 1);
  }
  __syncthreads();
  // the middle element by rank (P is odd)
  const int mid = P / 2;
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const double vi = v[i];
    int rank = 0;
    for (int j = 0; j < P; ++j) {
      const double vj = v[j];
      rank += (vj < vi) || (vj == vi && j < i);
    }
    if (rank == mid) med = 0.0 + vi;
  }
  __syncthreads();
  const double m = any_nan ? __builtin_nan("") : med;
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) v[i] = v[i] - m;
  __syncthreads();
  __shared__ double total;
  if (threadIdx.x == 0) total = np_pairwise_sum<4>(v, P);
  __syncthreads();
  const double t = total;
  if (threadIdx.x == 0) ok[s] = t > 0.0 ? 1 : 0;
  if (!(t > 0.0)) return;
  double* o = norm + size_t(s) * P;
  for (int i = threadIdx.x; i < P; i += blockDim.x) o[i] = v[i] / t;
}

// One workgroup per pixel: the median over the accepted sources, in source
// order (np.median along the stack axis).
__global__ __launch_bounds__(256) void k_psf_median(int n, int P, const double* __restrict__ norm,
                                                    const int* __restrict__ ok,
                                                    double* __restrict__ med,
                                                    int* __restrict__ n_ok) {
  const int p = blockIdx.x;
  __shared__ int cnt, any_nan;
  __shared__ double lo, hi;
  if (threadIdx.x == 0) {
    int c = 0;
    for (int s = 0; s < n; ++s) c += ok[s];
    cnt = c;
    any_nan = 0;
    lo = 0.0;
    hi = 0.0;
  }
  __syncthreads();
  const int m = cnt;
  if (m == 0) {
    if (p == 0 && threadIdx.x == 0) *n_ok = 0;
    return;
  }
  const int r_lo = (m - 1) / 2, r_hi = m / 2;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (!ok[i]) continue;
    const double vi = norm[size_t(i) * P + p];
    if (vi != vi) atomicOr(&any_nan, 1);
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      if (!ok[j]) continue;
      const double vj = norm[size_t(j) * P + p];
      rank += (vj < vi) || (vj == vi && j < i);
    }
    if (rank == r_lo) lo = vi;
    if (rank == r_hi) hi = vi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v;
    if (any_nan) {
      v = __builtin_nan("");
    } else if (r_lo == r_hi) {
      v = 0.0 + lo;
    } else {
      v = ((0.0 + lo) + hi) / 2.0;
    }
    med[p] = v;
    if (p == 0) *n_ok = m;
  }
}

// clip at 0, divide by the pairwise sum, round to float
__global__ __launch_bounds__(256) void k_psf_finish(int P, double* __restrict__ med,
                                                    float* __restrict__ psf) {
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const double v = med[i];
    med[i] = v != v ? v : (v > 0.0 ? v : 0.0);
  }
  __syncthreads();
  __shared__ double total;
  if (threadIdx.x == 0) total = np_pairwise_sum<4>(med, P);
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += blockDim.x) psf[i] = static_cast<float>(med[i] / total);
}

// Inverse-variance weighted mean of the unmasked pixels with positive
// variance: per-block double partial sums, added in block order afterwards.
__global__ __launch_bounds__(256) void k_wmean_part(size_t n, const float* __restrict__ d,
                                                    const float* __restrict__ v,
                                                    const uint8_t* __restrict__ m,
                                                    double* __restrict__ part) {
  __shared__ double sw[256], sx[256];
  double ws = 0.0, xs = 0.0;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    if (m[i] != 0 || !(v[i] > 0.0f)) continue;
    const double w = 1.0 / double(v[i]);
    ws += w;
    xs = fma(w, double(d[i]), xs);
  }
  sw[threadIdx.x] = ws;
  sx[threadIdx.x] = xs;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sw[threadIdx.x] += sw[threadIdx.x + o];
      sx[threadIdx.x] += sx[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sw[0];
    part[2 * blockIdx.x + 1] = sx[0];
  }
}

}  // namespace

void run_synth_variance(size_t n, const float* data, double gain, double read_noise, float* out,
                        hipStream_t s) {
  if (!(gain > 0.0)) invalid("gain must be > 0");
  if (n == 0) return;
  const float g = static_cast<float>(gain);
  const float rn2 = static_cast<float>(read_noise * read_noise);
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, 8192));
  hipLaunchKernelGGL(k_synth_variance, dim3(grid), dim3(256), 0, s, n, data, g, rn2, out);
  DG_LAUNCH_CHECK();
}

size_t variance_rescale_scratch_bytes() {
  return 2 * Arena::round(sizeof(unsigned) * kRsBins) + Arena::round(sizeof(unsigned long long));
}

// Returns the factor (1.0 when no pixel is good).  Device inputs.
double run_variance_rescale_factor(size_t n, const float* diff, const float* var,
                                   const uint8_t* mask, Arena& ar, hipStream_t s) {
  unsigned* h = ar.take<unsigned>(2 * kRsBins);
  unsigned long long* d_cnt = ar.take<unsigned long long>(1);
  const ZArgs za{n, diff, var, mask};
  const unsigned grid = unsigned(std::max<size_t>(1, std::min<size_t>((n + 255) / 256,
                                                                      size_t(4 * device_cu_count()))));
  static const int kShift[6] = {53, 42, 31, 20, 9, 0};
  static const int kBits[6] = {11, 11, 11, 11, 11, 9};
  uint64_t prefix[2] = {0, 0};
  uint64_t rank[2] = {0, 0};
  unsigned long long count = 0;
  std::vector<unsigned> hh(2 * kRsBins);
  for (int p = 0; p < 6; ++p) {
    const int shift = kShift[p], bits = kBits[p];
    const uint64_t hi_mask = shift + bits >= 64 ? 0 : ~((uint64_t(1) << (shift + bits)) - 1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, s, h, 2 * kRsBins,
                       p == 0 ? d_cnt : nullptr);
    hipLaunchKernelGGL(k_z_hist, dim3(grid), dim3(256), 0, s, za, shift, bits, hi_mask, prefix[0],
                       prefix[1], h, h + kRsBins, p == 0 ? d_cnt : nullptr);
    DG_LAUNCH_CHECK();
    d2h(hh.data(), h, hh.size(), s);
    if (p == 0) d2h(&count, d_cnt, 1, s);
    DG_HIP(hipStreamSynchronize(s));
    if (p == 0) {
      if (count == 0) return 1.0;
      rank[0] = (count - 1) / 2;  // lower middle
      rank[1] = count / 2;        // upper middle (equal for an odd count)
    }
    for (int t = 0; t < 2; ++t) {
      const unsigned* ht = hh.data() + t * kRsBins;
      uint64_t acc = 0;
      int b = 0;
      for (; b < (1 << bits); ++b) {
        if (acc + ht[b] > rank[t]) break;
        acc += ht[b];
      }
      if (b == (1 << bits)) runtime("variance rescale: order statistic not found");
      rank[t] -= acc;
      prefix[t] |= uint64_t(b) << shift;
    }
  }
  double lo, hi;
  std::memcpy(&lo, &prefix[0], sizeof(double));
  std::memcpy(&hi, &prefix[1], sizeof(double));
  const double median = lo == hi ? lo : (lo + hi) / 2.0;
  return median / 0.4549364231195724;  // median of chi-square with one dof
}

size_t estimate_psf_scratch_bytes(int n, int radius) {
  const size_t P = size_t(2 * radius + 1) * (2 * radius + 1);
  return 2 * Arena::round(sizeof(int) * std::max(n, 1)) +
         Arena::round(sizeof(double) * P * std::max(n, 1)) + Arena::round(sizeof(double) * P) +
         Arena::round(sizeof(float) * P) + Arena::round(sizeof(int));
}

// Median unit-sum stamp of the sources at (xs, ys) (host arrays) in the device
// image; psf (host) is (2r+1)^2 float.  Raises when no source is usable.
void run_estimate_psf(int w, int h, const float* img, const int* xs, const int* ys, int n,
                      int radius, float* psf, Arena& ar, hipStream_t s) {
  if (radius < 0) invalid("estimate_psf: radius must be >= 0");
  const int ks = 2 * radius + 1, P = ks * ks;
  if (P > kPsfMaxPix) invalid("estimate_psf: radius above 15 is not supported");
  if (n <= 0) invalid("no usable sources for PSF estimation");
  int* d_xs = ar.take<int>(n);
  int* d_ys = ar.take<int>(n);
  double* norm = ar.take<double>(size_t(P) * n);
  int* ok = ar.take<int>(n);
  double* med = ar.take<double>(P);
  float* d_psf = ar.take<float>(P);
  int* d_nok = ar.take<int>(1);
  h2d(d_xs, xs, n, s);
  h2d(d_ys, ys, n, s);
  hipLaunchKernelGGL(k_psf_stamps, dim3(n), dim3(256), 0, s, w, h, img,
                     static_cast<const int*>(d_xs), static_cast<const int*>(d_ys), radius, norm,
                     ok);
  hipLaunchKernelGGL(k_psf_median, dim3(P), dim3(256), 0, s, n, P,
                     static_cast<const double*>(norm), static_cast<const int*>(ok), med, d_nok);
  DG_LAUNCH_CHECK();
  int n_ok = 0;
  d2h(&n_ok, d_nok, 1, s);
  DG_HIP(hipStreamSynchronize(s));
  if (n_ok == 0) invalid("no usable sources for PSF estimation");
  hipLaunchKernelGGL(k_psf_finish, dim3(1), dim3(256), 0, s, P, med, d_psf);
  DG_LAUNCH_CHECK();
  d2h(psf, d_psf, P, s);
  DG_HIP(hipStreamSynchronize(s));
}

constexpr int kWmeanBlocks = 1024;

size_t weighted_mean_scratch_bytes() { return Arena::round(sizeof(double) * 2 * kWmeanBlocks); }

// sum(d / v) / sum(1 / v) over pixels with mask == 0 and v > 0 (0 if none).
// Device inputs; a host result.
double run_weighted_mean(size_t n, const float* d, const float* v, const uint8_t* m, Arena& ar,
                         hipStream_t s) {
  if (n == 0) return 0.0;
  double* part = ar.take<double>(2 * kWmeanBlocks);
  const unsigned grid = unsigned(std::min<size_t>((n + 255) / 256, kWmeanBlocks));
  hipLaunchKernelGGL(k_wmean_part, dim3(grid), dim3(256), 0, s, n, d, v, m, part);
  DG_LAUNCH_CHECK();
  std::vector<double> h(2 * grid);
  d2h(h.data(), static_cast<const double*>(part), h.size(), s);
  DG_HIP(hipStreamSynchronize(s));
  double ws = 0.0, xs = 0.0;
  for (unsigned b = 0; b < grid; ++b) {
    ws += h[2 * b];
    xs += h[2 * b + 1];
  }
  return ws > 0.0 ? xs / ws : 0.0;
}

}  // namespace dg
