// Caller-side image arithmetic of the reference's Python layer, on device:
//
//   synth_variance   Var = max(data, 0) / gain + read_noise^2 in FP32
//                    (python/delta/_inputs.py:75-86; NaN stays NaN, as np.clip)
//   variance_rescale median(d^2 / v) / median(chi2_1) over the good pixels
//                    (v > 0, d and v finite, unmasked), in FP64
//                    (python/delta/pipeline.py:66-86, np.median: the mean of
//                    the two middle elements for an even count)
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

}  // namespace dg
