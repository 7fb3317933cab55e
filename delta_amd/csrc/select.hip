// Exact order statistic on device: the k-th smallest of a strided, predicated
// float sample, by three most-significant-digit radix passes (11/11/10 bits)
// over an order-preserving integer key.  Used for the two median fills of the
// subtraction (reference src/subtract.cpp:20-51 and :919-945), whose sample is
// every 7th pixel that is finite and unmasked, and whose statistic is the
// element at rank size/2 (std::nth_element).  No host round trip: the bin
// search runs in a one-wave kernel and the result stays in device memory.
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kBins = 2048;

struct SelectState {
  uint32_t prefix;  // key bits fixed so far
  uint32_t rank;    // rank still to find inside the current prefix
  uint32_t count;   // valid samples (after pass 0)
  uint32_t done;    // 1 when the sample is empty
};

__device__ __forceinline__ uint32_t f2key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}

__global__ void k_sel_zero(uint32_t* hist, SelectState* st, int reset_state) {
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) hist[i] = 0;
  if (reset_state && threadIdx.x == 0) {
    st->prefix = 0;
    st->rank = 0;
    st->count = 0;
    st->done = 0;
  }
}

// pass 0: bits 31..21, pass 1: bits 20..10, pass 2: bits 9..0
__global__ __launch_bounds__(256) void k_sel_hist(const float* __restrict__ data,
                                                  const uint8_t* __restrict__ mask,
                                                  size_t nsamp, int step, int pass,
                                                  const SelectState* __restrict__ st,
                                                  uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[kBins];
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int shift = pass == 0 ? 21 : (pass == 1 ? 10 : 0);
  const uint32_t bmask = pass == 2 ? 0x3ffu : 0x7ffu;
  const int hi_shift = pass == 0 ? 32 : (pass == 1 ? 21 : 10);
  const uint32_t prefix = st->prefix;
  const bool skip = st->done != 0;
  if (!skip) {
    for (size_t j = blockIdx.x * size_t(blockDim.x) + threadIdx.x; j < nsamp;
         j += size_t(gridDim.x) * blockDim.x) {
      const size_t i = j * size_t(step);
      const float v = data[i];
      if (!isfinite(v)) continue;
      if (mask && mask[i] != 0) continue;
      const uint32_t key = f2key(v);
      if (hi_shift < 32 && (key >> hi_shift) != prefix) continue;
      atomicAdd(&sh[(key >> shift) & bmask], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBins; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// One wave: find the bin that holds the target rank, narrow the prefix.
__global__ __launch_bounds__(64) void k_sel_scan(const uint32_t* __restrict__ hist, int pass,
                                                 SelectState* st, float* out) {
  const int lane = threadIdx.x;
  constexpr int per = kBins / 64;  // 32 bins per lane
  const int nb = pass == 2 ? 1024 : 2048;
  uint32_t local = 0;
  for (int b = 0; b < per; ++b) {
    const int bin = lane * per + b;
    if (bin < nb) local += hist[bin];
  }
  // inclusive prefix over lanes
  uint32_t incl = local;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off, 64);
    if (lane >= off) incl += v;
  }
  __shared__ uint32_t s_total, s_rank, s_done, s_prefix;
  if (lane == 63) s_total = incl;
  __syncthreads();
  if (lane == 0) {
    if (pass == 0) {
      const uint32_t total = s_total;
      s_rank = total / 2;
      s_done = total == 0 ? 1u : 0u;
      s_prefix = 0;
      st->count = total;
      st->done = s_done;
      if (total == 0) *out = 0.0f;
    } else {
      s_rank = st->rank;
      s_done = st->done;
      s_prefix = st->prefix;
    }
  }
  __syncthreads();
  if (s_done) return;
  const uint32_t rank = s_rank;
  const uint32_t excl = incl - local;
  if (rank >= excl && rank < incl) {  // exactly one lane holds the target
    uint32_t cum = excl;
    for (int b = 0; b < per; ++b) {
      const int bin = lane * per + b;
      const uint32_t c = hist[bin];
      if (rank < cum + c) {
        const uint32_t bits = pass == 2 ? 10u : 11u;
        const uint32_t np = (pass == 0 ? 0u : (s_prefix << bits)) | uint32_t(bin);
        st->prefix = np;
        st->rank = rank - cum;
        if (pass == 2) *out = key2f(np);
        break;
      }
      cum += c;
    }
  }
}

}  // namespace

size_t select_scratch_bytes() {
  return Arena::round(sizeof(uint32_t) * kBins) + Arena::round(sizeof(SelectState));
}

void select_strided_median(const float* data, const uint8_t* mask, size_t n, int step,
                           float* out, Arena& ar, hipStream_t s) {
  uint32_t* hist = ar.take<uint32_t>(kBins);
  SelectState* st = ar.take<SelectState>(1);
  const size_t nsamp = (n + step - 1) / step;
  const int blocks = static_cast<int>(std::min<size_t>(1024, (nsamp + 255) / 256 + 1));
  for (int pass = 0; pass < 3; ++pass) {
    hipLaunchKernelGGL(k_sel_zero, dim3(1), dim3(256), 0, s, hist, st, pass == 0 ? 1 : 0);
    hipLaunchKernelGGL(k_sel_hist, dim3(blocks), dim3(256), 0, s, data, mask, nsamp, step, pass,
                       st, hist);
    hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(64), 0, s, hist, pass, st, out);
  }
  DG_LAUNCH_CHECK();
}

}  // namespace dg
