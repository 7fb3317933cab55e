// Exact order statistic on device: the k-th smallest of a strided, predicated
// float sample, by three most-significant-digit radix passes (11/11/10 bits)
// over an order-preserving integer key.  Used for the two median fills of the
// subtraction (reference src/subtract.cpp:20-51 and :919-945), whose sample is
// every 7th pixel that is finite and unmasked, and whose statistic is the
// element at rank size/2 (std::nth_element).  The first pass reads the frame
// and keeps the valid samples' keys in a compact buffer for the other two.
// No host round trip: the bin search runs in a one-wave kernel and the result
// stays in device memory.
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

__global__ void k_sel_zero(uint32_t* hist, SelectState* st) {
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) {
    st->prefix = 0;
    st->rank = 0;
    st->count = 0;
    st->done = 0;
  }
}

// Pass 0 (bits 31..21) over the strided sample: the histogram, and the keys of
// the valid samples kept for the later passes.  Workgroup b takes samples
// [b chunk, (b + 1) chunk) and packs their keys at the start of the same
// range of the key buffer (count in ncomp[b]), so no global counter is shared.
__global__ __launch_bounds__(256) void k_sel_gather(const float* __restrict__ data,
                                                    const uint8_t* __restrict__ mask,
                                                    size_t nsamp, int step, size_t chunk,
                                                    uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ ncomp) {
  __shared__ uint32_t sh[kBins];
  __shared__ uint32_t wtot[4];
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const size_t j0 = blockIdx.x * chunk, j1 = min(nsamp, j0 + chunk);
  uint32_t* out = keys + j0;
  uint32_t kept = 0;  // the workgroup's keys written so far (same in every thread)
  constexpr int kG = 8;  // rounds whose loads go out together
  for (size_t b0 = j0; b0 < j1; b0 += size_t(kG) * 256) {
    float v[kG];
    uint8_t mk[kG];
#pragma unroll
    for (int g = 0; g < kG; ++g) {
      const size_t j = b0 + size_t(g) * 256 + threadIdx.x;
      v[g] = 0.0f;
      mk[g] = 1;
      if (j < j1) {
        const size_t i = j * size_t(step);
        v[g] = data[i];
        mk[g] = mask ? mask[i] : uint8_t(0);
      }
    }
#pragma unroll
    for (int g = 0; g < kG; ++g) {
      if (b0 + size_t(g) * 256 >= j1) break;  // uniform over the workgroup
      const bool ok = mk[g] == 0 && isfinite(v[g]);
      const uint32_t key = f2key(v[g]);
      if (ok) atomicAdd(&sh[key >> 21], 1u);
      const unsigned long long m = __ballot(ok);
      if (lane == 0) wtot[wv] = uint32_t(__popcll(m));
      __syncthreads();
      uint32_t off = kept;
      for (int q = 0; q < wv; ++q) off += wtot[q];
      if (ok) out[off + __popcll(m & ((1ull << lane) - 1ull))] = key;
      kept += wtot[0] + wtot[1] + wtot[2] + wtot[3];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) ncomp[blockIdx.x] = kept;
  for (int i = threadIdx.x; i < kBins; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// Passes 1 (bits 20..10) and 2 (bits 9..0) over the kept keys that carry the
// prefix fixed so far (the same workgroup grid as the gather).
__global__ __launch_bounds__(256) void k_sel_hist(const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ ncomp,
                                                  size_t chunk, int pass,
                                                  const SelectState* __restrict__ st,
                                                  uint32_t* __restrict__ hist) {
  __shared__ uint32_t sh[kBins];
  for (int i = threadIdx.x; i < kBins; i += blockDim.x) sh[i] = 0;
  __syncthreads();
  const int shift = pass == 1 ? 10 : 0;
  const uint32_t bmask = pass == 2 ? 0x3ffu : 0x7ffu;
  const int hi_shift = pass == 1 ? 21 : 10;
  const uint32_t prefix = st->prefix;
  const uint32_t n = ncomp[blockIdx.x];
  const uint32_t* in = keys + blockIdx.x * chunk;
  if (st->done == 0) {
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
      const uint32_t key = in[j];
      if ((key >> hi_shift) != prefix) continue;
      atomicAdd(&sh[(key >> shift) & bmask], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBins; i += blockDim.x)
    if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// One wave: find the bin that holds the target rank, narrow the prefix, and
// clear the histogram the next pass accumulates into.
__global__ __launch_bounds__(64) void k_sel_scan(const uint32_t* __restrict__ hist, int pass,
                                                 SelectState* st, float* out,
                                                 uint32_t* __restrict__ next_hist) {
  const int lane = threadIdx.x;
  constexpr int per = kBins / 64;  // 32 bins per lane
  const int nb = pass == 2 ? 1024 : 2048;
  uint32_t local = 0;
  for (int b = 0; b < per; ++b) {
    const int bin = lane * per + b;
    if (bin < nb) local += hist[bin];
  }
  if (next_hist)
    for (int i = lane; i < kBins; i += 64) next_hist[i] = 0;
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

constexpr int kSelBlocks = 1024;  // most workgroups of a pass

size_t select_scratch_bytes(size_t n, int step) {
  const size_t nsamp = (n + step - 1) / step;
  return 2 * Arena::round(sizeof(uint32_t) * kBins) + Arena::round(sizeof(SelectState)) +
         Arena::round(sizeof(uint32_t) * kSelBlocks) +
         Arena::round(sizeof(uint32_t) * std::max<size_t>(nsamp, 1));
}

void select_strided_median(const float* data, const uint8_t* mask, size_t n, int step,
                           float* out, Arena& ar, hipStream_t s) {
  uint32_t* hist0 = ar.take<uint32_t>(kBins);
  uint32_t* hist1 = ar.take<uint32_t>(kBins);
  SelectState* st = ar.take<SelectState>(1);
  uint32_t* ncomp = ar.take<uint32_t>(kSelBlocks);
  const size_t nsamp = (n + step - 1) / step;
  uint32_t* keys = ar.take<uint32_t>(std::max<size_t>(nsamp, 1));
  const int blocks = static_cast<int>(
      std::max<size_t>(1, std::min<size_t>(kSelBlocks, (nsamp + 1023) / 1024)));
  const size_t chunk = (nsamp + blocks - 1) / blocks;
  hipLaunchKernelGGL(k_sel_zero, dim3(1), dim3(256), 0, s, hist0, st);
  hipLaunchKernelGGL(k_sel_gather, dim3(blocks), dim3(256), 0, s, data, mask, nsamp, step, chunk,
                     hist0, keys, ncomp);
  hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(64), 0, s, static_cast<const uint32_t*>(hist0), 0,
                     st, out, hist1);
  hipLaunchKernelGGL(k_sel_hist, dim3(blocks), dim3(256), 0, s,
                     static_cast<const uint32_t*>(keys), static_cast<const uint32_t*>(ncomp),
                     chunk, 1, st, hist1);
  hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(64), 0, s, static_cast<const uint32_t*>(hist1), 1,
                     st, out, hist0);
  hipLaunchKernelGGL(k_sel_hist, dim3(blocks), dim3(256), 0, s,
                     static_cast<const uint32_t*>(keys), static_cast<const uint32_t*>(ncomp),
                     chunk, 2, st, hist0);
  hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(64), 0, s, static_cast<const uint32_t*>(hist0), 2,
                     st, out, static_cast<uint32_t*>(nullptr));
  DG_LAUNCH_CHECK();
}

}  // namespace dg
