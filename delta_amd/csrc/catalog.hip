// Catalog summaries on device (reference src/catalog.cpp:262-697): the rows
// of build_catalog from the 8-connected labels of the score image.
//
// Each component's pixels are gathered as 64-bit keys (label : raster index)
// and sorted, which buckets them by label in raster order -- the order the
// reference's counting-sort scatter gives them on one thread, so the
// centroid sums round as there.  One thread per labelled pixel marks the
// dipoles and ORs the mask flags over footprint and boundary; one wave per
// component then walks its pixels (centroid, peak), forms the expected-pixel
// test and the aperture flux (a workgroup per component above
// kSequentialMax pixels); one thread per row runs the optional PSF-shape fit
// and bright-star distance.  Multiply-adds are fused exactly where the reference's compiler
// fuses them; the one library call on the path, log(), may round its last
// bit differently from glibc's.
#include "../../include/deltagpu.h"
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr double kFwhmPerSigma = 2.354820045030949;
constexpr double kPi = 3.14159265358979323846;
// Components up to this many pixels sum their centroid in raster order, one
// step at a time (bit-identical to the reference on one thread).  Larger ones
// sum per lane and then across the wave in a fixed order: the reference's own
// order for them depends on its thread schedule (its atomic bucket scatter
// interleaves the threads' row ranges), and a 64-lane wave stepping through
// hundreds of thousands of pixels one at a time would dominate the catalog.
constexpr uint32_t kSequentialMax = 8192;
// A giant component's sums: pixel k of the component (raster order) goes into
// slot k mod kGiantSlots (one thread each of a 1024-thread workgroup), each
// wave's 64 slots are combined by a butterfly (l + (l ^ 1), + (l ^ 2), ...,
// + (l ^ 32)), and the sixteen wave totals are added in wave order.
constexpr int kGiantSlots = 1024;
// Frame passes over the labels work in chunks of 4096 pixels (256 threads x
// 4 loads of 4); the keys are compacted chunk by chunk in raster order.
constexpr int kChunk = 4096;
constexpr int kFrameBlocks = 2048;  // workgroups of the chunked frame passes

__device__ __forceinline__ int4 load4_i32(const int32_t* __restrict__ p, size_t i, size_t n,
                                          bool vec) {
  if (vec && i + 4 <= n) return *reinterpret_cast<const int4*>(p + i);
  int4 r;
  r.x = i < n ? p[i] : 0;
  r.y = i + 1 < n ? p[i + 1] : 0;
  r.z = i + 2 < n ? p[i + 2] : 0;
  r.w = i + 3 < n ? p[i + 3] : 0;
  return r;
}

// Exclusive scan of v over a 256-thread workgroup; *total gets the sum.
// lds holds 4 entries; every thread must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off);
    if (lane >= off) inc += o;
  }
  __syncthreads();  // lds free from any earlier call
  if (lane == 63) lds[wv] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t x = lds[w];
    if (w < wv) before += x;
    all += x;
  }
  *total = all;
  return before + inc - v;
}

// Labelled pixels per chunk (and their total).  Load u of thread t covers
// pixels 4 (256 u + t) .. + 3 of the chunk, so each load is coalesced; each
// workgroup takes chunks blockIdx.x, + gridDim.x, ... (tiny workgroups would
// leave the frame pass bound by workgroup dispatch).
__global__ __launch_bounds__(256) void k_chunk_count(const int32_t* __restrict__ labels, size_t n,
                                                     int vec, int nchunk,
                                                     uint32_t* __restrict__ ccount,
                                                     uint32_t* __restrict__ total) {
  __shared__ uint32_t red[2][4];
  uint32_t mine = 0;
  int par = 0;
  for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x, par ^= 1) {
    const size_t c0 = size_t(ch) * kChunk;
    int4 q[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) q[u] = load4_i32(labels, c0 + 4 * (256 * u + threadIdx.x), n, vec != 0);
    uint32_t c = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) c += (q[u].x > 0) + (q[u].y > 0) + (q[u].z > 0) + (q[u].w > 0);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    // alternate halves of red: a chunk's partials are never overwritten by
    // the next chunk's before thread 0 has read them
    if ((threadIdx.x & 63) == 0) red[par][threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t sum = red[par][0] + red[par][1] + red[par][2] + red[par][3];
      ccount[ch] = sum;
      mine += sum;
    }
  }
  if (threadIdx.x == 0 && mine) atomicAdd(total, mine);
}

// Exclusive scan of the chunk counts (one workgroup of 1024 threads, each
// taking a run of consecutive chunks).
__global__ __launch_bounds__(1024) void k_chunk_scan(const uint32_t* __restrict__ ccount,
                                                     int nchunk, uint32_t* __restrict__ coff) {
  __shared__ uint32_t wsum[16];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int per = (nchunk + 1023) / 1024;
  const int a0 = min(nchunk, t * per), a1 = min(nchunk, a0 + per);
  uint32_t mine = 0;
  for (int c = a0; c < a1; ++c) mine += ccount[c];
  uint32_t inc = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off);
    if (lane >= off) inc += o;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t run = inc - mine;
  for (int w = 0; w < wv; ++w) run += wsum[w];
  for (int c = a0; c < a1; ++c) {
    coff[c] = run;
    run += ccount[c];
  }
}

// Keys (label << 32 | raster index) of each non-empty chunk's labelled pixels
// at the chunk's offset, in raster order (chunks strided over the
// workgroups as in k_chunk_count).
__global__ __launch_bounds__(256) void k_label_keys(const int32_t* __restrict__ labels, size_t n,
                                                    int vec, int nchunk,
                                                    const uint32_t* __restrict__ ccount,
                                                    const uint32_t* __restrict__ coff,
                                                    unsigned long long* __restrict__ keys) {
  __shared__ uint32_t red[4];
  for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    if (ccount[ch] == 0) continue;  // uniform over the workgroup
    const size_t c0 = size_t(ch) * kChunk;
    uint32_t pos = coff[ch];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t i0 = c0 + 4 * (256 * u + threadIdx.x);
      const int4 q = load4_i32(labels, i0, n, vec != 0);
      const int32_t v[4] = {q.x, q.y, q.z, q.w};
      uint32_t c = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) c += v[j] > 0;
      uint32_t all;
      uint32_t at = pos + block_excl_scan(c, red, &all);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (v[j] > 0)
          keys[at++] = (static_cast<unsigned long long>(static_cast<uint32_t>(v[j])) << 32) |
                       static_cast<uint32_t>(i0 + j);
      pos += all;
    }
  }
}

__global__ __launch_bounds__(256) void k_pad_tail(unsigned long long* __restrict__ keys,
                                                  uint32_t m, uint32_t npad) {
  const uint32_t i = m + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npad) keys[i] = ~0ull;
}

// start[l] = first sorted position of label l; start[nc + 1] = m.
__global__ __launch_bounds__(256) void k_bucket_starts(const unsigned long long* __restrict__ keys,
                                                       uint32_t m, int nc,
                                                       uint32_t* __restrict__ start) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const uint32_t l = static_cast<uint32_t>(keys[i] >> 32);
  if (i == 0 || static_cast<uint32_t>(keys[i - 1] >> 32) != l) start[l] = i;
  if (i == 0) start[nc + 1] = m;
}

struct SumArgs {
  const unsigned long long* keys;
  const uint32_t* start;
  const float* score;
  const float* diff;
  const uint8_t* mask;
  const int32_t* labels;
  int w, h, nc, polarity;
  dg_catalog_params p;
  dg_catalog_entry* out;
  uint32_t* giant_cnt;   // components above kSequentialMax pixels ...
  int32_t* giant_list;   // ... and their labels, for k_summarize_giant
  const uint32_t* flags; // per label: OR of the mask flags (k_mask_flags), or null
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// src/catalog.cpp:303-325, by a whole wave: the lanes load the disc's pixels
// 64 at a time in raster order (dy, then dx) and every lane adds them in that
// order as uniform steps, so the double sum rounds as the sequential loop's
// (a skipped pixel adds +0.0, which leaves a sum that starts at +0.0
// unchanged).
__device__ double aperture_flux_wave(const float* __restrict__ d, int w, int h, double cx,
                                     double cy, int radius, int lane) {
  const int x0 = static_cast<int>(llround(cx)), y0 = static_cast<int>(llround(cy));
  const double r2 = static_cast<double>(radius) * radius;
  const int side = 2 * radius + 1, total = side * side;
  double sum = 0.0;
  for (int base = 0; base < total; base += 64) {
    const int q = base + lane;
    double v = 0.0;
    if (q < total) {
      const int dy = q / side - radius, dx = q % side - radius;
      const int yy = y0 + dy, xx = x0 + dx;
      if (dx * dx + dy * dy <= r2 && yy >= 0 && yy < h && xx >= 0 && xx < w) {
        const float f = d[size_t(yy) * w + xx];
        if (isfinite(f)) v = static_cast<double>(f);
      }
    }
    const int cnt = min(64, total - base);
    for (int j = 0; j < cnt; ++j) sum += readlane_f64(v, j);
  }
  return sum;
}

// One step of the reference's sequential centroid sums, with lane j's pixel
// (every lane performs the same uniform step).
__device__ __forceinline__ void accumulate(double wt, int x, int y, int j, double& sum,
                                           double& sx, double& sy) {
  const double wj = readlane_f64(wt, j);
  sum += wj;
  sx = fma(wj, static_cast<double>(__builtin_amdgcn_readlane(x, j)), sx);
  sy = fma(wj, static_cast<double>(__builtin_amdgcn_readlane(y, j)), sy);
}

// Mask flags of pixel idx of component l: its own and those of its 8
// neighbours outside the component.
__device__ __forceinline__ uint32_t pixel_flags(const SumArgs& a, uint32_t idx, int x, int y,
                                                int l) {
  uint32_t flags = a.mask[idx];
  for (int dy = -1; dy <= 1; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= a.h) continue;
    for (int dx = -1; dx <= 1; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const int xx = x + dx;
      if (xx < 0 || xx >= a.w) continue;
      const size_t nb = size_t(yy) * a.w + xx;
      if (a.labels[nb] != l) flags |= a.mask[nb];
    }
  }
  return flags;
}

// The peak: the largest |score|, the first in raster order among equals.
__device__ __forceinline__ void peak_merge(double ob, uint32_t oi, double op, double& best,
                                           uint32_t& best_idx, double& peak) {
  if (ob > best || (ob == best && oi < best_idx)) {
    best = ob;
    best_idx = oi;
    peak = op;
  }
}

// Mask flags of every component: one thread per labelled pixel ORs its own
// flags and those of its neighbours outside the component into the label's
// word (the OR does not depend on the order).
__global__ __launch_bounds__(256) void k_mask_flags(SumArgs a, uint32_t m,
                                                    uint32_t* __restrict__ flags) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const unsigned long long key = a.keys[k];
  const int l = static_cast<int>(key >> 32);
  const uint32_t idx = static_cast<uint32_t>(key);
  const int x = static_cast<int>(idx % uint32_t(a.w)), y = static_cast<int>(idx / uint32_t(a.w));
  const uint32_t f = pixel_flags(a, idx, x, y, l);
  if (f) atomicOr(flags + l, f);
}

// The row of component l (members [b, e) of the sorted keys) from its sums,
// called by a whole wave with the same arguments on every lane; lane 0 stores.
__device__ void finish_entry(const SumArgs& a, int l, uint32_t b, uint32_t e, double sum,
                             double sx, double sy, uint32_t best_idx, double peak,
                             uint32_t flags, int lane) {
  const int w = a.w, h = a.h;
  const int px = static_cast<int>(best_idx % uint32_t(w));
  const int py = static_cast<int>(best_idx / uint32_t(w));
  dg_catalog_entry r;
  r.psf_chi2 = 0.0;
  r.psf_amplitude = 0.0;
  r.psf_amplitude_err = 0.0;
  r.bright_star_dist = __builtin_inf();
  r.psf_chi2_dof = 0;
  r.is_dipole = 0;
  r.shape_consistent = 1;
  r.near_bright_star = 0;
  r.pad[0] = r.pad[1] = 0;
  r.n_pix = static_cast<int32_t>(e - b);
  r.peak_x = px;
  r.peak_y = py;
  const uint32_t first = static_cast<uint32_t>(a.keys[b]);
  r.x = sum > 0.0 ? sx / sum : static_cast<double>(first % uint32_t(w));
  r.y = sum > 0.0 ? sy / sum : static_cast<double>(first / uint32_t(w));
  r.peak_snr = peak;
  r.mask_flags = static_cast<uint8_t>(flags);
  const double sigma = a.p.expected_fwhm / kFwhmPerSigma;
  const double thr = a.polarity > 0 ? a.p.threshold_sigma : a.p.threshold_sigma_dipole;
  const double ratio = thr > 0.0 ? fabs(r.peak_snr) / thr : 0.0;
  const double lt = ratio > 0.0 ? log(ratio * ratio) : 0.0;
  r.expected_n_pix = kPi * sigma * sigma * (lt > 0.0 ? lt : 0.0);
  r.fwhm_ratio = r.expected_n_pix > 0.0 ? r.n_pix / r.expected_n_pix : __builtin_inf();
  r.fwhm_consistent =
      r.fwhm_ratio >= a.p.fwhm_tolerance_lo && r.fwhm_ratio <= a.p.fwhm_tolerance_hi;
  int radius = a.p.aperture_radius;
  if (radius <= 0) {
    const long long rr = llround(3.0 * sigma);
    radius = rr > 1 ? static_cast<int>(rr) : 1;
  }
  r.flux = aperture_flux_wave(a.diff, w, h, r.x, r.y, radius, lane);
  r.quality = static_cast<uint8_t>((r.fwhm_consistent ? 0 : 1) | (r.mask_flags ? 4 : 0));
  if (lane == 0) a.out[l - 1] = r;
}

// summarize_component (src/catalog.cpp:335-418), one wave per component of
// up to kSequentialMax pixels.  The lanes load 64 member pixels at a time
// (raster order) and keep a per-lane peak (reduced at the end with the
// raster-order tie-break; the mask flags come from k_mask_flags); the centroid
// sums run through the 64 in lane order with every lane doing the same
// uniform steps, so they round exactly as the sequential loop does.  Larger
// components are listed for k_summarize_giant.
__global__ __launch_bounds__(256) void k_summarize(SumArgs a) {
  const int lane = threadIdx.x & 63;
  const int l = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) + 1;
  if (l > a.nc) return;
  const uint32_t b = a.start[l], e = a.start[l + 1];
  if (e - b > kSequentialMax) {
    if (lane == 0) a.giant_list[atomicAdd(a.giant_cnt, 1u)] = l;
    return;
  }
  const int w = a.w;
  double sum = 0.0, sx = 0.0, sy = 0.0, best = -1.0, peak = 0.0;
  uint32_t best_idx = 0xffffffffu;
  for (uint32_t c = b; c < e; c += 64) {
    const uint32_t k = c + lane;
    const bool valid = k < e;
    const uint32_t idx = valid ? static_cast<uint32_t>(a.keys[k]) : 0u;
    const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
    const double v = valid ? static_cast<double>(a.score[idx]) : 0.0;
    const double wt = fabs(v);
    if (valid) peak_merge(wt, idx, v, best, best_idx, peak);
    const int cnt = static_cast<int>(min(64u, e - c));
    if (cnt == 64) {
#pragma unroll 16
      for (int j = 0; j < 64; ++j) accumulate(wt, x, y, j, sum, sx, sy);
    } else {
      for (int j = 0; j < cnt; ++j) accumulate(wt, x, y, j, sum, sx, sy);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off), op = __shfl_xor(peak, off);
    const uint32_t oi = __shfl_xor(best_idx, off);
    peak_merge(ob, oi, op, best, best_idx, peak);
  }
  finish_entry(a, l, b, e, sum, sx, sy, best_idx, peak, a.flags ? a.flags[l] : 0u, lane);
}

// summarize_component for the listed components above kSequentialMax pixels,
// one 1024-thread workgroup each: thread t takes members b + t + 1024 i
// (slot t of the fixed order), sixteen at a time with their loads in flight
// together; the peak reduces over the workgroup.
__global__ __launch_bounds__(kGiantSlots) void k_summarize_giant(SumArgs a) {
  __shared__ double wsum[3][kGiantSlots / 64], wbest[kGiantSlots / 64], wpeak[kGiantSlots / 64];
  __shared__ uint32_t widx[kGiantSlots / 64];
  if (blockIdx.x >= *a.giant_cnt) return;  // uniform over the workgroup
  const int l = a.giant_list[blockIdx.x];
  const uint32_t b = a.start[l], e = a.start[l + 1];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int w = a.w;
  double sum = 0.0, sx = 0.0, sy = 0.0, best = -1.0, peak = 0.0;
  uint32_t best_idx = 0xffffffffu;
  constexpr int G = 16;
  for (uint32_t c = b + t; c < e; c += G * kGiantSlots) {
    uint32_t idx[G];
    float v[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      const uint32_t k = c + u * kGiantSlots;
      idx[u] = k < e ? static_cast<uint32_t>(a.keys[k]) : 0u;
    }
#pragma unroll
    for (int u = 0; u < G; ++u) v[u] = c + u * kGiantSlots < e ? a.score[idx[u]] : 0.0f;
#pragma unroll
    for (int u = 0; u < G; ++u) {
      if (c + u * kGiantSlots >= e) break;
      const int x = static_cast<int>(idx[u] % uint32_t(w));
      const int y = static_cast<int>(idx[u] / uint32_t(w));
      const double vv = static_cast<double>(v[u]);
      const double wt = fabs(vv);
      sum += wt;
      sx = fma(wt, static_cast<double>(x), sx);
      sy = fma(wt, static_cast<double>(y), sy);
      peak_merge(wt, idx[u], vv, best, best_idx, peak);
    }
  }
  for (int off = 1; off < 64; off <<= 1) {
    sum += __shfl_xor(sum, off);
    sx += __shfl_xor(sx, off);
    sy += __shfl_xor(sy, off);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off), op = __shfl_xor(peak, off);
    const uint32_t oi = __shfl_xor(best_idx, off);
    peak_merge(ob, oi, op, best, best_idx, peak);
  }
  if (lane == 0) {
    wsum[0][wv] = sum;
    wsum[1][wv] = sx;
    wsum[2][wv] = sy;
    wbest[wv] = best;
    wpeak[wv] = peak;
    widx[wv] = best_idx;
  }
  __syncthreads();
  if (wv != 0) return;
  // every lane of wave 0 forms the same totals (the waves' in order)
  sum = wsum[0][0];
  sx = wsum[1][0];
  sy = wsum[2][0];
  for (int q = 1; q < kGiantSlots / 64; ++q) {
    sum += wsum[0][q];
    sx += wsum[1][q];
    sy += wsum[2][q];
    peak_merge(wbest[q], widx[q], wpeak[q], best, best_idx, peak);
  }
  finish_entry(a, l, b, e, sum, sx, sy, best_idx, peak, a.flags ? a.flags[l] : 0u, lane);
}

// flag_dipole_positives (src/catalog.cpp:437-458) with the opposite seed
// evaluated in place: finite, beyond its float threshold, unmasked when the
// mask excludes.  The reference marks the labels in each opposite seed's 3x3
// neighbourhood; the neighbourhood relation is symmetric, so a component is
// marked exactly when one of its pixels has an opposite seed in its own 3x3
// neighbourhood (itself included), which one thread per labelled pixel tests.
__global__ __launch_bounds__(256) void k_dipole_touch(const float* __restrict__ score,
                                                      const uint8_t* __restrict__ seed_mask,
                                                      int w, int h, float thr, int polarity,
                                                      const unsigned long long* __restrict__ keys,
                                                      uint32_t m, uint8_t* __restrict__ touched) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const unsigned long long key = keys[k];
  const uint32_t l = static_cast<uint32_t>(key >> 32), idx = static_cast<uint32_t>(key);
  const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
  bool hit = false;
  for (int dy = -1; dy <= 1; ++dy) {
    const int yy = y + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -1; dx <= 1; ++dx) {
      const int xx = x + dx;
      if (xx < 0 || xx >= w) continue;
      const size_t i = size_t(yy) * w + xx;
      const float v = score[i];
      bool seed = isfinite(v) && (polarity > 0 ? v >= thr : v <= -thr);
      if (seed_mask) seed = seed && seed_mask[i] == 0;
      hit = hit || seed;
    }
  }
  if (hit) touched[l] = 1;
}

// fit_psf_shape (src/catalog.cpp:465-530) and nearest_bright_star_dist
// (:534-545) per row, after the dipole flag.
struct EnrichArgs {
  const float* diff;
  const float* var;
  const uint8_t* mask;
  const float* psf;
  int ps;
  const double* bx;
  const double* by;
  int nb;
  int w, h, nc;
  int do_shape, do_bright, enrich;
  double bright_radius, chi2_tol;
  const uint8_t* touched;
  dg_catalog_entry* out;
};

__device__ void fit_shape(const EnrichArgs& a, dg_catalog_entry& e) {
  const int r = a.ps / 2;
  const int x0 = static_cast<int>(llround(e.x)), y0 = static_cast<int>(llround(e.y));
  double swdp = 0.0, swpp = 0.0;
  int used = 0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= a.h) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x0 + dx;
      if (xx < 0 || xx >= a.w) continue;
      const size_t i = size_t(yy) * a.w + xx;
      if (a.mask && a.mask[i] != 0) continue;
      const float v = a.var[i], d = a.diff[i];
      if (!(v > 0.0f) || !isfinite(v) || !isfinite(d)) continue;
      const double p = static_cast<double>(a.psf[(dy + r) * a.ps + (dx + r)]);
      const double wgt = 1.0 / static_cast<double>(v);
      swdp = fma(wgt * static_cast<double>(d), p, swdp);
      swpp = fma(wgt * p, p, swpp);
      ++used;
    }
  }
  if (used < 3 || !(swpp > 0.0)) return;
  const double amp = swdp / swpp;
  double chi2 = 0.0;
  for (int dy = -r; dy <= r; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= a.h) continue;
    for (int dx = -r; dx <= r; ++dx) {
      const int xx = x0 + dx;
      if (xx < 0 || xx >= a.w) continue;
      const size_t i = size_t(yy) * a.w + xx;
      if (a.mask && a.mask[i] != 0) continue;
      const float v = a.var[i], d = a.diff[i];
      if (!(v > 0.0f) || !isfinite(v) || !isfinite(d)) continue;
      const double p = static_cast<double>(a.psf[(dy + r) * a.ps + (dx + r)]);
      const double res = fma(-amp, p, static_cast<double>(d));
      chi2 += res * res / static_cast<double>(v);
    }
  }
  e.psf_amplitude = amp;
  e.psf_amplitude_err = sqrt(1.0 / swpp);
  e.psf_chi2 = chi2;
  e.psf_chi2_dof = used - 1;
  e.shape_consistent = e.psf_chi2_dof > 0 ? (e.psf_chi2 / e.psf_chi2_dof <= a.chi2_tol) : 1;
  if (!e.shape_consistent) e.quality |= 8;
}

__global__ __launch_bounds__(64) void k_finish_rows(EnrichArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nc) return;
  dg_catalog_entry e = a.out[i];
  if (a.touched[i + 1]) {
    e.is_dipole = 1;
    e.quality |= 2;
  }
  if (a.enrich) {
    if (a.do_shape) fit_shape(a, e);
    if (a.do_bright) {
      double best2 = __builtin_inf();
      for (int k = 0; k < a.nb; ++k) {
        const double dx = e.x - a.bx[k], dy = e.y - a.by[k];
        double yy = dy * dy;
        asm volatile("" : "+v"(yy));  // fma(dx, dx, dy * dy), as the reference's compiler
        const double d2 = fma(dx, dx, yy);
        if (d2 < best2) best2 = d2;
      }
      e.bright_star_dist = sqrt(best2);
      e.near_bright_star = e.bright_star_dist <= a.bright_radius;
      if (e.near_bright_star) e.quality |= 16;
    }
  }
  a.out[i] = e;
}

uint32_t pow2_at_least(uint32_t v) {
  uint32_t p = 2;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

// Rows into `out` (device, capacity rows); returns the component count.
// With given_labels the seeds' labels (and their count n_given) come from the
// caller -- the labelling of the same seeds, score >= threshold_sigma (or
// <= -threshold_sigma_dipole) outside the seed mask -- instead of a new one.
int run_build_catalog(int w, int h, const float* score, const float* diff, const uint8_t* mask,
                      const dg_catalog_params& p, int return_negative, const float* var,
                      const float* psf, int ps, const double* bx, const double* by, int nb,
                      dg_catalog_entry* out, int capacity, hipStream_t s,
                      const int32_t* given_labels, int n_given) {
  if (w <= 0 || h <= 0) return 0;
  const size_t n = size_t(w) * h;
  const uint8_t* seed_mask = p.exclude_bad_pixels ? mask : nullptr;
  const int polarity = return_negative ? -1 : 1;
  ArenaLease lease;
  Arena& ar = *lease;
  const size_t nchunk = (n + kChunk - 1) / kChunk;
  ar.reserve((given_labels ? 0 : Arena::round(sizeof(int32_t) * n) + label_scratch_bytes(w, h)) +
             Arena::round(sizeof(uint32_t) * 4) + 2 * Arena::round(sizeof(uint32_t) * nchunk));
  const int32_t* labels = given_labels;
  uint32_t* cnt = ar.take<uint32_t>(4);  // [0] labelled pixels, [2] giant components
  uint32_t* ccount = ar.take<uint32_t>(nchunk);
  uint32_t* coff = ar.take<uint32_t>(nchunk);
  DG_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t) * 4, s));
  ProfScope prof("catalog", s);
  int nc = n_given;
  if (!given_labels) {
    int32_t* own = ar.take<int32_t>(n);
    const size_t mark = ar.mark();
    nc = run_label(w, h, score, seed_mask,
                   return_negative ? p.threshold_sigma_dipole : p.threshold_sigma, polarity, own,
                   ar, s);
    ar.rewind(mark);
    labels = own;
  }
  if (nc == 0) return 0;
  const auto aligned = [](const void* q, size_t b) {
    return reinterpret_cast<uintptr_t>(q) % b == 0;
  };
  const int lvec = aligned(labels, 16) ? 1 : 0;
  const int cgrid = static_cast<int>(std::min<size_t>(nchunk, kFrameBlocks));
  hipLaunchKernelGGL(k_chunk_count, dim3(cgrid), dim3(256), 0, s,
                     static_cast<const int32_t*>(labels), n, lvec, static_cast<int>(nchunk), ccount,
                     cnt);
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, s,
                     static_cast<const uint32_t*>(ccount), static_cast<int>(nchunk), coff);
  DG_LAUNCH_CHECK();
  uint32_t m = 0;
  d2h(&m, cnt, 1, s);
  DG_HIP(hipStreamSynchronize(s));

  const uint32_t npad = pow2_at_least(m);
  // every component above kSequentialMax pixels has more than that many
  const uint32_t ngmax = m / (kSequentialMax + 1);
  ArenaLease lease2;
  Arena& br = *lease2;
  const bool own_rows = capacity < nc;  // rows beyond the caller's buffer go to scratch
  br.reserve(Arena::round(sizeof(unsigned long long) * npad) +
             Arena::round(sizeof(uint32_t) * (size_t(nc) + 2)) + Arena::round(size_t(nc) + 1) +
             Arena::round(sizeof(int32_t) * (size_t(ngmax) + 1)) +
             (mask ? Arena::round(sizeof(uint32_t) * (size_t(nc) + 1)) : 0) +
             (own_rows ? Arena::round(sizeof(dg_catalog_entry) * nc) : 0));
  unsigned long long* keys = br.take<unsigned long long>(npad);
  uint32_t* start = br.take<uint32_t>(size_t(nc) + 2);
  uint8_t* touched = br.take<uint8_t>(size_t(nc) + 1);
  uint32_t* flags = mask ? br.take<uint32_t>(size_t(nc) + 1) : nullptr;
  int32_t* giant_list = br.take<int32_t>(size_t(ngmax) + 1);
  dg_catalog_entry* rows = own_rows ? br.take<dg_catalog_entry>(nc) : out;

  hipLaunchKernelGGL(k_label_keys, dim3(cgrid), dim3(256), 0, s,
                     static_cast<const int32_t*>(labels), n, lvec, static_cast<int>(nchunk),
                     static_cast<const uint32_t*>(ccount), static_cast<const uint32_t*>(coff),
                     keys);
  // dipole marks from the labelled pixels (raster order here)
  DG_HIP(hipMemsetAsync(touched, 0, size_t(nc) + 1, s));
  const double other_thr = return_negative ? p.threshold_sigma : p.threshold_sigma_dipole;
  if (m > 0)
    hipLaunchKernelGGL(k_dipole_touch, dim3((m + 255) / 256), dim3(256), 0, s, score, seed_mask,
                       w, h, static_cast<float>(other_thr), -polarity,
                       static_cast<const unsigned long long*>(keys), m, touched);
  if (npad > m)
    hipLaunchKernelGGL(k_pad_tail, dim3((npad - m + 255) / 256), dim3(256), 0, s, keys, m, npad);
  sort_keys_u64(keys, npad, s);
  hipLaunchKernelGGL(k_bucket_starts, dim3((m + 255) / 256), dim3(256), 0, s,
                     static_cast<const unsigned long long*>(keys), m, nc, start);
  SumArgs sa{keys, start, score, diff, mask, labels, w, h, nc, polarity, p, rows, cnt + 2,
             giant_list, flags};
  if (mask) {
    DG_HIP(hipMemsetAsync(flags, 0, sizeof(uint32_t) * (size_t(nc) + 1), s));
    if (m > 0) hipLaunchKernelGGL(k_mask_flags, dim3((m + 255) / 256), dim3(256), 0, s, sa, m, flags);
  }
  hipLaunchKernelGGL(k_summarize, dim3((nc + 3) / 4), dim3(256), 0, s, sa);
  if (ngmax > 0)
    hipLaunchKernelGGL(k_summarize_giant, dim3(ngmax), dim3(kGiantSlots), 0, s, sa);


  const bool do_shape = p.fit_psf_shape && psf && var && ps > 0;
  const bool do_bright = bx && by && nb > 0;
  const bool enrich = nc <= std::max(0, p.max_shape_fit) && (do_shape || do_bright);
  EnrichArgs ea{diff, var, mask, psf, ps, bx, by, nb, w, h, nc,
                do_shape ? 1 : 0, do_bright ? 1 : 0, enrich ? 1 : 0,
                p.bright_star_radius > 0.0 ? p.bright_star_radius : 3.0 * p.expected_fwhm,
                p.shape_chi2_tolerance, touched, rows};
  hipLaunchKernelGGL(k_finish_rows, dim3((nc + 63) / 64), dim3(64), 0, s, ea);
  DG_LAUNCH_CHECK();
  if (own_rows && capacity > 0)
    DG_HIP(hipMemcpyAsync(out, rows, sizeof(dg_catalog_entry) * capacity,
                          hipMemcpyDeviceToDevice, s));
  DG_HIP(hipStreamSynchronize(s));
  return nc;
}

}  // namespace dg
