// Catalog summaries on device (reference src/catalog.cpp:262-697): the rows
// of build_catalog from the 8-connected labels of the score image.
//
// Each component's pixels are gathered as 64-bit keys (label : raster index)
// and sorted, which buckets them by label in raster order -- the order the
// reference's counting-sort scatter gives them on one thread, so the
// centroid sums round as there.  One wave per component then walks its
// pixels (centroid, peak, mask flags over footprint and boundary), forms the
// expected-pixel test and the aperture flux; a frame pass marks the dipoles;
// one thread per row runs the optional PSF-shape fit and bright-star
// distance.  Multiply-adds are fused exactly where the reference's compiler
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

// Labelled pixels: count, then keys (label << 32 | raster index).
__global__ __launch_bounds__(256) void k_count_labelled(const int32_t* __restrict__ labels,
                                                        size_t n, uint32_t* __restrict__ total) {
  uint32_t c = 0;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    c += labels[i] > 0 ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(total, c);
}

__global__ __launch_bounds__(256) void k_label_keys(const int32_t* __restrict__ labels, size_t n,
                                                    unsigned long long* __restrict__ keys,
                                                    uint32_t* __restrict__ fill) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const int32_t l = labels[i];
    if (l > 0) {
      const uint32_t pos = atomicAdd(fill, 1u);
      keys[pos] = (static_cast<unsigned long long>(static_cast<uint32_t>(l)) << 32) |
                  static_cast<uint32_t>(i);
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

// src/catalog.cpp:303-325
__device__ double aperture_flux(const float* __restrict__ d, int w, int h, double cx, double cy,
                                int radius) {
  const int x0 = static_cast<int>(llround(cx)), y0 = static_cast<int>(llround(cy));
  const double r2 = static_cast<double>(radius) * radius;
  double sum = 0.0;
  for (int dy = -radius; dy <= radius; ++dy) {
    const int yy = y0 + dy;
    if (yy < 0 || yy >= h) continue;
    for (int dx = -radius; dx <= radius; ++dx) {
      if (dx * dx + dy * dy > r2) continue;
      const int xx = x0 + dx;
      if (xx < 0 || xx >= w) continue;
      const float v = d[size_t(yy) * w + xx];
      if (isfinite(v)) sum += static_cast<double>(v);
    }
  }
  return sum;
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
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
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

// summarize_component (src/catalog.cpp:335-418), one wave per component.  The
// lanes load 64 member pixels at a time (raster order), test their 3x3
// boundary for mask flags and keep a per-lane peak (reduced at the end with
// the raster-order tie-break); the centroid sums run through the 64 in lane
// order with every lane doing the same uniform steps, so they round exactly as
// the sequential loop does (up to kSequentialMax pixels).
__global__ __launch_bounds__(256) void k_summarize(SumArgs a) {
  const int lane = threadIdx.x & 63;
  const int l = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) + 1;
  if (l > a.nc) return;
  const uint32_t b = a.start[l], e = a.start[l + 1];
  const int w = a.w, h = a.h;
  double sum = 0.0, sx = 0.0, sy = 0.0, best = -1.0, peak = 0.0;
  uint32_t best_idx = 0xffffffffu;
  uint32_t flags = 0;
  for (uint32_t c = b; c < e; c += 64) {
    const uint32_t k = c + lane;
    const bool valid = k < e;
    const uint32_t idx = valid ? static_cast<uint32_t>(a.keys[k]) : 0u;
    const int x = static_cast<int>(idx % uint32_t(w)), y = static_cast<int>(idx / uint32_t(w));
    const double v = valid ? static_cast<double>(a.score[idx]) : 0.0;
    if (valid && a.mask) {
      flags |= a.mask[idx];
      for (int dy = -1; dy <= 1; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= h) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          if (dx == 0 && dy == 0) continue;
          const int xx = x + dx;
          if (xx < 0 || xx >= w) continue;
          const size_t nb = size_t(yy) * w + xx;
          if (a.labels[nb] != l) flags |= a.mask[nb];
        }
      }
    }
    // peak: the largest |score|, the first in raster order among equals
    const double wt = fabs(v);
    if (valid && (wt > best || (wt == best && idx < best_idx))) {
      best = wt;
      best_idx = idx;
      peak = v;
    }
    if (e - b > kSequentialMax) {
      // a giant component: per-lane partial sums, reduced in a fixed order below
      if (valid) {
        sum += wt;
        sx = fma(wt, static_cast<double>(x), sx);
        sy = fma(wt, static_cast<double>(y), sy);
      }
      continue;
    }
    const int cnt = static_cast<int>(min(64u, e - c));
    if (cnt == 64) {
#pragma unroll 16
      for (int j = 0; j < 64; ++j) accumulate(wt, x, y, j, sum, sx, sy);
    } else {
      for (int j = 0; j < cnt; ++j) accumulate(wt, x, y, j, sum, sx, sy);
    }
  }
  if (e - b > kSequentialMax) {
    for (int off = 1; off < 64; off <<= 1) {
      sum += __shfl_xor(sum, off);
      sx += __shfl_xor(sx, off);
      sy += __shfl_xor(sy, off);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    flags |= __shfl_xor(flags, off);
    const double ob = __shfl_xor(best, off), op = __shfl_xor(peak, off);
    const uint32_t oi = __shfl_xor(best_idx, off);
    if (ob > best || (ob == best && oi < best_idx)) {
      best = ob;
      best_idx = oi;
      peak = op;
    }
  }
  if (lane != 0) return;
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
  r.flux = aperture_flux(a.diff, w, h, r.x, r.y, radius);
  r.quality = static_cast<uint8_t>((r.fwhm_consistent ? 0 : 1) | (r.mask_flags ? 4 : 0));
  a.out[l - 1] = r;
}

// flag_dipole_positives (src/catalog.cpp:437-458) with the opposite seed
// evaluated in place: finite, beyond its float threshold, unmasked when the
// mask excludes.
__global__ __launch_bounds__(256) void k_dipole_touch(const float* __restrict__ score,
                                                      const uint8_t* __restrict__ seed_mask,
                                                      int w, int h, float thr, int polarity,
                                                      const int32_t* __restrict__ labels,
                                                      uint8_t* __restrict__ touched) {
  const size_t n = size_t(w) * h;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float v = score[i];
    bool seed = isfinite(v) && (polarity > 0 ? v >= thr : v <= -thr);
    if (seed_mask) seed = seed && seed_mask[i] == 0;
    if (!seed) continue;
    const int x = static_cast<int>(i % size_t(w)), y = static_cast<int>(i / size_t(w));
    for (int dy = -1; dy <= 1; ++dy) {
      const int yy = y + dy;
      if (yy < 0 || yy >= h) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx;
        if (xx < 0 || xx >= w) continue;
        const int32_t l = labels[size_t(yy) * w + xx];
        if (l > 0) touched[l] = 1;
      }
    }
  }
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

int grid_for(size_t n) {
  return static_cast<int>(std::min<size_t>(4096, std::max<size_t>(1, (n + 255) / 256)));
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
  ar.reserve((given_labels ? 0 : Arena::round(sizeof(int32_t) * n) + label_scratch_bytes(w, h)) +
             Arena::round(sizeof(uint32_t) * 4));
  const int32_t* labels = given_labels;
  uint32_t* cnt = ar.take<uint32_t>(4);
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
  hipLaunchKernelGGL(k_count_labelled, dim3(grid_for(n)), dim3(256), 0, s,
                     static_cast<const int32_t*>(labels), n, cnt);
  DG_LAUNCH_CHECK();
  uint32_t m = 0;
  d2h(&m, cnt, 1, s);
  DG_HIP(hipStreamSynchronize(s));

  const uint32_t npad = pow2_at_least(m);
  ArenaLease lease2;
  Arena& br = *lease2;
  const bool own_rows = capacity < nc;  // rows beyond the caller's buffer go to scratch
  br.reserve(Arena::round(sizeof(unsigned long long) * npad) +
             Arena::round(sizeof(uint32_t) * (size_t(nc) + 2)) + Arena::round(size_t(nc) + 1) +
             (own_rows ? Arena::round(sizeof(dg_catalog_entry) * nc) : 0));
  unsigned long long* keys = br.take<unsigned long long>(npad);
  uint32_t* start = br.take<uint32_t>(size_t(nc) + 2);
  uint8_t* touched = br.take<uint8_t>(size_t(nc) + 1);
  dg_catalog_entry* rows = own_rows ? br.take<dg_catalog_entry>(nc) : out;

  hipLaunchKernelGGL(k_label_keys, dim3(grid_for(n)), dim3(256), 0, s,
                     static_cast<const int32_t*>(labels), n, keys, cnt + 1);
  if (npad > m)
    hipLaunchKernelGGL(k_pad_tail, dim3((npad - m + 255) / 256), dim3(256), 0, s, keys, m, npad);
  sort_keys_u64(keys, npad, s);
  hipLaunchKernelGGL(k_bucket_starts, dim3((m + 255) / 256), dim3(256), 0, s,
                     static_cast<const unsigned long long*>(keys), m, nc, start);
  SumArgs sa{keys, start, score, diff, mask, labels, w, h, nc, polarity, p, rows};
  hipLaunchKernelGGL(k_summarize, dim3((nc + 3) / 4), dim3(256), 0, s, sa);

  DG_HIP(hipMemsetAsync(touched, 0, size_t(nc) + 1, s));
  const double other_thr = return_negative ? p.threshold_sigma : p.threshold_sigma_dipole;
  hipLaunchKernelGGL(k_dipole_touch, dim3(grid_for(n)), dim3(256), 0, s, score, seed_mask, w, h,
                     static_cast<float>(other_thr), -polarity,
                     static_cast<const int32_t*>(labels), touched);

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
