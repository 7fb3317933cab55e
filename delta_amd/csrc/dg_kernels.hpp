// Launchers shared between the HIP translation units of libdeltagpu.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dg_internal.hpp"

namespace dg {

constexpr uint8_t kMaskGood = 0, kMaskSaturated = 2, kMaskNonFinite = 8, kMaskEdge = 16;

// k-th (= size/2) smallest of the finite, unmasked samples data[i], i = step*j,
// written to *out (device).  0 when the sample is empty.
size_t select_scratch_bytes(size_t n, int step);
void select_strided_median(const float* data, const uint8_t* mask, size_t n, int step,
                           float* out, Arena& ar, hipStream_t s);

// Ascending in-place sort of npad (a power of two >= 2) 64-bit keys: a bitonic
// network, LDS-resident for every step below 2,048 keys.
void sort_keys_u64(unsigned long long* keys, uint32_t npad, hipStream_t s);

// Connected-component labels of the seed predicate (implemented in detect.hip).
size_t label_scratch_bytes(int w, int h);
int run_label(int w, int h, const float* score, const uint8_t* mask, double threshold,
              int polarity, int* labels, Arena& ar, hipStream_t s);
// The same on the device's static scratch array (no device allocation).
int run_label_noalloc(int w, int h, const float* score, const uint8_t* mask, double threshold,
                      int polarity, int* labels, hipStream_t s);

// Field lattice on device: either coarse (nf x ncy x ncx FP64 values plus the
// per-axis cell/weight tables) or exact (nf x h x w FP32 values per pixel).
struct DevFields {
  int exact;
  int nf, ncx, ncy;
  const double* lat;   // coarse values
  const int* cellx;    // w
  const float* wx;     // w
  const int* celly;    // h
  const float* wy;     // h
  const float* full;   // exact values
};

// Uploads the lattice axes and evaluates the lattice (or full fields) on device.
struct FieldsBuild {
  LatticeAxes axes;
  DevFields dev;
};
size_t fields_scratch_bytes(const Tps& tps, int nc, int w, int h);
// Field values at explicit (x, y) points (device npts x 2) -> out[f * npts + p], FP64.
// With a stage the knot, Z and theta uploads go through its pinned staging.
void eval_fields_points(const Tps& tps, const double* theta, int nf, const double* d_pts,
                        int npts, double* d_out, Arena& ar, hipStream_t s,
                        PinnedStage* stage = nullptr);
size_t eval_fields_points_bytes(const Tps& tps, int nf);

// Small uploads go through the thread's pinned staging in one copy; a caller
// may pass its own pending group, which is committed with them.  The staging
// is reused by the next call on this thread once the stream has drained.
FieldsBuild build_fields(const Tps& tps, const double* theta, int nc, int w, int h, Arena& ar,
                         hipStream_t s, UploadGroup* pending = nullptr);

// Bilinear cell ramp of field f at (x, y) (reference src/subtract.cpp:187-205).
__device__ __forceinline__ float field_at(const DevFields& F, int f, int x, int y, int w,
                                          int h) {
  if (F.exact) return F.full[(size_t(f) * h + y) * w + x];
  const int jx = F.cellx[x], jy = F.celly[y];
  const double* c0 = F.lat + (size_t(f) * F.ncy + jy) * F.ncx;
  const double* c1 = c0 + F.ncx;
  const float a00 = static_cast<float>(c0[jx]), a01 = static_cast<float>(c0[jx + 1]);
  const float a10 = static_cast<float>(c1[jx]), a11 = static_cast<float>(c1[jx + 1]);
  const float fy = F.wy[y];
  const float base = fmaf(a10 - a00, fy, a00);
  const float top = a01 - a00;
  const float slope = fmaf((a11 - a10) - top, fy, top);
  return fmaf(slope, F.wx[x], base);
}

}  // namespace dg
