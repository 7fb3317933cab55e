// extern "C" entry points of libdeltagpu (declared in include/deltagpu.h).
// Host-pointer entry points stage their buffers through one scratch arena on a
// private stream and call the device-resident implementation.
#include <cstring>

#include "../../include/deltagpu.h"
#include "dg_internal.hpp"

namespace dg {
void run_subtract(int w, int h, const float* sci, const float* ref, const float* svar,
                  const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                  const double* knots, int n_knots, const double* theta, int n_theta,
                  double beta, int n_max, int radius, double saturation, float* diff,
                  float* var, uint8_t* mask, hipStream_t s);

FitOut run_fit(int w, int h, const float* sci, const float* ref, const float* svar,
               const float* rvar, const uint8_t* smask, const uint8_t* rmask,
               const double* knots, int n_knots, const int* stamp_x, const int* stamp_y, int ns,
               int sr, double beta, int n_max, int radius, const double* grid, int ng,
               double clip_sigma, int clip_iterations, int min_stamps, int mode, hipStream_t s);
void run_solve_gls_gcv(const double* knots, int n_knots, const double* pts, const double* target,
                       const double* weights, const double* bn, int n, int nc,
                       const double* grid, int ng, int mode, double* theta, double* scal,
                       double* curve);
}  // namespace dg

namespace {
void export_fit(const dg::FitOut& f, double* theta, double* curve, double* scal, int32_t* cnt,
                int32_t* sx, int32_t* sy, double* chi2, uint8_t* acc) {
  std::memcpy(theta, f.theta.data(), sizeof(double) * f.theta.size());
  std::memcpy(curve, f.curve.data(), sizeof(double) * f.curve.size());
  scal[0] = f.lambda;
  scal[1] = f.gcv;
  scal[2] = f.edf;
  scal[3] = f.rss;
  scal[4] = f.reduced_chi2;
  cnt[0] = f.n_pixels;
  cnt[1] = f.n_used;
  cnt[2] = f.n_total;
  cnt[3] = f.n_rejected;
  cnt[4] = f.passes;
  for (size_t i = 0; i < f.stamp_x.size(); ++i) {
    sx[i] = f.stamp_x[i];
    sy[i] = f.stamp_y[i];
    chi2[i] = f.stamp_chi2[i];
    acc[i] = f.stamp_accepted[i];
  }
}
void check_precision(int p) {
  if (p != DG_PREC_FP32 && p != DG_PREC_SPLIT_BF16 && p != DG_PREC_FP64)
    dg::invalid("precision must be 0 (fp32), 1 (split-bf16) or 2 (fp64)");
}
}  // namespace

using namespace dg;

extern "C" {

int dg_version(void) { return 100; }

const char* dg_last_error(void) { return dg::last_error(); }

int dg_default_radius(double beta, int n_max) { return GhBasis::default_radius(beta, n_max); }

int dg_subtract_device(int w, int h, const float* sci, const float* ref, const float* svar,
                       const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                       const double* knots, int n_knots, const double* theta, int n_theta,
                       double beta, int n_max, int radius, double saturation, float* diff,
                       float* var, uint8_t* mask, void* stream) {
  return guarded([&] {
    if (!sci || !ref || !diff || !knots || !theta) invalid("subtract: null required argument");
    run_subtract(w, h, sci, ref, svar, rvar, smask, rmask, knots, n_knots, theta, n_theta, beta,
                 n_max, radius, saturation, diff, var, mask, static_cast<hipStream_t>(stream));
  });
}

int dg_subtract(int w, int h, const float* sci, const float* ref, const float* svar,
                const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                const double* knots, int n_knots, const double* theta, int n_theta, double beta,
                int n_max, int radius, double saturation, float* diff, float* var,
                uint8_t* mask) {
  return guarded([&] {
    if (!sci || !ref || !diff || !knots || !theta) invalid("subtract: null required argument");
    if (w <= 0 || h <= 0) invalid("subtract: empty frame");
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    const size_t fb = Arena::round(n * sizeof(float)), ub = Arena::round(n);
    size_t need = 3 * fb;  // sci, ref, diff
    need += (svar ? fb : 0) + (rvar ? fb : 0) + (var ? fb : 0);
    need += (smask ? ub : 0) + (rmask ? ub : 0) + (mask ? ub : 0);
    ar.reserve(need);
    float* d_sci = ar.take<float>(n);
    float* d_ref = ar.take<float>(n);
    float* d_diff = ar.take<float>(n);
    float* d_sv = svar ? ar.take<float>(n) : nullptr;
    float* d_rv = rvar ? ar.take<float>(n) : nullptr;
    float* d_var = var ? ar.take<float>(n) : nullptr;
    uint8_t* d_sm = smask ? ar.take<uint8_t>(n) : nullptr;
    uint8_t* d_rm = rmask ? ar.take<uint8_t>(n) : nullptr;
    uint8_t* d_mask = mask ? ar.take<uint8_t>(n) : nullptr;
    hipStream_t s = st.get();
    h2d(d_sci, sci, n, s);
    h2d(d_ref, ref, n, s);
    if (svar) h2d(d_sv, svar, n, s);
    if (rvar) h2d(d_rv, rvar, n, s);
    if (smask) h2d(d_sm, smask, n, s);
    if (rmask) h2d(d_rm, rmask, n, s);
    run_subtract(w, h, d_sci, d_ref, d_sv, d_rv, d_sm, d_rm, knots, n_knots, theta, n_theta, beta,
                 n_max, radius, saturation, d_diff, d_var, d_mask, s);
    d2h(diff, d_diff, n, s);
    if (var) d2h(var, d_var, n, s);
    if (mask) d2h(mask, d_mask, n, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_fit_kernel_device(int w, int h, const float* sci, const float* ref, const float* svar,
                         const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                         const double* knots, int n_knots, const int32_t* sx, const int32_t* sy,
                         int ns, int sr, double beta, int n_max, int radius, const double* grid,
                         int ng, double clip_sigma, int clip_iterations, int min_stamps,
                         int precision, double* theta, double* curve, double* scal, int32_t* cnt,
                         int32_t* sxo, int32_t* syo, double* chi2, uint8_t* acc, void* stream) {
  return guarded([&] {
    if (!sci || !ref || !knots || (ns > 0 && (!sx || !sy)) || !theta || !curve || !scal || !cnt)
      invalid("fit_kernel: null required argument");
    if (w <= 0 || h <= 0) invalid("fit_kernel: empty frame");
    check_precision(precision);
    const FitOut f = run_fit(w, h, sci, ref, svar, rvar, smask, rmask, knots, n_knots, sx, sy, ns,
                             sr, beta, n_max, radius, grid, ng, clip_sigma, clip_iterations,
                             min_stamps, precision, static_cast<hipStream_t>(stream));
    export_fit(f, theta, curve, scal, cnt, sxo, syo, chi2, acc);
  });
}

int dg_fit_kernel(int w, int h, const float* sci, const float* ref, const float* svar,
                  const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                  const double* knots, int n_knots, const int32_t* sx, const int32_t* sy, int ns,
                  int sr, double beta, int n_max, int radius, const double* grid, int ng,
                  double clip_sigma, int clip_iterations, int min_stamps, int precision,
                  double* theta, double* curve, double* scal, int32_t* cnt, int32_t* sxo,
                  int32_t* syo, double* chi2, uint8_t* acc) {
  return guarded([&] {
    if (!sci || !ref || !knots || (ns > 0 && (!sx || !sy)) || !theta || !curve || !scal || !cnt)
      invalid("fit_kernel: null required argument");
    if (w <= 0 || h <= 0) invalid("fit_kernel: empty frame");
    check_precision(precision);
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    const size_t fb = Arena::round(n * sizeof(float)), ub = Arena::round(n);
    ar.reserve(2 * fb + (svar ? fb : 0) + (rvar ? fb : 0) + (smask ? ub : 0) + (rmask ? ub : 0));
    hipStream_t s = st.get();
    float* d_sci = ar.take<float>(n);
    float* d_ref = ar.take<float>(n);
    float* d_sv = svar ? ar.take<float>(n) : nullptr;
    float* d_rv = rvar ? ar.take<float>(n) : nullptr;
    uint8_t* d_sm = smask ? ar.take<uint8_t>(n) : nullptr;
    uint8_t* d_rm = rmask ? ar.take<uint8_t>(n) : nullptr;
    h2d(d_sci, sci, n, s);
    h2d(d_ref, ref, n, s);
    if (svar) h2d(d_sv, svar, n, s);
    if (rvar) h2d(d_rv, rvar, n, s);
    if (smask) h2d(d_sm, smask, n, s);
    if (rmask) h2d(d_rm, rmask, n, s);
    const FitOut f = run_fit(w, h, d_sci, d_ref, d_sv, d_rv, d_sm, d_rm, knots, n_knots, sx, sy,
                             ns, sr, beta, n_max, radius, grid, ng, clip_sigma, clip_iterations,
                             min_stamps, precision, s);
    export_fit(f, theta, curve, scal, cnt, sxo, syo, chi2, acc);
  });
}

int dg_solve_gls_gcv(const double* knots, int n_knots, const double* points,
                     const double* target, const double* weights, const double* bn, int n,
                     int nc, const double* grid, int ng, int precision, double* theta,
                     double* scal, double* curve) {
  return guarded([&] {
    if (!knots || !points || !target || !weights || (nc > 0 && !bn) || !grid)
      invalid("solve_gls_gcv: null required argument");
    check_precision(precision);
    run_solve_gls_gcv(knots, n_knots, points, target, weights, bn, n, nc, grid, ng, precision,
                      theta, scal, curve);
  });
}

}  // extern "C"
