// extern "C" entry points of libdeltagpu (declared in include/deltagpu.h).
// Host-pointer entry points stage their buffers through one scratch arena on a
// private stream and call the device-resident implementation.
#include <algorithm>
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
               double clip_sigma, int clip_iterations, int min_stamps, int cv_folds, int mode,
               hipStream_t s);
void run_solve_gls_gcv(const double* knots, int n_knots, const double* pts, const double* target,
                       const double* weights, const double* bn, int n, int nc,
                       const double* grid, int ng, int mode, double* theta, double* scal,
                       double* curve);
void run_solve_gls_cv(const double* knots, int n_knots, const double* pts, const double* target,
                      const double* weights, const double* bn, int n, int nc,
                      const double* grid, int ng, const int* group, int n_groups, int mode,
                      double* theta, double* scal, double* curve);
void run_tps_design(const double* knots, int n_knots, const double* pts, int n,
                    const double* coef, double* out);
void run_tps_penalty(const double* knots, int n_knots, double* out);
void run_tps_fit(const double* knots, int n_knots, const double* pts, const double* values, int n,
                 double lam, double* theta);
void run_basis_convolve(int w, int h, const float* img, double beta, int n_max, int radius,
                        float* out, hipStream_t s);
void run_decorrelate_block(int n, const float* img, const float* kernel, int ks, double vs,
                           double vr, bool shift, float* out);
void run_synth_variance(size_t n, const float* data, double gain, double read_noise, float* out,
                        hipStream_t s);
size_t variance_rescale_scratch_bytes();
double run_variance_rescale_factor(size_t n, const float* diff, const float* var,
                                   const uint8_t* mask, Arena& ar, hipStream_t s);
size_t estimate_psf_scratch_bytes(int n, int radius);
void run_estimate_psf(int w, int h, const float* img, const int* xs, const int* ys, int n,
                      int radius, float* psf, Arena& ar, hipStream_t s);
size_t weighted_mean_scratch_bytes();
double run_weighted_mean(size_t n, const float* d, const float* v, const uint8_t* m, Arena& ar,
                         hipStream_t s);
void run_decorrelate(int w, int h, const float* diff, const double* knots, int n_knots,
                     const double* theta, int n_theta, double beta, int n_max, const float* var_s,
                     const float* var_r, int block, int radius, int kernel_cell_blocks,
                     float* out, float* out_var, hipStream_t s);
void run_whiten_score_psf(const float* psf, int ps, const double* knots, int n_knots,
                          const double* theta, int n_theta, double beta, int n_max,
                          const float* var_s, const float* var_r, int w, int h, int block,
                          int radius, float* out, hipStream_t s);
void run_matched_filter(int w, int h, const float* img, const float* psf, int ps, const float* var,
                        float* out, hipStream_t s);
size_t label_scratch_bytes(int w, int h);
int run_label(int w, int h, const float* score, const uint8_t* mask, double threshold,
              int polarity, int* labels, Arena& ar, hipStream_t s);
int run_label_noalloc(int w, int h, const float* score, const uint8_t* mask, double threshold,
                      int polarity, int* labels, hipStream_t s);
void run_estimate_background(const float* img, const uint8_t* mask, int w, int h, double* median,
                             double* sigma, hipStream_t s);
int run_detect_stamps(const float* img, const uint8_t* mask, int w, int h,
                      const dg_detect_params& p, int32_t* ox, int32_t* oy, double* flux,
                      double* snr, double* fwhm, hipStream_t s);
int run_select_stamps(int ws, int hs, const float* sci, const uint8_t* smask, int wr, int hr,
                      const float* ref, const uint8_t* rmask, const int32_t* cat_x,
                      const int32_t* cat_y, int n_cat, const dg_detect_params& p, int32_t* ox,
                      int32_t* oy, double* ofs, double* ofr, double* medians, int* direction,
                      hipStream_t s);
int run_build_catalog(int w, int h, const float* score, const float* diff, const uint8_t* mask,
                      const dg_catalog_params& p, int return_negative, const float* var,
                      const float* psf, int ps, const double* bx, const double* by, int nb,
                      dg_catalog_entry* out, int capacity, hipStream_t s,
                      const int32_t* given_labels = nullptr, int n_given = 0);
void run_photometric_scale(const double* knots, int n_knots, const double* theta, int n_theta,
                           const double* sums, int nc, int w, int h, float* out, hipStream_t s);
void run_photometric_scale_at(const double* knots, int n_knots, const double* theta,
                              int n_theta, const double* sums, int nc, const double* points,
                              int m, double* out);
void kernel_model_check(int w, int h, int n_max, int radius);
size_t kernel_model_scratch_bytes(int w, int h, int radius);
void run_kernel_model(int mode, int w, int h, const float* sci, const float* ref,
                      const float* weight, const float* upstream, const double* params,
                      int n_max, int radius, int want_grad, float* diff, double* out, Arena& ar,
                      hipStream_t s);
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
  scal[5] = f.cv_exact_design_bytes;
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

int dg_gauss_hermite_basis1d(double beta, int n_max, int radius, double* out) {
  return guarded([&] {
    if (!out) invalid("gauss_hermite_basis1d: null output");
    const GhBasis b(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
    std::copy(b.g.begin(), b.g.end(), out);
  });
}

int dg_component_sums(double beta, int n_max, int radius, double* out) {
  return guarded([&] {
    if (!out) invalid("component_sums: null output");
    const GhBasis b(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
    const std::vector<double> cs = b.component_sums();
    std::copy(cs.begin(), cs.end(), out);
  });
}

int dg_gauss_hermite_kernels(double beta, int n_max, int radius, int32_t* orders,
                             float* kernels) {
  return guarded([&] {
    if (!orders || !kernels) invalid("gauss_hermite_kernels: null output");
    const GhBasis b(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
    for (int c = 0; c < b.nc; ++c) {
      orders[2 * c] = b.order[c].first;
      orders[2 * c + 1] = b.order[c].second;
    }
    const std::vector<float> fp = b.footprints_f32();
    std::copy(fp.begin(), fp.end(), kernels);
  });
}

int dg_basis_convolve_device(int w, int h, const float* image, double beta, int n_max,
                             int radius, float* out, void* stream) {
  return guarded([&] {
    if (!image || !out) invalid("basis_convolve: null required argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    run_basis_convolve(w, h, image, beta, n_max, radius, out, s);
  });
}

int dg_basis_convolve(int w, int h, const float* image, double beta, int n_max, int radius,
                      float* out) {
  return guarded([&] {
    if (!image || !out) invalid("basis_convolve: null required argument");
    if (w <= 0 || h <= 0) invalid("basis_convolve: empty image");
    const GhBasis b(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(Arena::round(n * sizeof(float)) + Arena::round(n * b.nc * sizeof(float)));
    float* d_img = ar.take<float>(n);
    float* d_out = ar.take<float>(n * b.nc);
    hipStream_t s = st.get();
    h2d(d_img, image, n, s);
    run_basis_convolve(w, h, d_img, beta, n_max, radius, d_out, s);
    d2h(out, d_out, n * b.nc, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_tps_design(const double* knots, int n_knots, const double* points, int n, double* out) {
  return guarded([&] {
    if (!knots || (n > 0 && (!points || !out))) invalid("tps_design: null required argument");
    run_tps_design(knots, n_knots, points, n, nullptr, out);
  });
}

int dg_tps_evaluate(const double* knots, int n_knots, const double* points, int n,
                    const double* coeffs, int n_coeffs, double* out) {
  return guarded([&] {
    if (!knots || !coeffs || (n > 0 && (!points || !out)))
      invalid("tps_evaluate: null required argument");
    if (n_coeffs != n_knots) runtime("tps_evaluate: coeffs size must equal the number of knots");
    run_tps_design(knots, n_knots, points, n, coeffs, out);
  });
}

int dg_tps_penalty(const double* knots, int n_knots, double* out) {
  return guarded([&] {
    if (!knots || !out) invalid("tps_penalty: null required argument");
    run_tps_penalty(knots, n_knots, out);
  });
}

int dg_tps_fit(const double* knots, int n_knots, const double* points, const double* values,
               int n, double lam, double* theta) {
  return guarded([&] {
    if (!knots || !points || !values || !theta) invalid("tps_fit: null required argument");
    run_tps_fit(knots, n_knots, points, values, n, lam, theta);
  });
}

int dg_decorrelate_block(int n, const float* image, const float* kernel, int ksize,
                         double var_science, double var_reference, float* out) {
  return guarded([&] {
    if (!image || !kernel || !out) invalid("decorrelate_block: null required argument");
    run_decorrelate_block(n, image, kernel, ksize, var_science, var_reference, false, out);
  });
}

int dg_decorrelation_kernel(const float* kernel, int ksize, double var_science,
                            double var_reference, int n, float* out) {
  return guarded([&] {
    if (!kernel || !out) invalid("decorrelation_kernel: null required argument");
    run_decorrelate_block(n, nullptr, kernel, ksize, var_science, var_reference, true, out);
  });
}

int dg_synth_variance_device(int w, int h, const float* data, double gain, double read_noise,
                             float* out, void* stream) {
  return guarded([&] {
    if (!data || !out) invalid("synth_variance: null required argument");
    if (w < 0 || h < 0) invalid("synth_variance: negative size");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    run_synth_variance(size_t(w) * h, data, gain, read_noise, out, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_synth_variance(int w, int h, const float* data, double gain, double read_noise,
                      float* out) {
  return guarded([&] {
    if (!data || !out) invalid("synth_variance: null required argument");
    if (w < 0 || h < 0) invalid("synth_variance: negative size");
    if (!(gain > 0.0)) invalid("gain must be > 0");
    const size_t n = size_t(w) * h;
    if (n == 0) return;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(2 * Arena::round(n * sizeof(float)));
    float* d_in = ar.take<float>(n);
    float* d_out = ar.take<float>(n);
    hipStream_t s = st.get();
    h2d(d_in, data, n, s);
    run_synth_variance(n, d_in, gain, read_noise, d_out, s);
    d2h(out, d_out, n, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_variance_rescale_factor_device(int w, int h, const float* difference,
                                      const float* variance, const uint8_t* mask, double* factor,
                                      void* stream) {
  return guarded([&] {
    if (!difference || !variance || !factor) invalid("variance_rescale: null required argument");
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(variance_rescale_scratch_bytes());
    *factor = run_variance_rescale_factor(size_t(w) * h, difference, variance, mask, ar,
                                          static_cast<hipStream_t>(stream));
  });
}

int dg_variance_rescale_factor(int w, int h, const float* difference, const float* variance,
                               const uint8_t* mask, double* factor) {
  return guarded([&] {
    if (!difference || !variance || !factor) invalid("variance_rescale: null required argument");
    const size_t n = size_t(w) * h;
    *factor = 1.0;
    if (n == 0) return;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(2 * Arena::round(n * sizeof(float)) + Arena::round(n) +
               variance_rescale_scratch_bytes());
    float* d_d = ar.take<float>(n);
    float* d_v = ar.take<float>(n);
    uint8_t* d_m = mask ? ar.take<uint8_t>(n) : nullptr;
    hipStream_t s = st.get();
    h2d(d_d, difference, n, s);
    h2d(d_v, variance, n, s);
    if (mask) h2d(d_m, mask, n, s);
    *factor = run_variance_rescale_factor(n, d_d, d_v, d_m, ar, s);
  });
}

int dg_estimate_psf_device(int w, int h, const float* image, const int32_t* xs,
                           const int32_t* ys, int n, int radius, float* psf, void* stream) {
  return guarded([&] {
    if (!image || !psf || (n > 0 && (!xs || !ys))) invalid("estimate_psf: null required argument");
    if (radius < 0 || radius > 15) invalid("estimate_psf: radius must be in [0, 15]");
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(estimate_psf_scratch_bytes(n, radius));
    run_estimate_psf(w, h, image, xs, ys, n, radius, psf, ar, static_cast<hipStream_t>(stream));
  });
}

int dg_estimate_psf(int w, int h, const float* image, const int32_t* xs, const int32_t* ys,
                    int n, int radius, float* psf) {
  return guarded([&] {
    if (!image || !psf || (n > 0 && (!xs || !ys))) invalid("estimate_psf: null required argument");
    if (radius < 0 || radius > 15) invalid("estimate_psf: radius must be in [0, 15]");
    const size_t np = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(Arena::round(np * sizeof(float)) + estimate_psf_scratch_bytes(n, radius));
    float* d_img = ar.take<float>(std::max<size_t>(np, 1));
    hipStream_t s = st.get();
    h2d(d_img, image, np, s);
    run_estimate_psf(w, h, d_img, xs, ys, n, radius, psf, ar, s);
  });
}

int dg_weighted_mean(int w, int h, const float* data, const float* variance,
                     const uint8_t* mask, double* out) {
  return guarded([&] {
    if (!data || !variance || !mask || !out) invalid("weighted_mean: null required argument");
    const size_t n = size_t(w) * h;
    *out = 0.0;
    if (n == 0) return;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(2 * Arena::round(n * sizeof(float)) + Arena::round(n) +
               weighted_mean_scratch_bytes());
    float* d_d = ar.take<float>(n);
    float* d_v = ar.take<float>(n);
    uint8_t* d_m = ar.take<uint8_t>(n);
    hipStream_t s = st.get();
    h2d(d_d, data, n, s);
    h2d(d_v, variance, n, s);
    h2d(d_m, mask, n, s);
    *out = run_weighted_mean(n, d_d, d_v, d_m, ar, s);
  });
}

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
                         int cv_folds, int precision, double* theta, double* curve, double* scal,
                         int32_t* cnt, int32_t* sxo, int32_t* syo, double* chi2, uint8_t* acc,
                         void* stream) {
  return guarded([&] {
    if (!sci || !ref || !knots || (ns > 0 && (!sx || !sy)) || !theta || !curve || !scal || !cnt)
      invalid("fit_kernel: null required argument");
    if (w <= 0 || h <= 0) invalid("fit_kernel: empty frame");
    check_precision(precision);
    const FitOut f = run_fit(w, h, sci, ref, svar, rvar, smask, rmask, knots, n_knots, sx, sy, ns,
                             sr, beta, n_max, radius, grid, ng, clip_sigma, clip_iterations,
                             min_stamps, cv_folds, precision, static_cast<hipStream_t>(stream));
    export_fit(f, theta, curve, scal, cnt, sxo, syo, chi2, acc);
  });
}

int dg_fit_kernel(int w, int h, const float* sci, const float* ref, const float* svar,
                  const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                  const double* knots, int n_knots, const int32_t* sx, const int32_t* sy, int ns,
                  int sr, double beta, int n_max, int radius, const double* grid, int ng,
                  double clip_sigma, int clip_iterations, int min_stamps, int cv_folds,
                  int precision, double* theta, double* curve, double* scal, int32_t* cnt,
                  int32_t* sxo, int32_t* syo, double* chi2, uint8_t* acc) {
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
                             min_stamps, cv_folds, precision, s);
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

int dg_solve_gls_cv(const double* knots, int n_knots, const double* points, const double* target,
                    const double* weights, const double* bn, int n, int nc,
                    const double* lambda_grid, int n_lambda, const int32_t* group, int n_groups,
                    int precision, double* theta, double* scalars, double* curve) {
  return guarded([&] {
    if (!knots || !points || !target || !weights || (nc > 0 && !bn) || !lambda_grid ||
        !group || !theta || !scalars || !curve)
      invalid("solve_gls_cv: null required argument");
    check_precision(precision);
    run_solve_gls_cv(knots, n_knots, points, target, weights, bn, n, nc, lambda_grid, n_lambda,
                     group, n_groups, precision, theta, scalars, curve);
  });
}

int dg_decorrelate_device(int w, int h, const float* diff, const double* knots, int n_knots,
                          const double* theta, int n_theta, double beta, int n_max,
                          const float* var_s, const float* var_r, int block, int radius,
                          int kernel_cell_blocks, float* out, float* out_var, void* stream) {
  return guarded([&] {
    if (!diff || !knots || !theta || !var_s || !var_r || !out || !out_var)
      invalid("decorrelate: null required argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    run_decorrelate(w, h, diff, knots, n_knots, theta, n_theta, beta, n_max, var_s, var_r, block,
                    radius, kernel_cell_blocks, out, out_var, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_decorrelate(int w, int h, const float* diff, const double* knots, int n_knots,
                   const double* theta, int n_theta, double beta, int n_max, const float* var_s,
                   const float* var_r, int block, int radius, int kernel_cell_blocks, float* out,
                   float* out_var) {
  return guarded([&] {
    if (!diff || !knots || !theta || !var_s || !var_r || !out || !out_var)
      invalid("decorrelate: null required argument");
    if (w <= 0 || h <= 0) invalid("decorrelate: empty frame");
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(5 * Arena::round(n * sizeof(float)));
    float* d_diff = ar.take<float>(n);
    float* d_vs = ar.take<float>(n);
    float* d_vr = ar.take<float>(n);
    float* d_out = ar.take<float>(n);
    float* d_var = ar.take<float>(n);
    hipStream_t s = st.get();
    h2d(d_diff, diff, n, s);
    h2d(d_vs, var_s, n, s);
    h2d(d_vr, var_r, n, s);
    run_decorrelate(w, h, d_diff, knots, n_knots, theta, n_theta, beta, n_max, d_vs, d_vr, block,
                    radius, kernel_cell_blocks, d_out, d_var, s);
    d2h(out, d_out, n, s);
    d2h(out_var, d_var, n, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_whiten_score_psf(const float* psf, int ps, const double* knots, int n_knots,
                        const double* theta, int n_theta, double beta, int n_max,
                        const float* var_s, const float* var_r, int w, int h, int block,
                        int radius, float* out) {
  return guarded([&] {
    if (!psf || !knots || !theta || !var_s || !var_r || !out)
      invalid("whiten_score_psf: null required argument");
    if (ps <= 0) runtime("whiten_score_psf: psf must be square");
    if (w <= 0 || h <= 0) invalid("whiten_score_psf: empty frame");
    const size_t n = size_t(w) * h, np = size_t(ps) * ps;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(2 * Arena::round(n * sizeof(float)) + 2 * Arena::round(np * sizeof(float)));
    float* d_vs = ar.take<float>(n);
    float* d_vr = ar.take<float>(n);
    float* d_psf = ar.take<float>(np);
    float* d_out = ar.take<float>(np);
    hipStream_t s = st.get();
    h2d(d_vs, var_s, n, s);
    h2d(d_vr, var_r, n, s);
    h2d(d_psf, psf, np, s);
    run_whiten_score_psf(d_psf, ps, knots, n_knots, theta, n_theta, beta, n_max, d_vs, d_vr, w, h,
                         block, radius, d_out, s);
    d2h(out, d_out, np, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_whiten_score_psf_device(const float* psf, int ps, const double* knots, int n_knots,
                               const double* theta, int n_theta, double beta, int n_max,
                               const float* var_s, const float* var_r, int w, int h, int block,
                               int radius, float* out, void* stream) {
  return guarded([&] {
    if (!psf || !knots || !theta || !var_s || !var_r || !out)
      invalid("whiten_score_psf: null required argument");
    if (ps <= 0) runtime("whiten_score_psf: psf must be square");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    run_whiten_score_psf(psf, ps, knots, n_knots, theta, n_theta, beta, n_max, var_s, var_r, w, h,
                         block, radius, out, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_matched_filter_device(int w, int h, const float* img, const float* psf, int ps,
                             const float* var, float* out, void* stream) {
  return guarded([&] {
    if (!img || !psf || !var || !out) invalid("matched_filter: null required argument");
    if (w <= 0 || h <= 0) return;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    run_matched_filter(w, h, img, psf, ps, var, out, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_matched_filter(int w, int h, const float* img, const float* psf, int ps, const float* var,
                      float* out) {
  return guarded([&] {
    if (!img || !psf || !var || !out) invalid("matched_filter: null required argument");
    if (w <= 0 || h <= 0) return;
    const size_t n = size_t(w) * h, np = size_t(ps) * ps;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(3 * Arena::round(n * sizeof(float)) + Arena::round(np * sizeof(float)));
    float* d_img = ar.take<float>(n);
    float* d_var = ar.take<float>(n);
    float* d_out = ar.take<float>(n);
    float* d_psf = ar.take<float>(np);
    hipStream_t s = st.get();
    h2d(d_img, img, n, s);
    h2d(d_var, var, n, s);
    h2d(d_psf, psf, np, s);
    run_matched_filter(w, h, d_img, d_psf, ps, d_var, d_out, s);
    d2h(out, d_out, n, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_label_device(int w, int h, const float* score, const uint8_t* mask, double threshold,
                    int polarity, int32_t* labels, int* n_out, void* stream) {
  return guarded([&] {
    if (!score || !labels || !n_out) invalid("label: null required argument");
    *n_out = 0;
    if (w <= 0 || h <= 0) return;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(label_scratch_bytes(w, h));
    *n_out = run_label(w, h, score, mask, threshold, polarity, labels, ar,
                       static_cast<hipStream_t>(stream));
  });
}

int dg_label_device_noalloc(int w, int h, const float* score, const uint8_t* mask,
                            double threshold, int polarity, int32_t* labels, int* n_out,
                            void* stream) {
  return guarded([&] {
    if (w < 0 || h < 0) invalid("label_noalloc: negative frame size");
    if (!score || !labels || !n_out) invalid("label: null required argument");
    *n_out = 0;
    *n_out = run_label_noalloc(w, h, score, mask, threshold, polarity, labels,
                               static_cast<hipStream_t>(stream));
  });
}

int dg_label(int w, int h, const float* score, const uint8_t* mask, double threshold,
             int polarity, int32_t* labels, int* n_out) {
  return guarded([&] {
    if (!score || !labels || !n_out) invalid("label: null required argument");
    *n_out = 0;
    if (w <= 0 || h <= 0) return;
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(Arena::round(n * sizeof(float)) + Arena::round(n * sizeof(int32_t)) +
               (mask ? Arena::round(n) : 0) + label_scratch_bytes(w, h));
    float* d_score = ar.take<float>(n);
    int32_t* d_lab = ar.take<int32_t>(n);
    uint8_t* d_mask = mask ? ar.take<uint8_t>(n) : nullptr;
    hipStream_t s = st.get();
    h2d(d_score, score, n, s);
    if (mask) h2d(d_mask, mask, n, s);
    *n_out = run_label(w, h, d_score, d_mask, threshold, polarity, d_lab, ar, s);
    d2h(labels, d_lab, n, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

// ------------------------------------------------------------ photometric scale
int dg_photometric_scale_device(const double* knots, int n_knots, const double* theta,
                                int n_theta, const double* sums, int nc, int w, int h,
                                float* out, void* stream) {
  return guarded([&] {
    if (!knots || !theta || (nc > 0 && !sums) || !out)
      invalid("photometric_scale: null required argument");
    run_photometric_scale(knots, n_knots, theta, n_theta, sums, nc, w, h, out,
                          static_cast<hipStream_t>(stream));
  });
}

int dg_photometric_scale(const double* knots, int n_knots, const double* theta, int n_theta,
                         const double* sums, int nc, int w, int h, float* out) {
  return guarded([&] {
    if (!knots || !theta || (nc > 0 && !sums) || !out)
      invalid("photometric_scale: null required argument");
    if (w <= 0 || h <= 0) invalid("photometric_scale: empty frame");
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    lease->reserve(Arena::round(n * sizeof(float)));
    float* d = lease->take<float>(n);
    run_photometric_scale(knots, n_knots, theta, n_theta, sums, nc, w, h, d, st.get());
    d2h(out, static_cast<const float*>(d), n, st.get());
    DG_HIP(hipStreamSynchronize(st.get()));
  });
}

int dg_photometric_scale_at(const double* knots, int n_knots, const double* theta, int n_theta,
                            const double* sums, int nc, const double* points, int m,
                            double* out) {
  return guarded([&] {
    if (!knots || !theta || (nc > 0 && !sums) || (m > 0 && (!points || !out)))
      invalid("photometric_scale_at: null required argument");
    run_photometric_scale_at(knots, n_knots, theta, n_theta, sums, nc, points, m, out);
  });
}

// ------------------------------------------------------------ catalog
namespace {
void check_catalog_args(int w, int h, const float* score, const float* diff,
                        const dg_catalog_params* p, const dg_catalog_entry* e, int capacity,
                        const int* n_out, const double* bx, const double* by, int nb,
                        const float* psf, int ps) {
  if (!score || !diff || !p || !n_out || (capacity > 0 && !e))
    invalid("build_catalog: null required argument");
  if (w < 0 || h < 0 || capacity < 0) invalid("build_catalog: negative size");
  if (size_t(w) * size_t(h) >= (size_t(1) << 31)) invalid("build_catalog: frame too large");
  if ((bx == nullptr) != (by == nullptr) || (nb > 0 && !bx) || nb < 0)
    runtime("build_catalog: supply both bright_x and bright_y, or neither");
  if (psf && ps <= 0) invalid("build_catalog: psf_size must be > 0");
}
}  // namespace

int dg_build_catalog_device(int w, int h, const float* score, const float* diff,
                            const uint8_t* mask, const dg_catalog_params* p,
                            int return_negative, const float* var, const float* psf, int ps,
                            const double* bx, const double* by, int nb, dg_catalog_entry* e,
                            int capacity, int* n_out, void* stream) {
  return guarded([&] {
    check_catalog_args(w, h, score, diff, p, e, capacity, n_out, bx, by, nb, psf, ps);
    *n_out = run_build_catalog(w, h, score, diff, mask, *p, return_negative, var, psf, ps, bx,
                               by, nb, e, capacity, static_cast<hipStream_t>(stream));
  });
}

int dg_build_catalog_labelled_device(int w, int h, const float* score, const float* diff,
                                     const uint8_t* mask, const dg_catalog_params* p,
                                     int return_negative, const float* var, const float* psf,
                                     int ps, const double* bx, const double* by, int nb,
                                     const int32_t* labels, int n_labels, dg_catalog_entry* e,
                                     int capacity, int* n_out, void* stream) {
  return guarded([&] {
    check_catalog_args(w, h, score, diff, p, e, capacity, n_out, bx, by, nb, psf, ps);
    if (!labels && size_t(w) * size_t(h) > 0) invalid("build_catalog: null labels");
    if (n_labels < 0) invalid("build_catalog: negative label count");
    *n_out = run_build_catalog(w, h, score, diff, mask, *p, return_negative, var, psf, ps, bx,
                               by, nb, e, capacity, static_cast<hipStream_t>(stream), labels,
                               n_labels);
  });
}

int dg_build_catalog(int w, int h, const float* score, const float* diff, const uint8_t* mask,
                     const dg_catalog_params* p, int return_negative, const float* var,
                     const float* psf, int ps, const double* bx, const double* by, int nb,
                     dg_catalog_entry* e, int capacity, int* n_out) {
  return guarded([&] {
    check_catalog_args(w, h, score, diff, p, e, capacity, n_out, bx, by, nb, psf, ps);
    *n_out = 0;
    if (w == 0 || h == 0) return;
    const size_t n = size_t(w) * h;
    const size_t np = psf ? size_t(ps) * ps : 0;
    OwnedStream st;
    hipStream_t s = st.get();
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(Arena::round(n * sizeof(float)) * (var ? 3 : 2) + (mask ? Arena::round(n) : 0) +
               Arena::round(np * sizeof(float) + 4) + 2 * Arena::round(sizeof(double) * (nb + 1)) +
               Arena::round(sizeof(dg_catalog_entry) * (size_t(capacity) + 1)));
    float* ds = ar.take<float>(n);
    float* dd = ar.take<float>(n);
    h2d(ds, score, n, s);
    h2d(dd, diff, n, s);
    float* dv = nullptr;
    if (var) {
      dv = ar.take<float>(n);
      h2d(dv, var, n, s);
    }
    uint8_t* dm = nullptr;
    if (mask) {
      dm = ar.take<uint8_t>(n);
      h2d(dm, mask, n, s);
    }
    float* dp = nullptr;
    if (psf) {
      dp = ar.take<float>(np + 1);
      h2d(dp, psf, np, s);
    }
    double *dbx = nullptr, *dby = nullptr;
    if (bx && nb > 0) {
      dbx = ar.take<double>(nb + 1);
      dby = ar.take<double>(nb + 1);
      h2d(dbx, bx, nb, s);
      h2d(dby, by, nb, s);
    }
    dg_catalog_entry* de = ar.take<dg_catalog_entry>(size_t(capacity) + 1);
    const int nc = run_build_catalog(w, h, ds, dd, dm, *p, return_negative, dv, dp, ps, dbx, dby,
                                     bx ? nb : 0, de, capacity, s);
    d2h(e, static_cast<const dg_catalog_entry*>(de), size_t(std::min(nc, capacity)), s);
    DG_HIP(hipStreamSynchronize(s));
    *n_out = nc;
  });
}

// ------------------------------------------------------------ stamp selection
namespace {
void check_stamp_capacity(int capacity, int need, const char* what) {
  if (capacity < need)
    invalid(std::string(what) + ": output capacity below max_stamps / n_catalog");
}
// Host image (+ mask) staged into `ar`; returns device pointers.
std::pair<const float*, const uint8_t*> stage_image(const float* img, const uint8_t* mask, int w,
                                                    int h, Arena& ar, hipStream_t s) {
  const size_t n = size_t(w) * h;
  float* d = ar.take<float>(n);
  h2d(d, img, n, s);
  uint8_t* m = nullptr;
  if (mask) {
    m = ar.take<uint8_t>(n);
    h2d(m, mask, n, s);
  }
  return {d, m};
}
size_t image_bytes(int w, int h, bool mask) {
  const size_t n = size_t(std::max(w, 0)) * std::max(h, 0);
  return Arena::round(n * sizeof(float)) + (mask ? Arena::round(n) : 0);
}
}  // namespace

int dg_estimate_background_device(int w, int h, const float* data, const uint8_t* mask,
                                  double* median, double* sigma, void* stream) {
  return guarded([&] {
    if (!data || !median || !sigma) invalid("estimate_background: null required argument");
    run_estimate_background(data, mask, w, h, median, sigma, static_cast<hipStream_t>(stream));
  });
}

int dg_estimate_background(int w, int h, const float* data, const uint8_t* mask, double* median,
                           double* sigma) {
  return guarded([&] {
    if (!data || !median || !sigma) invalid("estimate_background: null required argument");
    if (w <= 0 || h <= 0) invalid("estimate_background: empty image");
    OwnedStream st;
    ArenaLease lease;
    lease->reserve(image_bytes(w, h, mask != nullptr));
    auto d = stage_image(data, mask, w, h, *lease, st.get());
    run_estimate_background(d.first, d.second, w, h, median, sigma, st.get());
  });
}

int dg_detect_stamps_device(int w, int h, const float* data, const uint8_t* mask,
                            const dg_detect_params* p, int32_t* x, int32_t* y, double* flux,
                            double* snr, double* fwhm, int capacity, int* n_out, void* stream) {
  return guarded([&] {
    if (!data || !p || !x || !y || !flux || !snr || !fwhm || !n_out)
      invalid("detect_stamps: null required argument");
    *n_out = 0;
    check_stamp_capacity(capacity, p->max_stamps, "detect_stamps");
    *n_out = run_detect_stamps(data, mask, w, h, *p, x, y, flux, snr, fwhm,
                               static_cast<hipStream_t>(stream));
  });
}

int dg_detect_stamps(int w, int h, const float* data, const uint8_t* mask,
                     const dg_detect_params* p, int32_t* x, int32_t* y, double* flux,
                     double* snr, double* fwhm, int capacity, int* n_out) {
  return guarded([&] {
    if (!data || !p || !x || !y || !flux || !snr || !fwhm || !n_out)
      invalid("detect_stamps: null required argument");
    *n_out = 0;
    check_stamp_capacity(capacity, p->max_stamps, "detect_stamps");
    if (w <= 0 || h <= 0) invalid("detect_stamps: empty image");
    const int k = std::max(p->max_stamps, 0) + 1;
    OwnedStream st;
    hipStream_t s = st.get();
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(image_bytes(w, h, mask != nullptr) + 2 * Arena::round(sizeof(int32_t) * k) +
               3 * Arena::round(sizeof(double) * k));
    auto d = stage_image(data, mask, w, h, ar, s);
    int32_t *dx = ar.take<int32_t>(k), *dy = ar.take<int32_t>(k);
    double *df = ar.take<double>(k), *ds = ar.take<double>(k), *dw = ar.take<double>(k);
    const int n = run_detect_stamps(d.first, d.second, w, h, *p, dx, dy, df, ds, dw, s);
    d2h(x, dx, n, s);
    d2h(y, dy, n, s);
    d2h(flux, static_cast<const double*>(df), n, s);
    d2h(snr, static_cast<const double*>(ds), n, s);
    d2h(fwhm, static_cast<const double*>(dw), n, s);
    DG_HIP(hipStreamSynchronize(s));
    *n_out = n;
  });
}

int dg_select_stamps_device(int ws, int hs, const float* sci, const uint8_t* smask, int wr,
                            int hr, const float* ref, const uint8_t* rmask,
                            const int32_t* cat_x, const int32_t* cat_y, int n_cat,
                            const dg_detect_params* p, int32_t* x, int32_t* y, double* fs,
                            double* fr, int capacity, int* n_out, double* medians,
                            int* direction, void* stream) {
  return guarded([&] {
    if (!sci || !ref || !p || !x || !y || !fs || !fr || !n_out || !medians || !direction)
      invalid("select_stamps: null required argument");
    *n_out = 0;
    const bool cat = cat_x && cat_y;
    check_stamp_capacity(capacity, cat ? n_cat : p->max_stamps, "select_stamps");
    *n_out = run_select_stamps(ws, hs, sci, smask, wr, hr, ref, rmask, cat_x, cat_y, n_cat, *p,
                               x, y, fs, fr, medians, direction,
                               static_cast<hipStream_t>(stream));
  });
}

int dg_select_stamps(int ws, int hs, const float* sci, const uint8_t* smask, int wr, int hr,
                     const float* ref, const uint8_t* rmask, const int32_t* cat_x,
                     const int32_t* cat_y, int n_cat, const dg_detect_params* p, int32_t* x,
                     int32_t* y, double* fs, double* fr, int capacity, int* n_out,
                     double* medians, int* direction) {
  return guarded([&] {
    if (!sci || !ref || !p || !x || !y || !fs || !fr || !n_out || !medians || !direction)
      invalid("select_stamps: null required argument");
    *n_out = 0;
    const bool cat = cat_x && cat_y;
    const int nc = cat ? std::max(n_cat, 0) : 0;
    const int cap = cat ? nc : std::max(p->max_stamps, 0);
    check_stamp_capacity(capacity, cat ? n_cat : p->max_stamps, "select_stamps");
    if (ws <= 0 || hs <= 0 || wr <= 0 || hr <= 0) invalid("select_stamps: empty image");
    OwnedStream st;
    hipStream_t s = st.get();
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(image_bytes(ws, hs, smask != nullptr) + image_bytes(wr, hr, rmask != nullptr) +
               4 * Arena::round(sizeof(int32_t) * (std::max(cap, nc) + 1)) +
               2 * Arena::round(sizeof(double) * (cap + 1)));
    auto ds = stage_image(sci, smask, ws, hs, ar, s);
    auto dr = stage_image(ref, rmask, wr, hr, ar, s);
    int32_t *dcx = nullptr, *dcy = nullptr;
    if (cat) {
      dcx = ar.take<int32_t>(nc + 1);
      dcy = ar.take<int32_t>(nc + 1);
      h2d(dcx, cat_x, nc, s);
      h2d(dcy, cat_y, nc, s);
    }
    int32_t *dx = ar.take<int32_t>(cap + 1), *dy = ar.take<int32_t>(cap + 1);
    double *dfs = ar.take<double>(cap + 1), *dfr = ar.take<double>(cap + 1);
    const int n = run_select_stamps(ws, hs, ds.first, ds.second, wr, hr, dr.first, dr.second,
                                    dcx, dcy, nc, *p, dx, dy, dfs, dfr, medians, direction, s);
    d2h(x, dx, n, s);
    d2h(y, dy, n, s);
    d2h(fs, static_cast<const double*>(dfs), n, s);
    d2h(fr, static_cast<const double*>(dfr), n, s);
    DG_HIP(hipStreamSynchronize(s));
    *n_out = n;
  });
}

// ------------------------------------------------------------ constant-kernel model
int dg_kernel_chi2_device(int w, int h, const float* sci, const float* ref, const float* weight,
                          const double* params, int n_max, int radius, int want_grad,
                          double* out, void* stream) {
  return guarded([&] {
    if (!sci || !ref || !params || !out) invalid("kernel_chi2: null required argument");
    kernel_model_check(w, h, n_max, radius);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(kernel_model_scratch_bytes(w, h, radius));
    run_kernel_model(0, w, h, sci, ref, weight, nullptr, params, n_max, radius, want_grad,
                     nullptr, out, ar, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_kernel_chi2(int w, int h, const float* sci, const float* ref, const float* weight,
                   const double* params, int n_max, int radius, int want_grad, double* out) {
  return guarded([&] {
    if (!sci || !ref || !params || !out) invalid("kernel_chi2: null required argument");
    kernel_model_check(w, h, n_max, radius);
    const int nc = (n_max + 1) * (n_max + 2) / 2;
    if (!(params[nc + 1] > 0.0)) invalid("beta must be > 0");
    const size_t n = size_t(w) * h;
    OwnedStream st;
    ArenaLease lease;
    Arena& ar = *lease;
    const size_t fb = Arena::round(n * sizeof(float));
    ar.reserve((weight ? 3 : 2) * fb + 2 * Arena::round((nc + 5) * sizeof(double)) +
               kernel_model_scratch_bytes(w, h, radius));
    float* d_sci = ar.take<float>(n);
    float* d_ref = ar.take<float>(n);
    float* d_w = weight ? ar.take<float>(n) : nullptr;
    double* d_par = ar.take<double>(nc + 4);
    double* d_out = ar.take<double>(nc + 5);
    hipStream_t s = st.get();
    h2d(d_sci, sci, n, s);
    h2d(d_ref, ref, n, s);
    if (weight) h2d(d_w, weight, n, s);
    h2d(d_par, params, nc + 4, s);
    run_kernel_model(0, w, h, d_sci, d_ref, d_w, nullptr, d_par, n_max, radius, want_grad,
                     nullptr, d_out, ar, s);
    d2h(out, d_out, want_grad ? nc + 5 : 1, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_kernel_difference_device(int w, int h, const float* sci, const float* ref,
                                const double* params, int n_max, int radius, float* diff,
                                void* stream) {
  return guarded([&] {
    if (!sci || !ref || !params || !diff) invalid("kernel_difference: null required argument");
    kernel_model_check(w, h, n_max, radius);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(kernel_model_scratch_bytes(w, h, radius));
    run_kernel_model(0, w, h, sci, ref, nullptr, nullptr, params, n_max, radius, 0, diff, nullptr,
                     ar, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

int dg_kernel_difference_vjp_device(int w, int h, const float* ref, const float* grad_diff,
                                    const double* params, int n_max, int radius, double* out,
                                    void* stream) {
  return guarded([&] {
    if (!ref || !grad_diff || !params || !out)
      invalid("kernel_difference_vjp: null required argument");
    kernel_model_check(w, h, n_max, radius);
    hipStream_t s = static_cast<hipStream_t>(stream);
    ArenaLease lease;
    Arena& ar = *lease;
    ar.reserve(kernel_model_scratch_bytes(w, h, radius));
    run_kernel_model(1, w, h, nullptr, ref, nullptr, grad_diff, params, n_max, radius, 1, nullptr,
                     out, ar, s);
    DG_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
