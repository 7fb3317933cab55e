/*
 * deltagpu.h -- C ABI of libdeltagpu, the MI355X (gfx950) implementation of
 * delta's PSF-matching hot path.
 *
 * Conventions (mirroring the reference's nanobind boundary, src/bindings.cpp):
 *   - images are float32, C-contiguous, row-major (index y*width + x);
 *     masks are uint8 bit flags (1 bad, 2 saturated, 4 cosmic, 8 non-finite,
 *     16 edge, 32 user; include/delta/image.hpp:14-22);
 *   - knots are (n_knots, 2) float64 row-major; theta is float64 of length
 *     n_knots * (nc + 1), ordered [c_0 | ... | c_{nc-1} | b] (solve.hpp:12-15);
 *     n_knots == 1 selects a spatially constant kernel;
 *   - optional inputs are NULL; outputs are caller-allocated;
 *   - plain entry points take HOST pointers, copy in, run on the current
 *     device on a private stream, copy out and return when done; `_device`
 *     entry points take DEVICE pointers for every image and run on the given
 *     hipStream_t (passed as void*), returning after the stream is drained;
 *   - every call returns 0 on success, DG_ERR_INVALID for arguments the
 *     reference rejects with std::invalid_argument (Python ValueError),
 *     DG_ERR_RUNTIME for its std::runtime_error cases (Python RuntimeError),
 *     DG_ERR_DEVICE for a HIP failure; dg_last_error() gives the message
 *     (thread-local).
 */
#ifndef DELTAGPU_H
#define DELTAGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DG_OK 0
#define DG_ERR_INVALID 1
#define DG_ERR_RUNTIME 2
#define DG_ERR_DEVICE 3

/* Library version (major*10000 + minor*100 + patch). */
int dg_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* dg_last_error(void);

/* Default kernel half-width max(ceil(3 beta sqrt(n_max+1)), 3).
 * Replaces GaussHermiteBasis::default_radius (src/basis.cpp:9-12). */
int dg_default_radius(double beta, int n_max);

/*
 * Sampled 1-D Gauss-Hermite functions, (n_max+1) x ksize FP64, ksize = 2r+1
 * (radius <= 0 selects dg_default_radius).  Replaces
 * _core.gauss_hermite_basis1d (src/bindings.cpp:171-185) /
 * GaussHermiteBasis::basis1d (src/basis.cpp:14-59).
 */
int dg_gauss_hermite_basis1d(double beta, int n_max, int radius, double* out);

/*
 * Footprint sums S_c = (sum_i g_nx[i]) (sum_i g_ny[i]) of the nc components,
 * each 1-D sum taken in index order in FP64.  Replaces
 * GaussHermiteBasis::component_sums (src/basis.cpp:81-93), which fit_kernel
 * returns as "component_sums" (src/bindings.cpp:668-692).
 */
int dg_component_sums(double beta, int n_max, int radius, double* out);

/*
 * The nc = (n_max+1)(n_max+2)/2 2-D components: orders[nc][2] = (nx, ny) by
 * total order then ny, kernels[nc][ksize][ksize] = (float)(g_ny[y] g_nx[x]).
 * Replaces _core.gauss_hermite_kernels (src/bindings.cpp:187-206) /
 * GaussHermiteBasis::kernel2d (src/basis.cpp:66-79).
 */
int dg_gauss_hermite_kernels(double beta, int n_max, int radius, int32_t* orders,
                             float* kernels);

/*
 * Image convolved with every basis component, out[nc][height][width] FP32:
 * zero-padded separable x- then y-passes, bit-equal to the reference.
 * Replaces _core.basis_convolve (src/bindings.cpp:209-227) /
 * delta::basis_convolve (src/convolve.cpp:24-113).  radius in [1, 20].
 */
int dg_basis_convolve(int width, int height, const float* image, double beta, int n_max,
                      int radius, float* out);
int dg_basis_convolve_device(int width, int height, const float* image, double beta,
                             int n_max, int radius, float* out, void* stream);

/*
 * Thin-plate regression spline over n_knots knots (x, y pairs):
 *   dg_tps_design   design rows of n points, out[n][n_knots]
 *                   (_core.tps_design, src/bindings.cpp:487-490; src/spatial.cpp:67-97);
 *   dg_tps_evaluate design(points) . coeffs, out[n]
 *                   (_core.tps_evaluate, src/bindings.cpp:502-506);
 *   dg_tps_penalty  bending-energy penalty, out[n_knots][n_knots]
 *                   (_core.tps_penalty, src/bindings.cpp:492-494; src/spatial.cpp:20-65);
 *   dg_tps_fit      theta = (D^T D + lam P)^-1 D^T values, theta[n_knots]
 *                   (_core.tps_fit, src/bindings.cpp:496-500; src/spatial.cpp:131-140).
 */
int dg_tps_design(const double* knots, int n_knots, const double* points, int n, double* out);
int dg_tps_evaluate(const double* knots, int n_knots, const double* points, int n,
                    const double* coeffs, int n_coeffs, double* out);
int dg_tps_penalty(const double* knots, int n_knots, double* out);
int dg_tps_fit(const double* knots, int n_knots, const double* points, const double* values,
               int n, double lam, double* theta);

/*
 * One square n x n block whitened with a constant ksize x ksize kernel and
 * scalar noise levels: r2c, times Phi = sqrt(vs + vr sum K^2) /
 * sqrt(vs + vr |Khat|^2), c2r / n^2.  Replaces _core.decorrelate_block
 * (src/bindings.cpp:725-745; src/noise.cpp:98-137).  n a power of two in
 * [8, 512] on the GPU.
 */
int dg_decorrelate_block(int n, const float* image, const float* kernel, int ksize,
                         double var_science, double var_reference, float* out);

/*
 * The real-space decorrelation kernel ifft(Phi) / n^2, fft-shifted to the
 * block centre, out[n][n].  Replaces _core.decorrelation_kernel
 * (src/bindings.cpp:748-763; src/noise.cpp:139-161).
 */
int dg_decorrelation_kernel(const float* kernel, int ksize, double var_science,
                            double var_reference, int n, float* out);

/*
 * Variance synthesised from gain and read noise, Var = max(data, 0) / gain +
 * read_noise^2 in FP32 (NaN stays NaN).  Replaces delta.synth_variance
 * (python/delta/_inputs.py:75-86), which the pipeline calls when a frame has
 * no variance layer.  gain <= 0 is an invalid argument.
 */
int dg_synth_variance(int width, int height, const float* data, double gain, double read_noise,
                      float* out);
int dg_synth_variance_device(int width, int height, const float* data, double gain,
                             double read_noise, float* out, void* stream);

/*
 * Scalar that brings the difference image's reduced chi^2 to one:
 * median(d^2 / v) / 0.4549364231195724 over pixels with v > 0, d and v
 * finite and mask == 0 (mask may be NULL); the median of an even count is the
 * mean of the two middle values; 1.0 when no pixel qualifies.  Replaces
 * _variance_rescale_factor (python/delta/pipeline.py:66-86).
 */
int dg_variance_rescale_factor(int width, int height, const float* difference,
                               const float* variance, const uint8_t* mask, double* factor);
int dg_variance_rescale_factor_device(int width, int height, const float* difference,
                                      const float* variance, const uint8_t* mask,
                                      double* factor, void* stream);

/*
 * Score PSF: the median over sources of each (2r+1)^2 cutout (fully inside
 * the frame) minus its median and divided by its sum (kept when the sum is
 * positive), clipped at 0 and normalised to unit sum; psf[(2r+1)^2] FP32,
 * numpy's arithmetic throughout.  xs, ys are host arrays of n positions.
 * Replaces _estimate_psf (python/delta/pipeline.py:285-302; it has no C++
 * counterpart in the reference).  No usable source is an invalid argument
 * ("no usable sources for PSF estimation").  radius in [0, 15].
 */
int dg_estimate_psf(int width, int height, const float* image, const int32_t* xs,
                    const int32_t* ys, int n, int radius, float* psf);
int dg_estimate_psf_device(int width, int height, const float* image, const int32_t* xs,
                           const int32_t* ys, int n, int radius, float* psf, void* stream);

/*
 * Inverse-variance weighted mean over pixels with mask == 0 and variance > 0:
 * sum(d / v) / sum(1 / v) in FP64, 0 when no pixel qualifies (summed in
 * parallel, so equal to the reference's sequential sum to rounding).
 * Replaces _core.weighted_mean (src/bindings.cpp:107-125).
 */
int dg_weighted_mean(int width, int height, const float* data, const float* variance,
                     const uint8_t* mask, double* out);

/*
 * Full-frame spatially varying subtraction
 *   D = S - b - sum_n a_n (phi_n (x) R),  Var(D) = Var(S) + K^2 (x) Var(R),
 * with mask growth.  Replaces _core.subtract (src/bindings.cpp:603-638) and
 * delta::subtract (include/delta/subtract.hpp:75-78, src/subtract.cpp:429-966).
 * radius <= 0 selects dg_default_radius.  `variance` must be non-NULL iff
 * science_var or reference_var is; `mask` non-NULL iff a mask is given or
 * saturation > 0 (the reference returns None otherwise).
 */
int dg_subtract(int width, int height, const float* science, const float* reference,
                const float* science_var, const float* reference_var,
                const uint8_t* science_mask, const uint8_t* reference_mask,
                const double* knots, int n_knots, const double* theta, int n_theta,
                double beta, int n_max, int radius, double saturation,
                float* difference, float* variance, uint8_t* mask);

/* Device-pointer form of dg_subtract (images, variances, masks and outputs on
 * the device; knots and theta on the host). */
int dg_subtract_device(int width, int height, const float* science, const float* reference,
                       const float* science_var, const float* reference_var,
                       const uint8_t* science_mask, const uint8_t* reference_mask,
                       const double* knots, int n_knots, const double* theta, int n_theta,
                       double beta, int n_max, int radius, double saturation,
                       float* difference, float* variance, uint8_t* mask, void* stream);

/* Precision of the normal-equation build M = X^T W X (the Cholesky, GCV and
 * residual sums are FP64 in every mode).  The thin-plate normal matrix is
 * ill-conditioned (cond(M + lambda P0) ~ 1e10 on a 4096^2 frame), so only
 * the FP64 build reproduces the reference's solution; the other two are
 * faster approximations whose fitted kernel differs measurably. */
#define DG_PREC_FP32 0        /* FP32 MFMA, FP32 accumulation flushed to FP64 */
#define DG_PREC_SPLIT_BF16 1  /* split-BF16 hi/lo, 3 MFMA products, FP32 acc */
#define DG_PREC_FP64 2        /* FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 sums */

/*
 * Kernel fit from stamps: stamp B_n gather, penalised GLS over the lambda
 * grid, IRLS reweighting and robust stamp clipping.  The smoothing is chosen
 * by GCV (cv_folds <= 1) or by leave-stamp-out CV over cv_folds folds (fold =
 * contributing stamp index mod cv_folds), built per stamp with the design
 * frozen at each stamp's pixel centroid when the knots are >= 16 stamp widths
 * apart (environment DELTA_STAMP_APPROX=1/0 forces it), else from the rows;
 * its gcv_curve is then the CV error, +inf where the coarse-to-fine search did
 * not evaluate it.  Replaces _core.fit_kernel (src/bindings.cpp:643-693) and
 * delta::fit_kernel (include/delta/fit.hpp:65-72, src/fit.cpp:63-511).
 *   science = the fit target, reference = the image convolved (as the
 *   reference's binding: the caller swaps them for direction "science").
 * Outputs (caller-allocated):
 *   theta[n_knots * (nc + 1)], gcv_curve[n_lambda],
 *   scalars[6] = {lambda, gcv, effective_dof, rss, reduced_chi2,
 *                 cv_exact_design_bytes},
 *   counts[5]  = {n_pixels, n_stamps_used, n_stamps_total, n_stamps_rejected,
 *                 IRLS passes},
 *   per contributing stamp (capacity n_stamps, first counts[2] valid):
 *   stamp_x_out, stamp_y_out, stamp_chi2, stamp_accepted.
 * Errors: DG_ERR_RUNTIME "fit_kernel: no usable stamp pixels" / "all stamp
 * pixels clipped" / "lambda_grid is empty" as the reference.
 */
int dg_fit_kernel(int width, int height, const float* science, const float* reference,
                  const float* science_var, const float* reference_var,
                  const uint8_t* science_mask, const uint8_t* reference_mask,
                  const double* knots, int n_knots, const int32_t* stamp_x,
                  const int32_t* stamp_y, int n_stamps, int stamp_radius, double beta,
                  int n_max, int radius, const double* lambda_grid, int n_lambda,
                  double clip_sigma, int clip_iterations, int min_stamps, int cv_folds,
                  int precision, double* theta, double* gcv_curve, double* scalars,
                  int32_t* counts, int32_t* stamp_x_out, int32_t* stamp_y_out,
                  double* stamp_chi2, uint8_t* stamp_accepted);

/* Device-pointer form of dg_fit_kernel (images on the device; stamps, knots,
 * grid and all outputs on the host). */
int dg_fit_kernel_device(int width, int height, const float* science, const float* reference,
                         const float* science_var, const float* reference_var,
                         const uint8_t* science_mask, const uint8_t* reference_mask,
                         const double* knots, int n_knots, const int32_t* stamp_x,
                         const int32_t* stamp_y, int n_stamps, int stamp_radius, double beta,
                         int n_max, int radius, const double* lambda_grid, int n_lambda,
                         double clip_sigma, int clip_iterations, int min_stamps, int cv_folds,
                         int precision, double* theta, double* gcv_curve, double* scalars,
                         int32_t* counts, int32_t* stamp_x_out, int32_t* stamp_y_out,
                         double* stamp_chi2, uint8_t* stamp_accepted, void* stream);

/*
 * Penalised GLS with GCV over lambda_grid on caller rows (host pointers):
 * points (n, 2), target (n), weights (n), bn (n, nc), all float64.
 * Replaces _core.solve_gls_gcv (src/bindings.cpp:540-551) / delta::solve_gls_gcv
 * (include/delta/solve.hpp:40-45).  scalars[4] = {lambda, gcv, effective_dof, rss}.
 */
int dg_solve_gls_gcv(const double* knots, int n_knots, const double* points,
                     const double* target, const double* weights, const double* bn, int n,
                     int nc, const double* lambda_grid, int n_lambda, int precision,
                     double* theta, double* scalars, double* gcv_curve);

/*
 * As dg_solve_gls_gcv with lambda chosen by leave-group-out cross-validation:
 * group[n] is each row's fold (0 .. n_groups-1; n_groups < 2 is GCV), the
 * lambdas are visited coarse-to-fine (every other one, then the +-2 window
 * of the best), each fold's error fits on the other folds; gcv_curve holds
 * the CV error, +inf where not evaluated.  Replaces _core.solve_gls_cv
 * (src/bindings.cpp:553-566) / delta::solve_gls_cv (src/solve.cpp:420-562).
 */
int dg_solve_gls_cv(const double* knots, int n_knots, const double* points,
                    const double* target, const double* weights, const double* bn, int n,
                    int nc, const double* lambda_grid, int n_lambda, const int32_t* group,
                    int n_groups, int precision, double* theta, double* scalars,
                    double* gcv_curve);

/*
 * Spatially varying noise decorrelation of a difference image with
 * Hann-apodised n x n FFT blocks at stride n/2 (n = block, a power of two,
 * 8..512), the decorrelation filter
 *   Phi(k) = sqrt(max(vs + vr sum K^2, 0)) / sqrt(vs + vr |Khat(k)|^2)
 * with K the local matching kernel cached per cell of kernel_cell_blocks^2
 * blocks (<= 0: auto) and vs, vr the robust median noise of each block.
 * Outputs the whitened difference and its post-whitening variance.
 * Replaces _core.decorrelate (src/bindings.cpp:766-792) and
 * delta::decorrelate (include/delta/noise.hpp:64-68, src/noise.cpp:228-483).
 */
int dg_decorrelate(int width, int height, const float* difference, const double* knots,
                   int n_knots, const double* theta, int n_theta, double beta, int n_max,
                   const float* var_science, const float* var_reference, int block, int radius,
                   int kernel_cell_blocks, float* out_difference, float* out_variance);
int dg_decorrelate_device(int width, int height, const float* difference, const double* knots,
                          int n_knots, const double* theta, int n_theta, double beta, int n_max,
                          const float* var_science, const float* var_reference, int block,
                          int radius, int kernel_cell_blocks, float* out_difference,
                          float* out_variance, void* stream);

/*
 * The score PSF whitened by Phi at the frame centre (kernel and block noise
 * over the central block), centre-cropped and normalised to unit sum.
 * `psf` is psf_size x psf_size; the variances are width x height (host).
 * Replaces _core.whiten_score_psf (src/bindings.cpp:796-874) and
 * delta::whiten_psf (src/noise.cpp:485-533).
 */
int dg_whiten_score_psf(const float* psf, int psf_size, const double* knots, int n_knots,
                        const double* theta, int n_theta, double beta, int n_max,
                        const float* var_science, const float* var_reference, int width,
                        int height, int block, int radius, float* out);
int dg_whiten_score_psf_device(const float* psf, int psf_size, const double* knots, int n_knots,
                               const double* theta, int n_theta, double beta, int n_max,
                               const float* var_science, const float* var_reference, int width,
                               int height, int block, int radius, float* out, void* stream);

/*
 * Match-filtered score image: score(x) = sum_u psf(u) image(x+u) /
 * sqrt(var(x) sum psf^2), zero-padded, 0 where var <= 0; psf_size odd or
 * even, at most 63.  Bit-identical to the reference's float arithmetic.
 * Replaces _core.matched_filter (src/bindings.cpp:878-898) and
 * delta::matched_filter (include/delta/noise.hpp:88-93, src/noise.cpp:535-632).
 */
int dg_matched_filter(int width, int height, const float* image, const float* psf,
                      int psf_size, const float* variance, float* out);
int dg_matched_filter_device(int width, int height, const float* image, const float* psf,
                             int psf_size, const float* variance, float* out, void* stream);

/*
 * Candidate seeds (finite, score >= threshold for polarity > 0 or
 * score <= -threshold for polarity < 0, threshold compared in float; with a
 * mask only mask == 0 pixels) and their 8-connected labels, numbered 1..n in
 * raster order of each component's first pixel (0 = background); *n_out gets
 * the component count.  Replaces seed_mask + label_components
 * (src/catalog.cpp:61-80, 105-261) inside delta::build_catalog (:600-697).
 */
int dg_label(int width, int height, const float* score, const uint8_t* mask, double threshold,
             int polarity, int32_t* labels, int* n_out);
int dg_label_device(int width, int height, const float* score, const uint8_t* mask,
                    double threshold, int polarity, int32_t* labels, int* n_out, void* stream);

/*
 * dg_label_device without any device allocation: the labelling's scratch (a
 * seed-occupancy byte per 32 x 32 tile, the per-row root counts and their
 * offsets) lives in a 1 MiB static device array that each GPU holds until the
 * process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Frames up to 2^28 pixels whose
 * scratch fits the array (every frame up to 16384 x 16384); others fail with
 * a clear error before anything runs.  Same arguments, labels and count as dg_label_device.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller.
 */
int dg_label_device_noalloc(int width, int height, const float* score, const uint8_t* mask,
                            double threshold, int polarity, int32_t* labels, int* n_out,
                            void* stream);

/*
 * Photometric scale field sum_c S_c a_c(x, y) over the frame (FP32, the
 * reference's coarse-lattice fields and float accumulation in component
 * order), S = component_sums (nc entries), theta of length n_knots * (nc + 1).
 * Replaces _core.photometric_scale (src/bindings.cpp:695-707) and
 * delta::photometric_scale (src/fit.cpp:513-530).  In the _device form `out`
 * is a device pointer (knots, theta and sums stay on the host).
 */
int dg_photometric_scale(const double* knots, int n_knots, const double* theta, int n_theta,
                         const double* component_sums, int nc, int width, int height,
                         float* out);
int dg_photometric_scale_device(const double* knots, int n_knots, const double* theta,
                                int n_theta, const double* component_sums, int nc, int width,
                                int height, float* out, void* stream);

/*
 * The same scale at m points (m x 2 float64 x, y), exact TPS fields in FP64.
 * Replaces _core.photometric_scale_at (src/bindings.cpp:710-719) and
 * delta::photometric_scale_at (src/fit.cpp:532-545).  Host pointers.
 */
int dg_photometric_scale_at(const double* knots, int n_knots, const double* theta, int n_theta,
                            const double* component_sums, int nc, const double* points, int m,
                            double* out);

/*
 * Catalog knobs, field for field delta::CatalogParams
 * (include/delta/catalog.hpp:57-84); booleans as int.
 */
typedef struct dg_catalog_params {
  double threshold_sigma;        /* positive seed/grow threshold (5.0) */
  double threshold_sigma_dipole; /* |score| threshold of the negative blobs (3.0) */
  double expected_fwhm;          /* px, for the consistency filter and aperture (3.5) */
  double fwhm_tolerance_lo;      /* fwhm_ratio bounds of "consistent" (0.3, 3.0) */
  double fwhm_tolerance_hi;
  int aperture_radius;           /* 0 -> max(1, lround(3 * expected_fwhm / 2.3548)) */
  int exclude_bad_pixels;        /* masked pixels neither seed nor grow (1) */
  int fit_psf_shape;             /* PSF chi^2 fit (needs psf and variance) (0) */
  double shape_chi2_tolerance;   /* reduced chi^2 cut of shape_consistent (3.0) */
  double bright_star_radius;     /* 0 -> 3 * expected_fwhm */
  int max_shape_fit;             /* enrichment skipped above this many entries (5000) */
} dg_catalog_params;

/*
 * One catalog row, the fields of delta::CatalogEntry
 * (include/delta/catalog.hpp:14-56) ordered doubles, int32, bytes (104 bytes,
 * no interior padding).
 */
typedef struct dg_catalog_entry {
  double x, y, peak_snr, flux, expected_n_pix, fwhm_ratio;
  double psf_chi2, psf_amplitude, psf_amplitude_err, bright_star_dist;
  int32_t peak_x, peak_y, n_pix, psf_chi2_dof;
  uint8_t fwhm_consistent, mask_flags, is_dipole, shape_consistent, near_bright_star, quality;
  uint8_t pad[2];
} dg_catalog_entry;

/*
 * Connected-component catalog of a score image: seeds (finite, score >=
 * threshold_sigma, or <= -threshold_sigma_dipole for return_negative; unmasked
 * when a mask is given and exclude_bad_pixels), 8-connected labels in raster
 * order of first pixels, and per component the |score|-weighted centroid
 * (summed in raster order), peak (first of equal |score|), pixel count,
 * expected count and FWHM consistency, the OR of mask flags over footprint and
 * boundary, aperture flux on `difference`, dipole flag (an opposite-sign seed
 * pixel in a member's 3x3 neighbourhood) and, for at most max_shape_fit rows,
 * the PSF-shape fit (psf_size^2 window, needs psf and variance) and the
 * nearest of n_bright bright stars.  Rows are written in label order, at most
 * `capacity` of them; *n_out gets the full count (call again with a larger
 * buffer when it exceeds capacity).  Replaces _core.build_catalog
 * (src/bindings.cpp:348-483) and delta::build_catalog (src/catalog.cpp:262-697).
 * In the _device form images, psf, bright_x/bright_y and `entries` are device
 * pointers; n_out stays a host pointer.
 */
int dg_build_catalog(int width, int height, const float* score, const float* difference,
                     const uint8_t* mask, const dg_catalog_params* params, int return_negative,
                     const float* variance, const float* psf, int psf_size,
                     const double* bright_x, const double* bright_y, int n_bright,
                     dg_catalog_entry* entries, int capacity, int* n_out);
int dg_build_catalog_device(int width, int height, const float* score,
                            const float* difference, const uint8_t* mask,
                            const dg_catalog_params* params, int return_negative,
                            const float* variance, const float* psf, int psf_size,
                            const double* bright_x, const double* bright_y, int n_bright,
                            dg_catalog_entry* entries, int capacity, int* n_out,
                            void* stream);
/*
 * dg_build_catalog_device for seeds the caller has already labelled: `labels`
 * (device int32, width x height) and n_labels must be what dg_label_device
 * returns for the same seeds -- the same score, threshold (threshold_sigma,
 * or threshold_sigma_dipole with return_negative), polarity and seed mask
 * (`mask` when exclude_bad_pixels, else none).  The rows are then those of
 * dg_build_catalog_device, without a second labelling.  The reference labels
 * inside delta::build_catalog (src/catalog.cpp:608-616); a caller that already
 * holds the labels of the same seeds (a pipeline that keeps the score's
 * 8-connected components as an output) passes them here.  labels is read
 * only.
 */
int dg_build_catalog_labelled_device(int width, int height, const float* score,
                                     const float* difference, const uint8_t* mask,
                                     const dg_catalog_params* params, int return_negative,
                                     const float* variance, const float* psf, int psf_size,
                                     const double* bright_x, const double* bright_y,
                                     int n_bright, const int32_t* labels, int n_labels,
                                     dg_catalog_entry* entries, int capacity, int* n_out,
                                     void* stream);

/*
 * Stamp selection knobs, field for field delta::DetectParams
 * (include/delta/detect.hpp:35-44); zero-valued knobs resolve as there.
 */
typedef struct dg_detect_params {
  int stamp_radius;       /* half-width of the stamp footprint (15) */
  double threshold_sigma; /* peak threshold above background, in sigma (5.0) */
  int max_stamps;         /* cap on returned stamps, brightest kept (200) */
  double saturation;      /* reject footprints with a pixel >= this (<= 0: off) */
  int isolation_radius;   /* min separation between stamps (0 -> 2 * stamp_radius) */
  int border;             /* edge exclusion (0 -> stamp_radius) */
  int fwhm_radius;        /* FWHM moment window (0 -> max(5, stamp_radius / 2)) */
  int match_radius;       /* science/reference cross-match tolerance, px (2) */
} dg_detect_params;

/*
 * Robust background: median and 1.4826 * MAD of every step-th pixel (flat
 * index, step = max(1, n / 2^18)) that is finite and unmasked; (0, 0) when
 * none is.  Replaces _core.estimate_background (src/bindings.cpp:231-236) and
 * delta::estimate_background (src/detect.cpp:84-113).  median and sigma are
 * host pointers in both forms.
 */
int dg_estimate_background(int width, int height, const float* data, const uint8_t* mask,
                           double* median, double* sigma);
int dg_estimate_background_device(int width, int height, const float* data,
                                  const uint8_t* mask, double* median, double* sigma,
                                  void* stream);

/*
 * Stamp detection: strict 3x3 local maxima above median + threshold_sigma *
 * sigma inside the border, whose footprint is on-frame, finite, unmasked and
 * unsaturated, taken brightest first while isolated from every stamp kept so
 * far, up to max_stamps.  Per stamp: x, y (int32), flux = v - median, snr,
 * fwhm (float64; NaN without positive flux), in selection order; equal-flux
 * candidates go by value, then raster index.  capacity (>= max_stamps) bounds
 * the outputs; *n_out gets the count.  The border must resolve to >= 1 pixel.
 * Replaces _core.detect_stamps (src/bindings.cpp:239-273) and
 * delta::detect_stamps (src/detect.cpp:159-241).  In the _device form the
 * image, mask and the five output arrays are device pointers.
 */
int dg_detect_stamps(int width, int height, const float* data, const uint8_t* mask,
                     const dg_detect_params* params, int32_t* x, int32_t* y, double* flux,
                     double* snr, double* fwhm, int capacity, int* n_out);
int dg_detect_stamps_device(int width, int height, const float* data, const uint8_t* mask,
                            const dg_detect_params* params, int32_t* x, int32_t* y,
                            double* flux, double* snr, double* fwhm, int capacity, int* n_out,
                            void* stream);

/*
 * Matched stamps of a science/reference pair.  Without a catalog (catalog_x
 * or catalog_y NULL): detection on both, then for each reference stamp, in
 * order, the nearest science stamp within match_radius (first of equal
 * distance).  With one: each of the first n_catalog positions whose footprint
 * passes in both images, FWHMs against each image's background median.
 * Outputs x, y, fwhm_science, fwhm_reference (capacity >= max_stamps, or >=
 * n_catalog), *n_out, medians[2] = upper medians of the finite FWHMs (NaN
 * when none), *direction = 0 ("reference": medians[1] <= medians[0]) or 1
 * ("science").  Replaces _core.select_stamps (src/bindings.cpp:277-320),
 * delta::select_stamps and select_stamps_from_catalog (src/detect.cpp:243-314).
 * In the _device form images, masks, the catalog and the four output arrays
 * are device pointers; n_out, medians and direction stay host pointers.
 */
int dg_select_stamps(int width_science, int height_science, const float* science,
                     const uint8_t* science_mask, int width_reference, int height_reference,
                     const float* reference, const uint8_t* reference_mask,
                     const int32_t* catalog_x, const int32_t* catalog_y, int n_catalog,
                     const dg_detect_params* params, int32_t* x, int32_t* y,
                     double* fwhm_science, double* fwhm_reference, int capacity, int* n_out,
                     double* medians, int* direction);
int dg_select_stamps_device(int width_science, int height_science, const float* science,
                            const uint8_t* science_mask, int width_reference,
                            int height_reference, const float* reference,
                            const uint8_t* reference_mask, const int32_t* catalog_x,
                            const int32_t* catalog_y, int n_catalog,
                            const dg_detect_params* params, int32_t* x, int32_t* y,
                            double* fwhm_science, double* fwhm_reference, int capacity,
                            int* n_out, double* medians, int* direction, void* stream);

/*
 * Differentiable constant-kernel model (no counterpart in the reference: it
 * is the reference's subtraction model, src/subtract.cpp:466-639, with one
 * spatial term, and a basis whose scale beta and sub-pixel shift (dx, dy)
 * are free parameters; src/basis.cpp:14-59 samples it at dx = dy = 0):
 *   K(u)  = sum_n a_n g_nx(ux - dx) g_ny(uy - dy),   u in [-radius, radius]^2
 *   D     = S - b - K (x) R                          (R zero outside the frame)
 *   chi^2 = sum w D^2 over pixels with w > 0        (weight NULL: w = 1)
 * params = [a_0 .. a_{nc-1}, b, beta, dx, dy] (nc + 4 float64, nc =
 * (n_max+1)(n_max+2)/2); out = [chi^2, dchi^2/da_0 .. dchi^2/da_{nc-1},
 * dchi^2/db, dchi^2/dbeta, dchi^2/ddx, dchi^2/ddy] (nc + 5 float64; only
 * out[0] when want_grad == 0).  The gradient comes from a hand-written
 * adjoint: the correlation C(u) = sum w D R(x - u), recomputed tile by tile
 * from the inputs, contracted with the basis and its Hermite derivatives.
 * n_max <= 10, 1 <= radius <= 15; beta must be > 0.
 * In the _device forms params and out are device pointers as well.
 */
int dg_kernel_chi2(int width, int height, const float* science, const float* reference,
                   const float* weight, const double* params, int n_max, int radius,
                   int want_grad, double* out);
int dg_kernel_chi2_device(int width, int height, const float* science, const float* reference,
                          const float* weight, const double* params, int n_max, int radius,
                          int want_grad, double* out, void* stream);

/* D of the same model (device pointers; params on the device). */
int dg_kernel_difference_device(int width, int height, const float* science,
                                const float* reference, const double* params, int n_max,
                                int radius, float* difference, void* stream);

/* Vector-Jacobian product of D: given grad_difference = dL/dD (device image),
 * out[1 .. nc+4] = dL/dparams (out[0] = 0), device pointers throughout. */
int dg_kernel_difference_vjp_device(int width, int height, const float* reference,
                                    const float* grad_difference, const double* params,
                                    int n_max, int radius, double* out, void* stream);

/*
 * Per-kernel timing (no counterpart in the reference; it replaces the
 * DELTA_TIMING stage timers of include/delta/timing.hpp:23-67 for device
 * work).  When enabled, each instrumented kernel launch records a HIP event
 * pair on its own stream; dg_profile_read sums the elapsed times of every
 * recorded launch of `name` ("model", "variance", "mask", "fields",
 * "select", ...).  dg_profile_reset frees the records.
 */
int dg_profile_enable(int on);
void dg_profile_reset(void);
int dg_profile_read(const char* name, int* count, double* total_ms);

#ifdef __cplusplus
}
#endif

#endif /* DELTAGPU_H */
