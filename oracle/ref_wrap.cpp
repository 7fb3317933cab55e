// C entry points onto the reference's own Eigen-free code (basis.cpp,
// convolve.cpp, detect.cpp and catalog.cpp's anonymous-namespace CCL), for pinning the
// oracle in this container.  Built by `make ref` into oracle/_ref/; never
// shipped, never loaded by the product.
#define DELTA_STR2(x) #x
#define DELTA_STR(x) DELTA_STR2(x)
#include DELTA_STR(DELTA_REF_SRC/catalog.cpp)

#include <cstring>
#include <omp.h>

#include "delta/basis.hpp"
#include "delta/convolve.hpp"
#include "delta/detect.hpp"

namespace {
delta::ImageF ref_image(const float* img, const unsigned char* mask, int w, int h) {
  delta::ImageF im(w, h, img);
  if (mask) im.mask().assign(mask, mask + static_cast<size_t>(w) * h);
  return im;
}
// iparams: stamp_radius, max_stamps, isolation_radius, border, fwhm_radius,
// match_radius; dparams: threshold_sigma, saturation.
delta::DetectParams ref_params(const int* ip, const double* dp) {
  delta::DetectParams p;
  p.stamp_radius = ip[0];
  p.max_stamps = ip[1];
  p.isolation_radius = ip[2];
  p.border = ip[3];
  p.fwhm_radius = ip[4];
  p.match_radius = ip[5];
  p.threshold_sigma = dp[0];
  p.saturation = dp[1];
  return p;
}
}  // namespace

extern "C" {

void ref_basis1d(double beta, int n_max, int radius, double* out) {
  delta::GaussHermiteBasis b(beta, n_max, radius);
  for (int n = 0; n <= n_max; ++n) {
    const auto& v = b.basis1d(n);
    std::memcpy(out + static_cast<size_t>(n) * b.ksize(), v.data(), sizeof(double) * v.size());
  }
}

int ref_default_radius(double beta, int n_max) {
  return delta::GaussHermiteBasis::default_radius(beta, n_max);
}

void ref_component_sums(double beta, int n_max, int radius, double* out) {
  delta::GaussHermiteBasis b(beta, n_max, radius);
  const auto s = b.component_sums();
  std::memcpy(out, s.data(), sizeof(double) * s.size());
}

// B_n = phi_n (x) image for all components, nc x h x w.
void ref_basis_convolve(const float* img, int w, int h, double beta, int n_max,
                        int radius, float* out) {
  delta::GaussHermiteBasis b(beta, n_max, radius);
  delta::ImageF im(w, h, img);
  const auto bn = delta::basis_convolve(im, b);
  for (size_t c = 0; c < bn.size(); ++c)
    std::memcpy(out + c * static_cast<size_t>(w) * h, bn[c].data(), sizeof(float) * w * h);
}

// 8-connected labels of the seed predicate (score >= thr, finite, mask good).
int ref_label(const float* score, const unsigned char* mask, int w, int h,
              double threshold, int polarity, int* labels) {
  delta::ImageF s(w, h, score);
  if (mask) {
    s.allocate_mask();
    std::memcpy(s.mask().data(), mask, static_cast<size_t>(w) * h);
  }
  const auto seed = delta::seed_mask(s, threshold, polarity, mask != nullptr);
  const auto res = delta::label_components(seed, w, h);
  std::memcpy(labels, res.labels.data(), sizeof(int) * static_cast<size_t>(w) * h);
  return res.n_components;
}

void ref_estimate_background(const float* img, const unsigned char* mask, int w, int h,
                             double* out) {
  const auto bg = delta::estimate_background(ref_image(img, mask, w, h));
  out[0] = bg.median;
  out[1] = bg.sigma;
}

double ref_estimate_fwhm(const float* img, int w, int h, int x, int y, int radius, double bg) {
  return delta::estimate_fwhm(ref_image(img, nullptr, w, h), x, y, radius, bg);
}

// Stamps in selection order; returns the full count (cap bounds the copy).
int ref_detect_stamps(const float* img, const unsigned char* mask, int w, int h, const int* ip,
                      const double* dp, int cap, int* xs, int* ys, double* flux, double* snr,
                      double* fwhm) {
  const auto st = delta::detect_stamps(ref_image(img, mask, w, h), ref_params(ip, dp));
  const int n = static_cast<int>(st.size());
  for (int i = 0; i < n && i < cap; ++i) {
    xs[i] = st[i].x;
    ys[i] = st[i].y;
    flux[i] = st[i].flux;
    snr[i] = st[i].snr;
    fwhm[i] = st[i].fwhm;
  }
  return n;
}

// direction: 0 = reference, 1 = science.  cat_x == nullptr: detection.
int ref_select_stamps(const float* sci, const unsigned char* sci_mask, int ws, int hs,
                      const float* ref, const unsigned char* ref_mask, int wr, int hr,
                      const int* cat_x, const int* cat_y, int n_cat, const int* ip,
                      const double* dp, int cap, int* xs, int* ys, double* fsci, double* fref,
                      double* medians, int* direction) {
  const auto s = ref_image(sci, sci_mask, ws, hs);
  const auto r = ref_image(ref, ref_mask, wr, hr);
  const auto p = ref_params(ip, dp);
  delta::StampSelection sel;
  if (cat_x) {
    const std::vector<int> cx(cat_x, cat_x + n_cat), cy(cat_y, cat_y + n_cat);
    sel = delta::select_stamps_from_catalog(s, r, cx, cy, p);
  } else {
    sel = delta::select_stamps(s, r, p);
  }
  const int n = static_cast<int>(sel.x.size());
  for (int i = 0; i < n && i < cap; ++i) {
    xs[i] = sel.x[i];
    ys[i] = sel.y[i];
    fsci[i] = sel.fwhm_science[i];
    fref[i] = sel.fwhm_reference[i];
  }
  medians[0] = sel.median_fwhm_science;
  medians[1] = sel.median_fwhm_reference;
  *direction = sel.direction == delta::ConvolveDirection::kConvolveReference ? 0 : 1;
  return n;
}

// The same layout as the oracle's or_catalog_entry.
struct RefCatalogEntry {
  double x, y, peak_snr, flux, expected_n_pix, fwhm_ratio;
  double psf_chi2, psf_amplitude, psf_amplitude_err, bright_star_dist;
  int32_t peak_x, peak_y, n_pix, psf_chi2_dof;
  uint8_t fwhm_consistent, mask_flags, is_dipole, shape_consistent, near_bright_star, quality;
  uint8_t pad[2];
};

// build_catalog on one OpenMP thread, so each component's pixels reach
// summarize_component in raster order.  dp: threshold_sigma,
// threshold_sigma_dipole, expected_fwhm, fwhm_tolerance_lo, fwhm_tolerance_hi,
// shape_chi2_tolerance, bright_star_radius; ip: aperture_radius,
// exclude_bad_pixels, fit_psf_shape, max_shape_fit.  Returns the entry count
// (cap bounds the copy).
int ref_build_catalog(const float* score, const float* diff, const unsigned char* mask, int w,
                      int h, const double* dp, const int* ip, int return_negative,
                      const float* var, const float* psf, int ps, const double* bx,
                      const double* by, int nb, RefCatalogEntry* out, int cap) {
  delta::ImageF s(w, h, score), d(w, h, diff), v;
  if (mask) s.mask().assign(mask, mask + static_cast<size_t>(w) * h);
  delta::CatalogParams p;
  p.threshold_sigma = dp[0];
  p.threshold_sigma_dipole = dp[1];
  p.expected_fwhm = dp[2];
  p.fwhm_tolerance_lo = dp[3];
  p.fwhm_tolerance_hi = dp[4];
  p.shape_chi2_tolerance = dp[5];
  p.bright_star_radius = dp[6];
  p.aperture_radius = ip[0];
  p.exclude_bad_pixels = ip[1] != 0;
  p.fit_psf_shape = ip[2] != 0;
  p.max_shape_fit = ip[3];
  delta::CatalogAux aux;
  std::vector<float> pv;
  std::vector<double> bxv, byv;
  if (var) {
    v = delta::ImageF(w, h, var);
    aux.variance = &v;
  }
  if (psf) {
    pv.assign(psf, psf + static_cast<size_t>(ps) * ps);
    aux.psf = &pv;
    aux.psf_size = ps;
  }
  if (bx) {
    bxv.assign(bx, bx + nb);
    byv.assign(by, by + nb);
    aux.bright_x = &bxv;
    aux.bright_y = &byv;
  }
  const int threads = omp_get_max_threads();
  omp_set_num_threads(1);
  const auto e = delta::build_catalog(s, d, p, return_negative != 0, aux);
  omp_set_num_threads(threads);
  const int n = static_cast<int>(e.size());
  for (int i = 0; i < n && i < cap; ++i) {
    RefCatalogEntry& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.x = e[i].x;
    o.y = e[i].y;
    o.peak_snr = e[i].peak_snr;
    o.flux = e[i].flux;
    o.expected_n_pix = e[i].expected_n_pix;
    o.fwhm_ratio = e[i].fwhm_ratio;
    o.psf_chi2 = e[i].psf_chi2;
    o.psf_amplitude = e[i].psf_amplitude;
    o.psf_amplitude_err = e[i].psf_amplitude_err;
    o.bright_star_dist = e[i].bright_star_dist;
    o.peak_x = e[i].peak_x;
    o.peak_y = e[i].peak_y;
    o.n_pix = e[i].n_pix;
    o.psf_chi2_dof = e[i].psf_chi2_dof;
    o.fwhm_consistent = e[i].fwhm_consistent;
    o.mask_flags = e[i].mask_flags;
    o.is_dipole = e[i].is_dipole;
    o.shape_consistent = e[i].shape_consistent;
    o.near_bright_star = e[i].near_bright_star;
    o.quality = e[i].quality;
  }
  return n;
}
}
