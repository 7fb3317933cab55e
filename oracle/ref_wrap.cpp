// C entry points onto the reference's own Eigen-free code (basis.cpp,
// convolve.cpp, detect.cpp and catalog.cpp's anonymous-namespace CCL), for pinning the
// oracle in this container.  Built by `make ref` into oracle/_ref/; never
// shipped, never loaded by the product.
#define DELTA_STR2(x) #x
#define DELTA_STR(x) DELTA_STR2(x)
#include DELTA_STR(DELTA_REF_SRC/catalog.cpp)

#include <cstring>

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
}
