// C entry points onto the reference's own Eigen-free code (basis.cpp,
// convolve.cpp and catalog.cpp's anonymous-namespace CCL), for pinning the
// oracle in this container.  Built by `make ref` into oracle/_ref/; never
// shipped, never loaded by the product.
#define DELTA_STR2(x) #x
#define DELTA_STR(x) DELTA_STR2(x)
#include DELTA_STR(DELTA_REF_SRC/catalog.cpp)

#include <cstring>

#include "delta/basis.hpp"
#include "delta/convolve.hpp"

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
}
