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
}  // namespace dg

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

}  // extern "C"
