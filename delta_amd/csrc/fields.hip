// Thin-plate-spline coefficient fields on device (reference
// src/subtract.cpp:107-316): exact FP64 evaluation of
// f(x,y) = [U(r) Z | 1 u v](x,y) . theta_f on the coarse lattice (or per pixel
// for the exact small-frame mode), consumed by the model and variance kernels
// through the FP32 bilinear cell ramp (field_at).
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kMaxKnots = 128;

struct PointArgs {
  int k, nr, nf;
  double cx, cy, scale;
  const double* kn;
  const double* Z;
  const double* theta;
  int exact, w, h, ncx, ncy;
  const int* gx;
  const int* gy;
  double* lat;
  float* full;
  const double* pts;  // explicit (x, y) points when non-null: npts of them, values -> lat
  int npts;
};

// kFieldPts points per workgroup; the per-point work is spread over the
// workgroup in three phases through LDS: the kernel values U(r_j) of every
// (knot, point), the design entries [U Z | 1 u v] of every (column, point),
// then the field values of every (field, point).  Every sum keeps the serial
// order of one point's evaluation.
constexpr int kFieldPts = 64;
constexpr int kFieldThreads = 256;

__global__ __launch_bounds__(kFieldThreads) void k_field_points(PointArgs a) {
  extern __shared__ double fsm[];
  double* sig = fsm;                         // k x kFieldPts
  double* row = fsm + a.k * kFieldPts;       // k x kFieldPts
  __shared__ double ux_s[kFieldPts], uy_s[kFieldPts];
  const size_t npts = a.pts ? size_t(a.npts) : a.exact ? size_t(a.w) * a.h : size_t(a.ncx) * a.ncy;
  const size_t p0 = size_t(blockIdx.x) * kFieldPts;
  const int t = threadIdx.x;
  if (t < kFieldPts) {
    const size_t p = min(p0 + t, npts - 1);
    double px, py;
    if (a.pts) {
      px = a.pts[2 * p];
      py = a.pts[2 * p + 1];
    } else if (a.exact) {
      px = double(p % a.w);
      py = double(p / a.w);
    } else {
      px = a.gx[p % a.ncx];
      py = a.gy[p / a.ncx];
    }
    ux_s[t] = (px - a.cx) / a.scale;
    uy_s[t] = (py - a.cy) / a.scale;
  }
  __syncthreads();
  if (a.k == 1) {
    if (t < kFieldPts) row[t] = 1.0;
  } else {
    for (int e = t; e < a.k * kFieldPts; e += kFieldThreads) {
      const int j = e / kFieldPts, q = e % kFieldPts;
      const double dx = ux_s[q] - a.kn[2 * j], dy = uy_s[q] - a.kn[2 * j + 1];
      const double r2 = dx * dx + dy * dy;
      sig[e] = r2 > 0.0 ? 0.5 * r2 * log(r2) : 0.0;
    }
    __syncthreads();
    for (int e = t; e < a.k * kFieldPts; e += kFieldThreads) {
      const int m = e / kFieldPts, q = e % kFieldPts;
      double s;
      if (m < a.nr) {
        s = 0.0;
        for (int j = 0; j < a.k; ++j) s += sig[j * kFieldPts + q] * a.Z[j * a.nr + m];
      } else {
        s = m == a.nr ? 1.0 : (m == a.nr + 1 ? ux_s[q] : uy_s[q]);
      }
      row[e] = s;
    }
  }
  __syncthreads();
  for (int e = t; e < a.nf * kFieldPts; e += kFieldThreads) {
    const int f = e / kFieldPts, q = e % kFieldPts;
    const size_t p = p0 + q;
    if (p >= npts) continue;
    double s = 0.0;
    for (int m = 0; m < a.k; ++m) s += row[m * kFieldPts + q] * a.theta[size_t(f) * a.k + m];
    if (a.exact && !a.pts)
      a.full[size_t(f) * npts + p] = static_cast<float>(s);
    else
      a.lat[size_t(f) * npts + p] = s;
  }
}

void launch_field_points(const PointArgs& a, size_t npts, hipStream_t s) {
  if (npts == 0) return;
  const int lds = int(sizeof(double) * 2 * a.k * kFieldPts);
  set_max_lds(reinterpret_cast<const void*>(&k_field_points), lds);
  hipLaunchKernelGGL(k_field_points, dim3(unsigned((npts + kFieldPts - 1) / kFieldPts)),
                     dim3(kFieldThreads), lds, s, a);
}

}  // namespace

size_t fields_scratch_bytes(const Tps& tps, int nc, int w, int h) {
  const LatticeAxes L = lattice_axes(tps, w, h);
  const int k = tps.k, nf = nc + 1;
  // (an upload group takes at least one element per buffer: an empty Z too)
  size_t b = Arena::round(sizeof(double) * 2 * k) +
             Arena::round(sizeof(double) * std::max<size_t>(tps.Z.size(), 1)) +
             Arena::round(sizeof(double) * size_t(nf) * k);
  if (L.exact) return b + Arena::round(sizeof(float) * size_t(nf) * w * h);
  b += Arena::round(sizeof(int) * L.ncx) + Arena::round(sizeof(int) * L.ncy);
  b += Arena::round(sizeof(double) * size_t(nf) * L.ncx * L.ncy);
  b += 2 * Arena::round(sizeof(int) * w) + 2 * Arena::round(sizeof(float) * h) +
       Arena::round(sizeof(float) * w) + Arena::round(sizeof(int) * h);
  return b;
}

FieldsBuild build_fields(const Tps& tps, const double* theta, int nc, int w, int h, Arena& ar,
                         hipStream_t s, UploadGroup* pending) {
  if (tps.k > kMaxKnots) invalid("at most 128 knots are supported");
  FieldsBuild fb;
  fb.axes = lattice_axes(tps, w, h);
  const LatticeAxes& L = fb.axes;
  const int k = tps.k, nf = nc + 1;
  // every small upload of this build (and the caller's pending ones) in one
  // copy from pinned staging, committed before the first launch
  UploadGroup own(ar);
  UploadGroup& up = pending ? *pending : own;
  double* d_kn = up.add(tps.kn.data(), size_t(2) * k);
  double* d_Z = up.add(tps.Z.data(), tps.Z.size());
  double* d_th = up.add(theta, size_t(nf) * k);
  PointArgs a{};
  a.k = k;
  a.nr = std::max(k - 3, 0);
  a.nf = nf;
  a.cx = tps.cx;
  a.cy = tps.cy;
  a.scale = tps.scale;
  a.kn = d_kn;
  a.Z = d_Z;
  a.theta = d_th;
  a.w = w;
  a.h = h;
  DevFields& F = fb.dev;
  F = DevFields{};
  F.nf = nf;
  if (L.exact) {
    float* full = ar.take<float>(size_t(nf) * w * h);
    a.exact = 1;
    a.full = full;
    F.exact = 1;
    F.full = full;
    up.commit(PinnedStage::local(), s);
    const size_t npts = size_t(w) * h;
    launch_field_points(a, npts, s);
    DG_LAUNCH_CHECK();
    return fb;
  }
  int* d_gx = up.add(L.gx.data(), size_t(L.ncx));
  int* d_gy = up.add(L.gy.data(), size_t(L.ncy));
  int* d_cellx = up.add(L.cellx.data(), size_t(w));
  float* d_wx = up.add(L.wx.data(), size_t(w));
  int* d_celly = up.add(L.celly.data(), size_t(h));
  float* d_wy = up.add(L.wy.data(), size_t(h));
  double* lat = ar.take<double>(size_t(nf) * L.ncx * L.ncy);  // written on device
  up.commit(PinnedStage::local(), s);
  a.exact = 0;
  a.ncx = L.ncx;
  a.ncy = L.ncy;
  a.gx = d_gx;
  a.gy = d_gy;
  a.lat = lat;
  const size_t npts = size_t(L.ncx) * L.ncy;
  launch_field_points(a, npts, s);
  DG_LAUNCH_CHECK();
  F.exact = 0;
  F.ncx = L.ncx;
  F.ncy = L.ncy;
  F.lat = lat;
  F.cellx = d_cellx;
  F.wx = d_wx;
  F.celly = d_celly;
  F.wy = d_wy;
  return fb;
}

// Field values a_f(x, y) = [U(r) Z | 1 u v](x, y) . theta_f at explicit
// (x, y) points (device, npts x 2), FP64: out[f * npts + p].
void eval_fields_points(const Tps& tps, const double* theta, int nf, const double* d_pts,
                        int npts, double* d_out, Arena& ar, hipStream_t s, PinnedStage* stage) {
  if (tps.k > kMaxKnots) invalid("at most 128 knots are supported");
  const int k = tps.k;
  double* d_kn = ar.take<double>(2 * k);
  double* d_Z = ar.take<double>(std::max<size_t>(tps.Z.size(), 1));
  double* d_th = ar.take<double>(size_t(nf) * k);
  const auto up = [&](double* d, const double* src, size_t n) {
    if (stage)
      stage->h2d(d, src, n, s);
    else
      h2d(d, src, n, s);
  };
  up(d_kn, tps.kn.data(), 2 * size_t(k));
  if (!tps.Z.empty()) up(d_Z, tps.Z.data(), tps.Z.size());
  up(d_th, theta, size_t(nf) * k);
  PointArgs a{};
  a.k = k;
  a.nr = std::max(k - 3, 0);
  a.nf = nf;
  a.cx = tps.cx;
  a.cy = tps.cy;
  a.scale = tps.scale;
  a.kn = d_kn;
  a.Z = d_Z;
  a.theta = d_th;
  a.pts = d_pts;
  a.npts = npts;
  a.lat = d_out;
  launch_field_points(a, npts, s);
  DG_LAUNCH_CHECK();
}

size_t eval_fields_points_bytes(const Tps& tps, int nf) {
  return Arena::round(sizeof(double) * 2 * tps.k) +
         Arena::round(sizeof(double) * std::max<size_t>(tps.Z.size(), 1)) +
         Arena::round(sizeof(double) * size_t(nf) * tps.k);
}

// ------------------------------------------------------------ photometric scale
namespace {

// sum_c (float)S_c * a_c(x, y) per pixel, accumulated in float in component
// order with the multiply-add fused (src/fit.cpp:513-530).
__global__ __launch_bounds__(256) void k_photometric(DevFields F, int w, int h, int nc,
                                                     const float* __restrict__ sums,
                                                     float* __restrict__ out) {
  const size_t n = size_t(w) * h;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const int x = static_cast<int>(i % size_t(w)), y = static_cast<int>(i / size_t(w));
    float acc = 0.0f;
    for (int c = 0; c < nc; ++c) acc = fmaf(sums[c], field_at(F, c, x, y, w, h), acc);
    out[i] = acc;
  }
}

// sum_c a_c(p) * S_c at explicit points, FP64 (src/fit.cpp:532-545).
__global__ __launch_bounds__(256) void k_photometric_at(const double* __restrict__ fields, int m,
                                                        int nc, const double* __restrict__ sums,
                                                        double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  double acc = 0.0;
  for (int c = 0; c < nc; ++c) acc = fma(fields[size_t(c) * m + p], sums[c], acc);
  out[p] = acc;
}

}  // namespace

void run_photometric_scale(const double* knots, int n_knots, const double* theta, int n_theta,
                           const double* sums, int nc, int w, int h, float* out, hipStream_t s) {
  if (w <= 0 || h <= 0 || nc < 0) invalid("photometric_scale: empty frame");
  const Tps tps(knots, n_knots);
  if (n_theta != n_knots * (nc + 1)) runtime("photometric_scale: theta size mismatch");
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(fields_scratch_bytes(tps, nc, w, h) + Arena::round(sizeof(float) * (nc + 1)));
  std::vector<float> sf(nc + 1);
  for (int c = 0; c < nc; ++c) sf[c] = static_cast<float>(sums[c]);
  float* d_s = ar.take<float>(nc + 1);
  h2d(d_s, sf.data(), nc + 1, s);
  FieldsBuild fb = build_fields(tps, theta, nc, w, h, ar, s);
  const size_t n = size_t(w) * h;
  hipLaunchKernelGGL(k_photometric, dim3(unsigned(std::min<size_t>(8192, (n + 255) / 256))),
                     dim3(256), 0, s, fb.dev, w, h, nc, static_cast<const float*>(d_s), out);
  DG_LAUNCH_CHECK();
  DG_HIP(hipStreamSynchronize(s));
}

void run_photometric_scale_at(const double* knots, int n_knots, const double* theta,
                              int n_theta, const double* sums, int nc, const double* points,
                              int m, double* out) {
  const Tps tps(knots, n_knots);
  if (n_theta != n_knots * (nc + 1)) runtime("photometric_scale: theta size mismatch");
  if (m <= 0) return;
  OwnedStream st;
  hipStream_t s = st.get();
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(eval_fields_points_bytes(tps, nc + 1) + Arena::round(sizeof(double) * 2 * m) +
             Arena::round(sizeof(double) * size_t(nc + 1) * m) +
             Arena::round(sizeof(double) * (nc + 1)) + Arena::round(sizeof(double) * m));
  double* d_pts = ar.take<double>(2 * size_t(m));
  double* d_f = ar.take<double>(size_t(nc + 1) * m);
  double* d_s = ar.take<double>(nc + 1);
  double* d_out = ar.take<double>(m);
  h2d(d_pts, points, 2 * size_t(m), s);
  if (nc > 0) h2d(d_s, sums, nc, s);
  eval_fields_points(tps, theta, nc + 1, d_pts, m, d_f, ar, s);
  hipLaunchKernelGGL(k_photometric_at, dim3((m + 255) / 256), dim3(256), 0, s,
                     static_cast<const double*>(d_f), m, nc, static_cast<const double*>(d_s),
                     d_out);
  DG_LAUNCH_CHECK();
  d2h(out, static_cast<const double*>(d_out), m, s);
  DG_HIP(hipStreamSynchronize(s));
}

}  // namespace dg
