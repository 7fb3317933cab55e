// Host-side setup of libdeltagpu: error text, the scratch arena pool, the
// Gauss-Hermite basis, the thin-plate spline and the field lattice axes.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <memory>

#include "dg_internal.hpp"

namespace dg {

// ---------------------------------------------------------------- errors
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& m) { g_last_error = m; }
const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------- arenas
Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

void Arena::reserve(size_t bytes) {
  used_ = 0;
  if (bytes <= cap_) return;
  if (base_) DG_HIP(hipFree(base_));
  base_ = nullptr;
  cap_ = 0;
  DG_HIP(hipMalloc(reinterpret_cast<void**>(&base_), bytes));
  cap_ = bytes;
}

void set_max_lds(const void* kernel, int bytes) {
  static std::mutex mu;
  static std::vector<std::pair<const void*, int>> done;
  std::lock_guard<std::mutex> lk(mu);
  for (auto& e : done)
    if (e.first == kernel) {
      if (e.second >= bytes) return;
      DG_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
      e.second = bytes;
      return;
    }
  DG_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  done.emplace_back(kernel, bytes);
}

PinnedStage& PinnedStage::local() {
  thread_local PinnedStage stage;
  return stage;
}

char* PinnedStage::grab(size_t bytes) {
  constexpr size_t kCap = size_t(4) << 20;
  if (!base_) {
    if (hipHostMalloc(reinterpret_cast<void**>(&base_), kCap, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();  // leave no sticky error for a later launch check
      base_ = nullptr;
      return nullptr;
    }
    cap_ = kCap;
  }
  const size_t need = (bytes + 255) & ~size_t(255);
  if (used_ + need > cap_) return nullptr;
  char* p = base_ + used_;
  used_ += need;
  return p;
}

namespace {
std::mutex g_pool_mu;
std::vector<std::unique_ptr<Arena>> g_free;
}  // namespace

ArenaLease::ArenaLease() {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (g_free.empty()) {
    a_ = new Arena();
  } else {
    a_ = g_free.back().release();
    g_free.pop_back();
  }
}

ArenaLease::~ArenaLease() {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  g_free.emplace_back(a_);
}

// ---------------------------------------------------------------- basis
int GhBasis::default_radius(double beta, int n_max) {
  // beta * sqrt(n_max + 1) first, then the factor 3, as src/basis.cpp:9-12
  // associates it: (3 beta) sqrt(n+1) lands on the other side of an integer
  // for some beta, and ceil then gives another radius
  const double scale = beta * std::sqrt(double(n_max + 1));
  const int r = static_cast<int>(std::ceil(3.0 * scale));
  return std::max(r, 3);
}

GhBasis::GhBasis(double beta_, int n_max_, int radius_)
    : beta(beta_), n_max(n_max_), radius(radius_) {
  if (!(beta > 0.0)) invalid("beta must be > 0");
  if (n_max < 0) invalid("n_max must be >= 0");
  if (radius < 1) invalid("radius must be >= 1");
  ks = 2 * radius + 1;
  nc = (n_max + 1) * (n_max + 2) / 2;
  g.assign(static_cast<size_t>(n_max + 1) * ks, 0.0);
  // log of the per-order normalisation [2^n n! sqrt(pi) beta]^(-1/2).  The
  // rounding follows the reference's compiled basis.cpp (g++ contracts
  // n ln2 + lgamma and 2t H_n - 2n H_{n-1} into one fused multiply-add each,
  // nothing else): the explicit fma and the volatile products pin that form
  // under this compiler's own contraction, so the samples are bit-equal.
  volatile double half_ln_pi = 0.5 * std::log(std::acos(-1.0));
  std::vector<double> lognorm(n_max + 1);
  for (int n = 0; n <= n_max; ++n)
    lognorm[n] = -0.5 * ((std::fma(double(n), std::log(2.0), std::lgamma(n + 1.0)) + half_ln_pi) +
                         std::log(beta));
  for (int i = 0; i < ks; ++i) {
    const double t = double(i - radius) / beta;
    const double e = std::exp(-0.5 * t * t);
    // physicists' Hermite recurrence H_{n+1} = 2t H_n - 2n H_{n-1}
    double hprev = 1.0, hcur = 2.0 * t;
    g[i] = std::exp(lognorm[0]) * hprev * e;
    if (n_max >= 1) g[ks + i] = std::exp(lognorm[1]) * hcur * e;
    for (int n = 2; n <= n_max; ++n) {
      volatile double lower = 2.0 * (n - 1) * hprev;
      const double hnext = std::fma(2.0 * t, hcur, -lower);
      g[size_t(n) * ks + i] = std::exp(lognorm[n]) * hnext * e;
      hprev = hcur;
      hcur = hnext;
    }
  }
  for (int tot = 0; tot <= n_max; ++tot)
    for (int ny = 0; ny <= tot; ++ny) order.emplace_back(tot - ny, ny);
}

std::vector<float> GhBasis::taps_f32() const { return std::vector<float>(g.begin(), g.end()); }

std::vector<float> GhBasis::footprints_f32() const {
  std::vector<float> out(size_t(nc) * ks * ks);
  for (int c = 0; c < nc; ++c) {
    const double* gx = g.data() + size_t(order[c].first) * ks;
    const double* gy = g.data() + size_t(order[c].second) * ks;
    float* o = out.data() + size_t(c) * ks * ks;
    for (int iy = 0; iy < ks; ++iy)
      for (int ix = 0; ix < ks; ++ix) o[iy * ks + ix] = static_cast<float>(gy[iy] * gx[ix]);
  }
  return out;
}

std::vector<double> GhBasis::component_sums() const {
  std::vector<double> s1(n_max + 1, 0.0), out;
  for (int n = 0; n <= n_max; ++n)
    for (int i = 0; i < ks; ++i) s1[n] += g[size_t(n) * ks + i];
  for (auto [nx, ny] : order) out.push_back(s1[nx] * s1[ny]);
  return out;
}

// ---------------------------------------------------------------- thin-plate spline
namespace {
double tps_green(double r2) { return r2 > 0.0 ? 0.5 * r2 * std::log(r2) : 0.0; }
}  // namespace

Tps::Tps(const double* knots, int k_) : k(k_) {
  if (k != 1 && k < 3) invalid("need at least 3 knots");
  kn.assign(size_t(k) * 2, 0.0);
  pen.assign(size_t(k) * k, 0.0);
  if (k == 1) return;  // constant field
  double sx = 0, sy = 0;
  for (int i = 0; i < k; ++i) {
    sx += knots[2 * i];
    sy += knots[2 * i + 1];
  }
  cx = sx / k;
  cy = sy / k;
  double ex = 0, ey = 0;
  for (int i = 0; i < k; ++i) {
    ex = std::max(ex, std::fabs(knots[2 * i] - cx));
    ey = std::max(ey, std::fabs(knots[2 * i + 1] - cy));
  }
  scale = std::max({ex, ey, 1e-12});
  for (int i = 0; i < k; ++i) {
    kn[2 * i] = (knots[2 * i] - cx) / scale;
    kn[2 * i + 1] = (knots[2 * i + 1] - cy) / scale;
  }
  // Null space of the affine trend [1 u v]^T: the trailing k-3 columns of the
  // complete Q of its Householder QR (reflector beta = -sign(alpha)||x||).
  std::vector<double> a(size_t(k) * 3), v(size_t(k) * 3, 0.0);
  double tau[3] = {0, 0, 0};
  for (int i = 0; i < k; ++i) {
    a[3 * i] = 1.0;
    a[3 * i + 1] = kn[2 * i];
    a[3 * i + 2] = kn[2 * i + 1];
  }
  for (int j = 0; j < 3; ++j) {
    const double alpha = a[3 * j + j];
    double tail = 0;
    for (int i = j + 1; i < k; ++i) tail += a[3 * i + j] * a[3 * i + j];
    v[3 * j + j] = 1.0;
    if (tail <= DBL_MIN) continue;
    double beta = std::sqrt(alpha * alpha + tail);
    if (alpha >= 0) beta = -beta;
    for (int i = j + 1; i < k; ++i) v[3 * i + j] = a[3 * i + j] / (alpha - beta);
    tau[j] = (beta - alpha) / beta;
    a[3 * j + j] = beta;
    for (int c = j + 1; c < 3; ++c) {
      double s = 0;
      for (int i = j; i < k; ++i) s += v[3 * i + j] * a[3 * i + c];
      s *= tau[j];
      for (int i = j; i < k; ++i) a[3 * i + c] -= s * v[3 * i + j];
    }
  }
  std::vector<double> q(size_t(k) * k, 0.0);
  for (int i = 0; i < k; ++i) q[size_t(i) * k + i] = 1.0;
  for (int j = 2; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double s = 0;
      for (int i = j; i < k; ++i) s += v[3 * i + j] * q[size_t(i) * k + c];
      s *= tau[j];
      for (int i = j; i < k; ++i) q[size_t(i) * k + c] -= s * v[3 * i + j];
    }
  }
  const int nr = k - 3;
  Z.assign(size_t(k) * std::max(nr, 1), 0.0);
  for (int i = 0; i < k; ++i)
    for (int m = 0; m < nr; ++m) Z[size_t(i) * nr + m] = q[size_t(i) * k + 3 + m];
  if (nr > 0) {
    std::vector<double> e(size_t(k) * k), ez(size_t(k) * nr);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) {
        const double dx = kn[2 * i] - kn[2 * j], dy = kn[2 * i + 1] - kn[2 * j + 1];
        e[size_t(i) * k + j] = tps_green(dx * dx + dy * dy);
      }
    for (int i = 0; i < k; ++i)
      for (int m = 0; m < nr; ++m) {
        double s = 0;
        for (int j = 0; j < k; ++j) s += e[size_t(i) * k + j] * Z[size_t(j) * nr + m];
        ez[size_t(i) * nr + m] = s;
      }
    for (int p = 0; p < nr; ++p)
      for (int r = 0; r < nr; ++r) {
        double s = 0;
        for (int i = 0; i < k; ++i) s += Z[size_t(i) * nr + p] * ez[size_t(i) * nr + r];
        pen[size_t(p) * k + r] = s;
      }
    for (int p = 0; p < k; ++p)
      for (int r = p + 1; r < k; ++r) {
        const double m = 0.5 * (pen[size_t(p) * k + r] + pen[size_t(r) * k + p]);
        pen[size_t(p) * k + r] = pen[size_t(r) * k + p] = m;
      }
  }
}

void Tps::design_row(double px, double py, double* row) const {
  if (k == 1) {
    row[0] = 1.0;
    return;
  }
  const int nr = k - 3;
  const double ux = (px - cx) / scale, uy = (py - cy) / scale;
  std::vector<double> sig(k);
  for (int j = 0; j < k; ++j) {
    const double dx = ux - kn[2 * j], dy = uy - kn[2 * j + 1];
    sig[j] = tps_green(dx * dx + dy * dy);
  }
  for (int m = 0; m < nr; ++m) {
    double s = 0;
    for (int j = 0; j < k; ++j) s += sig[j] * Z[size_t(j) * nr + m];
    row[m] = s;
  }
  row[nr] = 1.0;
  row[nr + 1] = ux;
  row[nr + 2] = uy;
}

double Tps::min_spacing() const {
  if (k < 2) return 0.0;
  double best = INFINITY;
  for (int i = 0; i < k; ++i)
    for (int j = i + 1; j < k; ++j) {
      const double dx = kn[2 * i] - kn[2 * j], dy = kn[2 * i + 1] - kn[2 * j + 1];
      best = std::min(best, dx * dx + dy * dy);
    }
  return std::isfinite(best) ? std::sqrt(best) * scale : 0.0;
}

// ---------------------------------------------------------------- lattice axes
namespace {
std::vector<int> axis(int len, int stride) {
  std::vector<int> g;
  for (int p = 0; p < len - 1; p += stride) g.push_back(p);
  g.push_back(len - 1);
  return g;
}
}  // namespace

int device_cu_count() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  DG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  if (static_cast<int>(cache.size()) <= dev) cache.resize(dev + 1, 0);
  if (cache[dev] == 0) {
    int n = 0;
    DG_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    cache[dev] = std::max(n, 1);
  }
  return cache[dev];
}

LatticeAxes lattice_axes(const Tps& tps, int w, int h) {
  LatticeAxes L;
  int stride = 1;
  const double sp = tps.min_spacing();
  if (sp > 0.0) stride = std::clamp(static_cast<int>(sp / 16.0), 1, 64);
  L.stride = stride;
  if (tps.k == 1) {
    // Constant field: a 2x2 lattice of equal values reproduces it exactly.
    L.gx = {0, w - 1};
    L.gy = {0, h - 1};
    stride = std::max(w, h) + 1;
  } else if (stride < 4 || w <= 2 * stride || h <= 2 * stride) {
    L.exact = true;
    return L;
  } else {
    L.gx = axis(w, stride);
    L.gy = axis(h, stride);
  }
  L.ncx = static_cast<int>(L.gx.size());
  L.ncy = static_cast<int>(L.gy.size());
  L.cellx.resize(w);
  L.wx.resize(w);
  for (int x = 0; x < w; ++x) {
    const int c0 = std::min(x / stride, L.ncx - 2);
    const int x0 = L.gx[c0], x1 = L.gx[c0 + 1];
    L.cellx[x] = c0;
    L.wx[x] = x1 > x0 ? static_cast<float>(x - x0) / static_cast<float>(x1 - x0) : 0.0f;
  }
  L.celly.resize(h);
  L.wy.resize(h);
  for (int y = 0; y < h; ++y) {
    const int c0 = std::min(y / stride, L.ncy - 2);
    const int y0 = L.gy[c0], y1 = L.gy[c0 + 1];
    L.celly[y] = c0;
    L.wy[y] = y1 > y0 ? static_cast<float>(y - y0) / static_cast<float>(y1 - y0) : 0.0f;
  }
  return L;
}

}  // namespace dg

// ---------------------------------------------------------------- profiling
namespace dg {
namespace {
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRec> g_prof;
thread_local std::vector<size_t> t_open;
}  // namespace

bool prof_enabled() { return g_prof_on; }

void prof_begin(const char* name, hipStream_t s) {
  ProfRec r{name, nullptr, nullptr};
  DG_HIP(hipEventCreate(&r.a));
  DG_HIP(hipEventCreate(&r.b));
  DG_HIP(hipEventRecord(r.a, s));
  std::lock_guard<std::mutex> lk(g_prof_mu);
  g_prof.push_back(r);
  t_open.push_back(g_prof.size() - 1);
}

void prof_end(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (t_open.empty()) return;
  const size_t i = t_open.back();
  t_open.pop_back();
  (void)hipEventRecord(g_prof[i].b, s);
}
}  // namespace dg

extern "C" {
int dg_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(dg::g_prof_mu);
  dg::g_prof_on = on != 0;
  return 0;
}

void dg_profile_reset(void) {
  std::lock_guard<std::mutex> lk(dg::g_prof_mu);
  for (auto& r : dg::g_prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  dg::g_prof.clear();
}

int dg_profile_read(const char* name, int* count, double* total_ms) {
  std::lock_guard<std::mutex> lk(dg::g_prof_mu);
  int c = 0;
  double t = 0;
  for (auto& r : dg::g_prof) {
    if (r.name != name) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return dg::kHip;
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return dg::kHip;
    ++c;
    t += ms;
  }
  *count = c;
  *total_ms = t;
  return 0;
}
}
