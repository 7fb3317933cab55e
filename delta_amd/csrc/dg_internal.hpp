// Internal host-side declarations of libdeltagpu: error plumbing, device
// buffers, the Gauss-Hermite basis and thin-plate-spline setup shared by every
// entry point.  Nothing here is part of the C ABI (include/deltagpu.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace dg {

// ---------------------------------------------------------------- errors
enum Status : int { kOk = 0, kInvalid = 1, kRuntime = 2, kHip = 3 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void invalid(const std::string& m) { throw Error(kInvalid, m); }
[[noreturn]] inline void runtime(const std::string& m) { throw Error(kRuntime, m); }

void set_last_error(const std::string& m);
const char* last_error();

#define DG_HIP(expr)                                                           \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::dg::Error(::dg::kHip, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define DG_LAUNCH_CHECK() DG_HIP(hipGetLastError())

// Runs `fn`, converting exceptions into a status code + last_error text.
template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return kOk;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return kRuntime;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return kRuntime;
  }
}

// ---------------------------------------------------------------- device memory
// Grow-only scratch arena.  A call takes one arena from a process-wide pool for
// its duration, so concurrent calls never share scratch, and repeated calls of
// the same shape reuse their allocation instead of paying hipMalloc each time.
class Arena {
 public:
  ~Arena();
  void reserve(size_t bytes);  // ensure capacity (may reallocate)
  void reset() { used_ = 0; }
  template <class T>
  T* take(size_t n) {
    const size_t bytes = ((n * sizeof(T)) + 255) & ~size_t(255);
    if (used_ + bytes > cap_) runtime("internal: arena overflow");
    T* p = reinterpret_cast<T*>(base_ + used_);
    used_ += bytes;
    return p;
  }
  static size_t round(size_t bytes) { return (bytes + 255) & ~size_t(255); }
  // Bump-pointer checkpoints: buffers taken after mark() are released by
  // rewind(); stream order keeps reuse safe for work queued on one stream.
  size_t mark() const { return used_; }
  void rewind(size_t m) { used_ = m; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};

class ArenaLease {
 public:
  ArenaLease();
  ~ArenaLease();
  Arena& operator*() { return *a_; }
  Arena* operator->() { return a_; }

 private:
  Arena* a_;
};

// Device copy helpers on a stream.
template <class T>
void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) DG_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* dst, const T* src, size_t n, hipStream_t s) {
  if (n) DG_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

// Raise a kernel's dynamic-LDS limit to `bytes` unless an earlier call in
// this process already raised it at least that far (the attribute call is a
// driver round trip; launch paths call this on every launch).
void set_max_lds(const void* kernel, int bytes);

// Per-thread pinned staging for the small host-to-device copies of one call
// that drains its stream before returning: copies from pinned memory are
// asynchronous, where a copy from pageable memory makes the host wait for its
// staging.  reset() at the start of such a call (every earlier copy from this
// thread has completed by then); a copy that no longer fits goes through the
// pageable path.  The 4 MiB buffer of a thread is kept until the process
// ends (it is not freed from thread-exit destructors, which can run after the
// HIP runtime has shut down).
class PinnedStage {
 public:
  static PinnedStage& local();
  void reset() { used_ = 0; }
  template <class T>
  void h2d(T* dst, const T* src, size_t n, hipStream_t s) {
    const size_t bytes = n * sizeof(T);
    if (!bytes) return;
    char* p = grab(bytes);
    if (!p) {
      DG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
      return;
    }
    std::memcpy(p, src, bytes);
    DG_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s));
  }
  // Pinned bytes for the caller to fill and copy from (nullptr if full).
  char* grab(size_t bytes);

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
};

// Device buffers taken consecutively from an arena and filled from host
// arrays by one copy: commit() packs the host data into pinned staging at the
// buffers' own relative offsets and copies the whole span at once.
class UploadGroup {
 public:
  explicit UploadGroup(Arena& ar) : ar_(ar) {}
  template <class T>
  T* add(const T* src, size_t n) {
    T* d = ar_.take<T>(n > 0 ? n : 1);
    items_.push_back(Item{reinterpret_cast<char*>(d), src, n * sizeof(T)});
    return d;
  }
  void commit(PinnedStage& stage, hipStream_t s) {
    if (items_.empty()) return;
    char* base = items_.front().dst;
    const Item& last = items_.back();
    const size_t span = size_t(last.dst - base) + last.bytes;
    char* p = span ? stage.grab(span) : nullptr;
    if (!p) {
      for (const Item& it : items_)
        if (it.bytes) DG_HIP(hipMemcpyAsync(it.dst, it.src, it.bytes, hipMemcpyHostToDevice, s));
    } else {
      for (const Item& it : items_)
        if (it.bytes) std::memcpy(p + (it.dst - base), it.src, it.bytes);
      DG_HIP(hipMemcpyAsync(base, p, span, hipMemcpyHostToDevice, s));
    }
    items_.clear();
  }

 private:
  struct Item {
    char* dst;
    const void* src;
    size_t bytes;
  };
  Arena& ar_;
  std::vector<Item> items_;
};

// Owning stream for host-pointer entry points (one per call, default priority).
class OwnedStream {
 public:
  OwnedStream() { DG_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~OwnedStream() { (void)hipStreamDestroy(s_); }
  hipStream_t get() const { return s_; }

 private:
  hipStream_t s_{};
};

// ---------------------------------------------------------------- basis / TPS
// Cartesian Gauss-Hermite basis (reference src/basis.cpp:9-59).
struct GhBasis {
  double beta = 0;
  int n_max = 0, radius = 0, ks = 0, nc = 0;
  std::vector<double> g;              // (n_max+1) x ks samples, FP64
  std::vector<std::pair<int, int>> order;  // component -> (nx, ny)
  GhBasis(double beta, int n_max, int radius);
  static int default_radius(double beta, int n_max);
  std::vector<float> taps_f32() const;      // (n_max+1) x ks float copy
  std::vector<float> footprints_f32() const;  // nc x ks x ks, (float)(gy*gx)
  std::vector<double> component_sums() const;
};

// Low-rank thin-plate spline over k knots (reference src/spatial.cpp:20-111).
// k == 1 is the constant-field mode: design row [1], zero penalty.
struct Tps {
  int k = 0;
  double cx = 0, cy = 0, scale = 1;
  std::vector<double> kn;   // k x 2 normalised knots
  std::vector<double> Z;    // k x (k-3), row-major
  std::vector<double> pen;  // k x k
  Tps(const double* knots, int k);
  void design_row(double px, double py, double* row) const;
  double min_spacing() const;
};

// Host description of the coarse field lattice (reference
// src/subtract.cpp:113-316): per-axis sample coordinates, bracketing cells and
// fractional weights.  The lattice values themselves are evaluated on device.
struct LatticeAxes {
  bool exact = false;
  int stride = 1, ncx = 0, ncy = 0;
  std::vector<int> gx, gy, cellx, celly;
  std::vector<float> wx, wy;
};
LatticeAxes lattice_axes(const Tps& tps, int w, int h);

// Compute-unit count of the current device (cached per device).
int device_cu_count();

// Result of the kernel fit (the reference's KernelFit, include/delta/fit.hpp:16-28).
struct FitOut {
  std::vector<double> theta, curve;
  double lambda = 0, gcv = 0, edf = 0, rss = 0, reduced_chi2 = 0;
  double cv_exact_design_bytes = 0;  // N x P x 8 of the last exact CV pass, else 0
  int n_pixels = 0, n_used = 0, n_total = 0, n_rejected = 0, passes = 0;
  std::vector<int> stamp_x, stamp_y;
  std::vector<double> stamp_chi2;
  std::vector<uint8_t> stamp_accepted;
};

// Device-side parameters of the spatial model (uploaded per call).
struct DevTps {
  int k, nr;
  double cx, cy, scale;
  const double* kn;     // k x 2
  const double* Z;      // k x nr
  const double* theta;  // nf x k
};

}  // namespace dg

namespace dg {
// Optional per-kernel timing with HIP events on the launch stream.  Off by
// default; when enabled every instrumented launch records an event pair that
// dg_profile_read() resolves (after the caller's own synchronisation).
bool prof_enabled();
void prof_begin(const char* name, hipStream_t s);
void prof_end(hipStream_t s);

struct ProfScope {
  hipStream_t s;
  bool on;
  ProfScope(const char* name, hipStream_t st) : s(st), on(prof_enabled()) {
    if (on) prof_begin(name, s);
  }
  ~ProfScope() {
    if (on) prof_end(s);
  }
};
}  // namespace dg
