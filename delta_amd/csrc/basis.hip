// Gauss-Hermite basis convolution of a whole frame (reference
// src/convolve.cpp:24-113, binding src/bindings.cpp:209-227):
//
//   B_c = g_ny(c) *_y (g_nx(c) *_x image),  c = 0 .. nc-1
//
// both passes zero-padded, 'same' size, true convolution, FP32 with the taps
// accumulated in ascending order as fused multiply-adds (the reference's
// tap-outer axpy loops, contracted by its compiler), so every B_c is bit-equal
// to the reference's.  Out-of-frame source pixels contribute exact zeros, the
// same as the reference skipping them.
//
// One workgroup per 64x64 output tile: the haloed input window is staged in
// LDS (transposed, so the sliding x-pass reads are conflict-free), the x-pass
// of one nx goes to a second LDS buffer over all staged rows, and each thread
// runs the y-passes of the components sharing that nx on a 16-row column strip
// held in registers, storing each B_c strip straight to HBM.  The stack is
// write-bound: 4 bytes in and 4*nc bytes out per pixel.
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kBTile = 64;
constexpr int kBStrip = 16;
constexpr int kBXg = 8;
constexpr int kBThreads = 256;

struct BasisArgs {
  int w, h, n_max;
  const float* img;
  const float* taps;  // (n_max+1) x KS
  const int* comp;    // (n_max+1) x (n_max+1): component of (nx, ny), -1 if none
  float* out;         // nc x h x w
};

template <int KS>
struct BasisGeom {
  static constexpr int R = KS / 2;
  static constexpr int NR = kBTile + 2 * R;
  static constexpr int NC = kBTile + 2 * R;
  static constexpr int LDR = NR + 1;
  static constexpr int LDT = kBTile + 1;
  static constexpr size_t lds_bytes = sizeof(float) * (size_t(NC) * LDR + size_t(NR) * LDT);
};

template <int KS>
__global__ __launch_bounds__(kBThreads) void k_basis_convolve(BasisArgs a) {
  using G = BasisGeom<KS>;
  constexpr int R = G::R;
  extern __shared__ float bsm[];
  float* win = bsm;                          // [NC][LDR]: column-major staged window
  float* tx = bsm + size_t(G::NC) * G::LDR;  // [NR][LDT]
  const int tid = threadIdx.x;
  const int bx = blockIdx.x * kBTile, by = blockIdx.y * kBTile;
  const int w = a.w, h = a.h;
  // 12 loads per thread in flight before any is stored
  {
    constexpr int kTot = G::NR * G::NC, kB = 12;
    for (int base = 0; base < kTot; base += kBThreads * kB) {
      float v[kB];
#pragma unroll
      for (int j = 0; j < kB; ++j) {
        const int idx = base + tid + kBThreads * j;
        const int rr = idx / G::NC, cc = idx - rr * G::NC;
        const int gy = by - R + rr, gx = bx - R + cc;
        v[j] = (idx < kTot && gy >= 0 && gy < h && gx >= 0 && gx < w)
                   ? a.img[size_t(gy) * w + gx] : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < kB; ++j) {
        const int idx = base + tid + kBThreads * j;
        if (idx >= kTot) break;
        const int rr = idx / G::NC, cc = idx - rr * G::NC;
        win[cc * G::LDR + rr] = v[j];
      }
    }
  }
  const int c = tid & (kBTile - 1);
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int x = bx + c;
  const int y0 = by + q * kBStrip;
  const size_t plane = size_t(w) * h;
  for (int nx = 0; nx <= a.n_max; ++nx) {
    __syncthreads();
    const float* gx = a.taps + nx * KS;
    for (int u = tid; u < G::NR * (kBTile / kBXg); u += kBThreads) {
      const int r = u % G::NR, s = u / G::NR;
      float in[kBXg + KS - 1];
#pragma unroll
      for (int i = 0; i < kBXg + KS - 1; ++i) in[i] = win[(s * kBXg + i) * G::LDR + r];
      float o[kBXg];
#pragma unroll
      for (int t = 0; t < kBXg; ++t) o[t] = 0.0f;
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const float g = gx[j];
#pragma unroll
        for (int t = 0; t < kBXg; ++t) o[t] = fmaf(g, in[t + 2 * R - j], o[t]);
      }
#pragma unroll
      for (int t = 0; t < kBXg; ++t) tx[r * G::LDT + s * kBXg + t] = o[t];
    }
    __syncthreads();
    float col[kBStrip + KS - 1];
#pragma unroll
    for (int i = 0; i < kBStrip + KS - 1; ++i) col[i] = tx[(q * kBStrip + i) * G::LDT + c];
    for (int ny = 0; ny <= a.n_max - nx; ++ny) {
      const float* gy = a.taps + ny * KS;
      float b[kBStrip];
#pragma unroll
      for (int r = 0; r < kBStrip; ++r) b[r] = 0.0f;
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const float g = gy[j];
#pragma unroll
        for (int r = 0; r < kBStrip; ++r) b[r] = fmaf(g, col[r + 2 * R - j], b[r]);
      }
      if (x < w) {
        float* o = a.out + size_t(a.comp[nx * (a.n_max + 1) + ny]) * plane;
#pragma unroll
        for (int r = 0; r < kBStrip; ++r)
          if (y0 + r < h) __builtin_nontemporal_store(b[r], o + size_t(y0 + r) * w + x);
      }
    }
  }
}

template <int KS>
void launch_basis(const BasisArgs& a, hipStream_t s) {
  const size_t lds = BasisGeom<KS>::lds_bytes;
  set_max_lds(reinterpret_cast<const void*>(&k_basis_convolve<KS>), int(lds));
  dim3 grid((a.w + kBTile - 1) / kBTile, (a.h + kBTile - 1) / kBTile);
  hipLaunchKernelGGL(k_basis_convolve<KS>, grid, dim3(kBThreads), lds, s, a);
  DG_LAUNCH_CHECK();
}

void dispatch_basis(int ks, const BasisArgs& a, hipStream_t s) {
  switch (ks) {
#define DG_BKS(K)          \
  case K:                  \
    launch_basis<K>(a, s); \
    return;
    DG_BKS(3) DG_BKS(5) DG_BKS(7) DG_BKS(9) DG_BKS(11) DG_BKS(13) DG_BKS(15) DG_BKS(17)
    DG_BKS(19) DG_BKS(21) DG_BKS(23) DG_BKS(25) DG_BKS(27) DG_BKS(29) DG_BKS(31) DG_BKS(33)
    DG_BKS(35) DG_BKS(37) DG_BKS(39) DG_BKS(41)
#undef DG_BKS
    default:
      invalid("basis_convolve: kernel radius must be in [1, 20] on the GPU");
  }
}

}  // namespace

size_t basis_convolve_scratch_bytes(int n_max, int ks) {
  return Arena::round(sizeof(float) * size_t(n_max + 1) * ks) +
         Arena::round(sizeof(int) * size_t(n_max + 1) * (n_max + 1)) + (1 << 12);
}

// Device-resident basis convolution: img is w x h, out is nc x h x w, both in
// HBM; runs on stream s.
void run_basis_convolve(int w, int h, const float* img, double beta, int n_max, int radius,
                        float* out, hipStream_t s) {
  if (w <= 0 || h <= 0) invalid("basis_convolve: empty image");
  const GhBasis basis(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
  const int ks = basis.ks;
  if (ks > 41) invalid("basis_convolve: kernel radius must be in [1, 20] on the GPU");
  ArenaLease lease;
  Arena& ar = *lease;
  ar.reserve(basis_convolve_scratch_bytes(n_max, ks));
  const std::vector<float> taps = basis.taps_f32();
  std::vector<int> comp(size_t(n_max + 1) * (n_max + 1), -1);
  for (int c = 0; c < basis.nc; ++c)
    comp[basis.order[c].first * (n_max + 1) + basis.order[c].second] = c;
  float* d_taps = ar.take<float>(taps.size());
  int* d_comp = ar.take<int>(comp.size());
  h2d(d_taps, taps.data(), taps.size(), s);
  h2d(d_comp, comp.data(), comp.size(), s);
  BasisArgs a{};
  a.w = w;
  a.h = h;
  a.n_max = n_max;
  a.img = img;
  a.taps = d_taps;
  a.comp = d_comp;
  a.out = out;
  ProfScope ps("basis_convolve", s);
  dispatch_basis(ks, a, s);
  // the staged taps live in the leased arena: finish before it is returned
  DG_HIP(hipStreamSynchronize(s));
}

}  // namespace dg
