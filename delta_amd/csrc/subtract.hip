// Full-frame spatially varying subtraction on gfx950 (reference
// src/subtract.cpp:429-966):
//
//   k_model     D = S - b - sum_n a_n (g_ny *_y (g_nx *_x R_clean))  (FP32)
//   k_variance  Var(D) = (K_c^2 (x) Var(R)) + Var(S), K frozen per 64x64 tile
//   k_mask_*    mask growth, edge border, non-finite flags and median fill
//
// Data layout: every image is row-major FP32 (masks uint8), index y*w + x, as
// the reference's ImageView.  Each model workgroup owns a 64x64 output tile:
// it stages the sanitised, haloed reference window in LDS (transposed, so the
// sliding x-pass reads are conflict-free), runs the x-pass of one nx into a
// second LDS buffer, and every thread then runs the y-passes of the components
// sharing that nx over a 16-row column strip held in registers, accumulating
// a_n * B_n straight into its 16 difference pixels.  Nothing but S, R (and the
// mask) is read from HBM and only D is written.
#include "dg_kernels.hpp"

namespace dg {
namespace {

constexpr int kTile = 64;    // output tile edge
constexpr int kStrip = 16;   // rows per thread in the y-pass
constexpr int kXg = 8;       // columns per x-pass task
constexpr int kThreads = 256;
constexpr int kStageBatch = 12;  // global loads per thread in flight while staging a window

struct ModelArgs {
  int w, h, n_max, nc;
  const float* sci;
  const float* ref;
  const uint8_t* rmask;
  const float* fill;   // device scalar: reference fill value
  const float* taps;   // (n_max+1) x KS
  const int* comp;     // (n_max+1) x (n_max+1): component index of (nx, ny), -1 if none
  DevFields F;
  float* diff;
};

template <int KS>
struct ModelGeom {
  static constexpr int R = KS / 2;
  static constexpr int NR = kTile + 2 * R;    // staged rows
  static constexpr int NC = kTile + 2 * R;    // staged columns
  static constexpr int LDR = NR + 1;          // RwT leading dim (column-major)
  static constexpr int LDT = kTile + 1;       // T leading dim (row-major)
  static constexpr size_t lds_bytes = sizeof(float) * (size_t(NC) * LDR + size_t(NR) * LDT);
};

template <int KS>
__global__ __launch_bounds__(kThreads) void k_model(ModelArgs a) {
  using G = ModelGeom<KS>;
  constexpr int R = G::R;
  extern __shared__ float smem[];
  float* rwt = smem;                          // [NC][LDR]: rwt[c*LDR + r]
  float* tx = smem + size_t(G::NC) * G::LDR;  // [NR][LDT]
  const int tid = threadIdx.x;
  const int bx = blockIdx.x * kTile, by = blockIdx.y * kTile;
  const int w = a.w, h = a.h;

  // 1. stage the haloed reference window, sanitised (masked / non-finite ->
  //    fill value, off-frame -> 0)
  //    kStageBatch loads per thread go out before any is used, so the
  //    window costs a few memory round trips rather than one per element
  const float fillv = *a.fill;
  {
    constexpr int kTot = G::NR * G::NC;
    for (int base = 0; base < kTot; base += kThreads * kStageBatch) {
      float v[kStageBatch];
      uint8_t bad[kStageBatch];
#pragma unroll
      for (int j = 0; j < kStageBatch; ++j) {
        const int idx = base + tid + kThreads * j;
        const int rr = idx / G::NC, cc = idx - rr * G::NC;
        const int gy = by - R + rr, gx = bx - R + cc;
        v[j] = 0.0f;
        bad[j] = 0;
        if (idx < kTot && gy >= 0 && gy < h && gx >= 0 && gx < w) {
          const size_t i = size_t(gy) * w + gx;
          v[j] = a.ref[i];
          bad[j] = a.rmask ? a.rmask[i] : uint8_t(kMaskGood);
        }
      }
#pragma unroll
      for (int j = 0; j < kStageBatch; ++j) {
        const int idx = base + tid + kThreads * j;
        if (idx >= kTot) break;
        const int rr = idx / G::NC, cc = idx - rr * G::NC;
        const float x = (bad[j] != kMaskGood || !isfinite(v[j])) ? fillv : v[j];
        rwt[cc * G::LDR + rr] = x;
      }
    }
  }

  // 2. per-thread strip: column c, rows q*16 .. q*16+15
  const int c = tid & (kTile - 1);
  const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index: uniform
  const int x = bx + c;
  const int xc = x < w ? x : w - 1;
  const int y0 = by + q * kStrip;
  int jx = 0;
  float wxv = 0.0f;
  if (!a.F.exact) {
    jx = a.F.cellx[xc];
    wxv = a.F.wx[xc];
  }
  int jy[kStrip];
  float fy[kStrip];
#pragma unroll
  for (int r = 0; r < kStrip; ++r) {
    const int yc = min(y0 + r, h - 1);
    jy[r] = a.F.exact ? 0 : a.F.celly[yc];
    fy[r] = a.F.exact ? 0.0f : a.F.wy[yc];
  }
  // field value of f at (x, y0 + r) for every strip row
  auto field_strip = [&](int f, float (&out)[kStrip]) {
    if (a.F.exact) {
#pragma unroll
      for (int r = 0; r < kStrip; ++r)
        out[r] = a.F.full[(size_t(f) * h + min(y0 + r, h - 1)) * w + xc];
      return;
    }
    const double* base_f = a.F.lat + size_t(f) * a.F.ncy * a.F.ncx;
    float a00 = 0, a01 = 0, a10 = 0, a11 = 0;
    int cur = -1;
#pragma unroll
    for (int r = 0; r < kStrip; ++r) {
      if (jy[r] != cur) {
        cur = jy[r];
        const double* c0 = base_f + size_t(cur) * a.F.ncx + jx;
        const double* c1 = c0 + a.F.ncx;
        a00 = static_cast<float>(c0[0]);
        a01 = static_cast<float>(c0[1]);
        a10 = static_cast<float>(c1[0]);
        a11 = static_cast<float>(c1[1]);
      }
      const float bse = fmaf(a10 - a00, fy[r], a00);
      const float top = a01 - a00;
      const float slope = fmaf((a11 - a10) - top, fy[r], top);
      out[r] = fmaf(slope, wxv, bse);
    }
  };

  float d[kStrip];
  {
    float bg[kStrip];
    field_strip(a.nc, bg);
#pragma unroll
    for (int r = 0; r < kStrip; ++r) {
      const int yc = min(y0 + r, h - 1);
      d[r] = a.sci[size_t(yc) * w + xc] - bg[r];
    }
  }

  for (int nx = 0; nx <= a.n_max; ++nx) {
    __syncthreads();  // staged window ready / previous T consumed
    // 3. x-pass of order nx over all staged rows: T[r][c] = sum_j g[j] W[r][c+2R-j]
    const float* gx = a.taps + nx * KS;
    for (int u = tid; u < G::NR * (kTile / kXg); u += kThreads) {
      const int r = u % G::NR, s = u / G::NR;
      float in[kXg + KS - 1];
#pragma unroll
      for (int i = 0; i < kXg + KS - 1; ++i) in[i] = rwt[(s * kXg + i) * G::LDR + r];
      float o[kXg];
#pragma unroll
      for (int t = 0; t < kXg; ++t) o[t] = 0.0f;
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const float g = gx[j];
#pragma unroll
        for (int t = 0; t < kXg; ++t) o[t] = fmaf(g, in[t + 2 * R - j], o[t]);
      }
#pragma unroll
      for (int t = 0; t < kXg; ++t) tx[r * G::LDT + s * kXg + t] = o[t];
    }
    __syncthreads();
    // 4. y-passes of every component with this nx
    float col[kStrip + KS - 1];
#pragma unroll
    for (int i = 0; i < kStrip + KS - 1; ++i) col[i] = tx[(q * kStrip + i) * G::LDT + c];
    for (int ny = 0; ny <= a.n_max - nx; ++ny) {
      const float* gy = a.taps + ny * KS;
      float b[kStrip];
#pragma unroll
      for (int r = 0; r < kStrip; ++r) b[r] = 0.0f;
#pragma unroll
      for (int j = 0; j < KS; ++j) {
        const float g = gy[j];
#pragma unroll
        for (int r = 0; r < kStrip; ++r) b[r] = fmaf(g, col[r + 2 * R - j], b[r]);
      }
      float av[kStrip];
      field_strip(a.comp[nx * (a.n_max + 1) + ny], av);
#pragma unroll
      for (int r = 0; r < kStrip; ++r) d[r] = fmaf(-av[r], b[r], d[r]);
    }
  }
  if (x < w) {
#pragma unroll
    for (int r = 0; r < kStrip; ++r)
      if (y0 + r < h) a.diff[size_t(y0 + r) * w + x] = d[r];
  }
}

// ------------------------------------------------------------------ variance
constexpr int kVb = 64;  // variance tile edge

struct VarArgs {
  int w, h, nc, ks;
  int r;              // true kernel radius (<= RK; phi zero-padded to RK)
  const float* vr;    // may be null
  const float* vs;    // may be null
  const float* phi;   // nc x ks x ks footprints (float)(gy*gx)
  DevFields F;
  float* var;
};

template <int RK>
struct VarGeom {
  static constexpr int KS = 2 * RK + 1;
  static constexpr int KK = KS * KS;
  static constexpr int NW = kVb + 2 * RK;  // window edge
  static constexpr int LDW = NW + 1;
  // doubles first (box sums, the wave partials of Q and the centred row sums)
  // so they stay 8-byte aligned
  static constexpr size_t lds_doubles = (RK + 2) + 4 + size_t(KS) * (RK + 1);
  static constexpr size_t lds_bytes =
      sizeof(double) * lds_doubles +
      sizeof(float) * (size_t(NW) * LDW + 2 * KK + 64);
  static_assert(size_t(NW) * LDW >= size_t(kVb) * (kVb + 1),
                "the output staging tile reuses the window's LDS");
};

template <int RK>
__global__ __launch_bounds__(kThreads) void k_variance(VarArgs a) {
  using G = VarGeom<RK>;
  constexpr int KS = G::KS, KK = G::KK;
  extern __shared__ double smem_d[];
  double* box = smem_d;                        // box sums per hw, then Q
  double* qw = smem_d + RK + 2;                // per-wave partials of Q
  double* rowc = qw + 4;                       // [KS][RK+1] centred row sums
  float* win = reinterpret_cast<float*>(smem_d + G::lds_doubles);  // [NW][LDW]
  float* kf = win + size_t(G::NW) * G::LDW;    // mirrored K^2
  float* kv = kf + KK;                         // K itself (unmirrored)
  float* outt = win;                           // [64][65] output staging, once win is read
  float* an = kv + KK;                         // nc field values (<= 64)
  __shared__ float s_min[kThreads / 64], s_max[kThreads / 64];
  __shared__ int s_hw;
  __shared__ float s_q;
  __shared__ int s_flat;

  const int tid = threadIdx.x;
  const int bx = blockIdx.x * kVb, by = blockIdx.y * kVb;
  const int w = a.w, h = a.h;
  const int bw = min(kVb, w - bx), bh = min(kVb, h - by);

  // (a) field values at the tile centre
  const int cpx = min(bx + bw / 2, w - 1), cpy = min(by + bh / 2, h - 1);
  if (tid < a.nc) an[tid] = field_at(a.F, tid, cpx, cpy, w, h);
  __syncthreads();
  // (b) K = sum_n a_n phi_n (component order), mirrored K^2.  Each thread
  //     holds its taps i = tid + 256 u; the footprint loads of 8 components
  //     go out together, so a tile waits on nc / 8 round trips, not nc.
  {
    constexpr int kTaps = (KK + kThreads - 1) / kThreads;
    constexpr int kCb = 8;
    float k[kTaps];
#pragma unroll
    for (int u = 0; u < kTaps; ++u) k[u] = 0.0f;
    for (int n0 = 0; n0 < a.nc; n0 += kCb) {
      float ph[kCb][kTaps];
#pragma unroll
      for (int j = 0; j < kCb; ++j)
#pragma unroll
        for (int u = 0; u < kTaps; ++u) {
          const int i = tid + kThreads * u;
          ph[j][u] = (n0 + j < a.nc && i < KK) ? a.phi[size_t(n0 + j) * KK + i] : 0.0f;
        }
#pragma unroll
      for (int j = 0; j < kCb; ++j) {
        if (n0 + j >= a.nc) break;
        const float anj = an[n0 + j];
#pragma unroll
        for (int u = 0; u < kTaps; ++u) k[u] = fmaf(anj, ph[j][u], k[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kTaps; ++u) {
      const int i = tid + kThreads * u;
      if (i < KK) {
        kv[i] = k[u];
        kf[KK - 1 - i] = k[u] * k[u];
      }
    }
  }
  // (c) gather the zero-padded Var(R) window and its min / max, kStageBatch
  //     loads per thread in flight together
  float lmin = INFINITY, lmax = -INFINITY;
  {
    constexpr int kTot = G::NW * G::NW;
    // the flat test looks at the window of the true radius only; with a
    // zero-padded footprint (r < RK) it sits off = RK - r into the staged one
    const int off = RK - a.r;
    for (int base = 0; base < kTot; base += kThreads * kStageBatch) {
      float v[kStageBatch];
#pragma unroll
      for (int j = 0; j < kStageBatch; ++j) {
        const int idx = base + tid + kThreads * j;
        const int iy = idx / G::NW, ix = idx - iy * G::NW;
        const int gy = by - RK + iy, gx = bx - RK + ix;
        v[j] = 0.0f;
        if (idx < kTot && iy >= off && ix >= off && iy < off + bh + 2 * a.r &&
            ix < off + bw + 2 * a.r && gy >= 0 && gy < h && gx >= 0 && gx < w)
          v[j] = a.vr[size_t(gy) * w + gx];
      }
#pragma unroll
      for (int j = 0; j < kStageBatch; ++j) {
        const int idx = base + tid + kThreads * j;
        if (idx >= kTot) break;
        const int iy = idx / G::NW, ix = idx - iy * G::NW;
        if (iy >= off && ix >= off && iy < off + bh + 2 * a.r && ix < off + bw + 2 * a.r) {
          lmin = (v[j] < lmin) ? v[j] : lmin;
          lmax = (lmax < v[j]) ? v[j] : lmax;
        }
        win[iy * G::LDW + ix] = v[j];
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float mn = __shfl_xor(lmin, off, 64), mx = __shfl_xor(lmax, off, 64);
    lmin = (mn < lmin) ? mn : lmin;
    lmax = (lmax < mx) ? mx : lmax;
  }
  if ((tid & 63) == 0) {
    s_min[tid >> 6] = lmin;
    s_max[tid >> 6] = lmax;
  }
  __syncthreads();
  // (d) Q = sum K^2 (each square exact in FP64) and the box-shrink sums of the
  //     mirrored float K^2, in FP64 by deterministic parallel reductions:
  //     per row yy the centred sums rowc[yy][h] = sum_{|xx - RK| <= h} kf,
  //     then box(h) = sum_{|yy - RK| <= h} rowc[yy][h]; Q by a tree.
  {
    double q = 0.0;
    for (int i = tid; i < KK; i += kThreads) q = fma(double(kv[i]), double(kv[i]), q);
    for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off, 64);
    if ((tid & 63) == 0) qw[tid >> 6] = q;
    if (tid < KS) {
      const float* krow = kf + tid * KS;
      double c = double(krow[RK]);
      rowc[tid * (RK + 1)] = c;
      for (int hh = 1; hh <= RK; ++hh) {
        c += double(krow[RK - hh]) + double(krow[RK + hh]);
        rowc[tid * (RK + 1) + hh] = c;
      }
    }
  }
  __syncthreads();
  if (tid < RK) {
    double bs = 0.0;
    for (int yy = RK - tid; yy <= RK + tid; ++yy) bs += rowc[yy * (RK + 1) + tid];
    box[tid] = bs;
  } else if (tid == RK + 1) {
    box[RK + 1] = (qw[0] + qw[1]) + (qw[2] + qw[3]);
  }
  __syncthreads();
  if (tid == 0) {
    const double qsum = box[RK + 1];
    const double target = (1.0 - double(1e-4f)) * qsum;
    int hw = 0;
    for (; hw < a.r; ++hw)
      if (box[hw] >= target) break;
    s_hw = hw;
    s_q = static_cast<float>(qsum);
    float wmin = INFINITY, wmax = -INFINITY;
    for (int i = 0; i < kThreads / 64; ++i) {
      wmin = (s_min[i] < wmin) ? s_min[i] : wmin;
      wmax = (wmax < s_max[i]) ? s_max[i] : wmax;
    }
    s_flat = (wmax > 0.0f && (wmax - wmin) <= 1e-3f * wmax) ? 1 : 0;
  }
  __syncthreads();

  // (e) per thread: output row oy = tid % 64, 16 columns ox0 .. ox0+15
  const int oy = tid & 63;
  const int ox0 = (tid >> 6) * kStrip;
  float acc[kStrip];
  if (s_flat) {
    const float qf = s_q;
#pragma unroll
    for (int t = 0; t < kStrip; ++t) acc[t] = fmaf(qf, win[(oy + RK) * G::LDW + RK + ox0 + t], 0.0f);
  } else {
    const int hw = s_hw, lo = RK - hw, hi = RK + hw;
#pragma unroll
    for (int t = 0; t < kStrip; ++t) acc[t] = 0.0f;
    for (int ly = lo; ly <= hi; ++ly) {
      float rowv[kStrip + KS - 1];
      const float* src = win + (oy + ly) * G::LDW + ox0;
      // only the columns the box [lo, hi] reads (lo, hi are uniform)
#pragma unroll
      for (int i = 0; i < kStrip + KS - 1; ++i)
        if (i >= lo && i < hi + kStrip) rowv[i] = src[i];
      const float* krow = kf + ly * KS;
#pragma unroll
      for (int lx = 0; lx < KS; ++lx) {
        if (lx < lo || lx > hi) continue;
        const float cf = krow[lx];
        if (cf == 0.0f) continue;
#pragma unroll
        for (int t = 0; t < kStrip; ++t) acc[t] = fmaf(cf, rowv[t + lx], acc[t]);
      }
    }
  }
  __syncthreads();  // every thread is done reading win before outt overwrites it
#pragma unroll
  for (int t = 0; t < kStrip; ++t) outt[oy * (kVb + 1) + ox0 + t] = acc[t];
  __syncthreads();
  // (f) coalesced store, adding Var(S) (its loads issued together)
  {
    constexpr int kIt = kVb * kVb / kThreads;
    float sv[kIt];
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
      const int idx = tid + kThreads * j, ly = idx >> 6, lx = idx & 63;
      sv[j] = (a.vs && ly < bh && lx < bw) ? a.vs[size_t(by + ly) * w + bx + lx] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
      const int idx = tid + kThreads * j, ly = idx >> 6, lx = idx & 63;
      if (ly < bh && lx < bw) {
        float v = outt[ly * (kVb + 1) + lx];
        if (a.vs) v += sv[j];
        a.var[size_t(by + ly) * w + bx + lx] = v;
      }
    }
  }
}

__global__ void k_var_science_only(const float* __restrict__ vs, float* __restrict__ var,
                                   size_t n) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i < n) var[i] = 0.0f + vs[i];
}

// ------------------------------------------------------------------ mask growth
struct MaskInitArgs {
  size_t n;
  const float* sci;
  const float* ref;
  const uint8_t* sm;
  const uint8_t* rm;
  int has_sat;
  float sat;
  uint8_t* out;
  uint8_t* grow;  // may be null
  uint8_t* eb;
};

__global__ void k_mask_init(MaskInitArgs a) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i >= a.n) return;
  uint8_t direct = 0, grow = 0, eb = 0;
  if (a.sm && a.sm[i]) {
    direct |= a.sm[i];
    eb = 1;
  }
  if (a.rm && a.rm[i]) {
    grow |= a.rm[i];
    eb = 1;
  }
  if (a.has_sat) {
    if (a.ref[i] >= a.sat) {
      grow |= kMaskSaturated;
      eb = 1;
    }
    if (a.sci[i] >= a.sat) {
      direct |= kMaskSaturated;
      eb = 1;
    }
  }
  a.out[i] = direct;
  if (a.grow) a.grow[i] = grow;
  a.eb[i] = eb;
}

// x-dilation: OR over [x-r, x+r] clamped to the row
__global__ __launch_bounds__(256) void k_dilate_x(const uint8_t* __restrict__ in,
                                                  uint8_t* __restrict__ out, int w, int h,
                                                  int r) {
  extern __shared__ uint8_t row[];
  const int y = blockIdx.y;
  const int x0 = blockIdx.x * 256;
  const int span = 256 + 2 * r;
  for (int i = threadIdx.x; i < span; i += 256) {
    const int gx = x0 - r + i;
    row[i] = (gx >= 0 && gx < w) ? in[size_t(y) * w + gx] : 0;
  }
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= w) return;
  uint8_t acc = 0;
  for (int i = 0; i <= 2 * r; ++i) acc |= row[threadIdx.x + i];
  out[size_t(y) * w + x] = acc;
}

// y-dilation + edge border + non-finite difference flags
__global__ void k_mask_finish(const uint8_t* __restrict__ gx, const float* __restrict__ diff,
                              uint8_t* __restrict__ out, uint8_t* __restrict__ eb, int w, int h,
                              int r, int has_grow) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w) return;
  const size_t i = size_t(y) * w + x;
  uint8_t m = out[i];
  if (has_grow) {
    const int ya = max(0, y - r), yb = min(h - 1, y + r);
    uint8_t acc = 0;
    for (int yy = ya; yy <= yb; ++yy) acc |= gx[size_t(yy) * w + x];
    m |= acc;
  }
  if (x < r || x >= w - r || y < r || y >= h - r) m |= kMaskEdge;
  if (!isfinite(diff[i])) {
    m |= kMaskNonFinite;
    eb[i] = 1;
  }
  out[i] = m;
}

__global__ void k_mask_fill(const uint8_t* __restrict__ out, const uint8_t* __restrict__ eb,
                            const float* __restrict__ fill, float* __restrict__ diff,
                            float* __restrict__ var, size_t n) {
  const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  if (out[i] != kMaskGood && eb[i]) {
    diff[i] = *fill;
    if (var) var[i] = 0.0f;
  }
}

template <int KS>
void launch_model(const ModelArgs& a, hipStream_t s) {
  const size_t lds = ModelGeom<KS>::lds_bytes;
  set_max_lds(reinterpret_cast<const void*>(&k_model<KS>), int(lds));
  dim3 grid((a.w + kTile - 1) / kTile, (a.h + kTile - 1) / kTile);
  hipLaunchKernelGGL(k_model<KS>, grid, dim3(kThreads), lds, s, a);
  DG_LAUNCH_CHECK();
}

template <int RK>
void launch_variance(const VarArgs& a, hipStream_t s) {
  const size_t lds = VarGeom<RK>::lds_bytes;
  set_max_lds(reinterpret_cast<const void*>(&k_variance<RK>), int(lds));
  dim3 grid((a.w + kVb - 1) / kVb, (a.h + kVb - 1) / kVb);
  hipLaunchKernelGGL(k_variance<RK>, grid, dim3(kThreads), lds, s, a);
  DG_LAUNCH_CHECK();
}

// Radii above 20 run on the radius-25 or radius-30 instantiation with the taps
// and footprints zero-padded on both sides.  That is bit-exact: the model's
// staged window is sanitised (finite), so each leading zero tap leaves the
// running sum at +0 and each trailing one adds a signed zero to a nonzero or
// +0 sum; the variance kernel takes the true radius for its flat-window test
// and box-shrink search, and padded K^2 taps are skipped.
constexpr int kMaxRadius = 30;
int padded_radius(int r) { return r <= 20 ? r : (r <= 25 ? 25 : kMaxRadius); }

void dispatch_model(int ks, const ModelArgs& a, hipStream_t s) {
  switch (ks) {
#define DG_KS(K) \
  case K:        \
    launch_model<K>(a, s); \
    return;
    DG_KS(3) DG_KS(5) DG_KS(7) DG_KS(9) DG_KS(11) DG_KS(13) DG_KS(15) DG_KS(17) DG_KS(19)
    DG_KS(21) DG_KS(23) DG_KS(25) DG_KS(27) DG_KS(29) DG_KS(31) DG_KS(33) DG_KS(35) DG_KS(37)
    DG_KS(39) DG_KS(41) DG_KS(51) DG_KS(61)
#undef DG_KS
    default:
      invalid("kernel radius must be in [1, 30] for the GPU subtraction");
  }
}

void dispatch_variance(int rk, const VarArgs& a, hipStream_t s) {
  switch (rk) {
#define DG_RK(K) \
  case K:        \
    launch_variance<K>(a, s); \
    return;
    DG_RK(1) DG_RK(2) DG_RK(3) DG_RK(4) DG_RK(5) DG_RK(6) DG_RK(7) DG_RK(8) DG_RK(9) DG_RK(10)
    DG_RK(11) DG_RK(12) DG_RK(13) DG_RK(14) DG_RK(15) DG_RK(16) DG_RK(17) DG_RK(18) DG_RK(19)
    DG_RK(20) DG_RK(25) DG_RK(30)
#undef DG_RK
    default:
      invalid("kernel radius must be in [1, 30] for the GPU subtraction");
  }
}

// rows of ks taps -> rows of ks + 2 pad taps, zeros on both sides
std::vector<float> pad_taps(const std::vector<float>& t, int rows, int ks, int pad) {
  if (pad == 0) return t;
  const int kp = ks + 2 * pad;
  std::vector<float> out(size_t(rows) * kp, 0.0f);
  for (int r = 0; r < rows; ++r)
    std::copy(t.begin() + size_t(r) * ks, t.begin() + size_t(r + 1) * ks,
              out.begin() + size_t(r) * kp + pad);
  return out;
}

// nc footprints of ks x ks -> (ks + 2 pad)^2, centred, zeros around
std::vector<float> pad_footprints(const std::vector<float>& f, int nc, int ks, int pad) {
  if (pad == 0) return f;
  const int kp = ks + 2 * pad;
  std::vector<float> out(size_t(nc) * kp * kp, 0.0f);
  for (int c = 0; c < nc; ++c)
    for (int y = 0; y < ks; ++y)
      std::copy(f.begin() + (size_t(c) * ks + y) * ks, f.begin() + (size_t(c) * ks + y + 1) * ks,
                out.begin() + (size_t(c) * kp + y + pad) * kp + pad);
  return out;
}

}  // namespace

// Device-resident subtraction: all image pointers are device pointers.
void run_subtract(int w, int h, const float* sci, const float* ref, const float* svar,
                  const float* rvar, const uint8_t* smask, const uint8_t* rmask,
                  const double* knots, int n_knots, const double* theta, int n_theta,
                  double beta, int n_max, int radius, double saturation, float* diff,
                  float* var, uint8_t* mask, hipStream_t s) {
  if (w <= 0 || h <= 0) invalid("subtract: empty frame");
  const GhBasis basis(beta, n_max, radius > 0 ? radius : GhBasis::default_radius(beta, n_max));
  const Tps tps(knots, n_knots);
  const int nc = basis.nc, nf = nc + 1;
  if (n_theta != n_knots * nf)
    runtime("subtract: theta size does not match k*(nc+1)");
  if (nc > 64) invalid("subtract: at most 64 basis components are supported");
  const bool want_var = (svar != nullptr) || (rvar != nullptr);
  const bool want_mask = (smask != nullptr) || (rmask != nullptr) || saturation > 0.0;
  if (want_var != (var != nullptr)) invalid("subtract: variance output must be given iff a variance input is");
  if (want_mask != (mask != nullptr)) invalid("subtract: mask output must be given iff a mask input or saturation is");

  if (basis.radius > kMaxRadius)
    invalid("subtract: kernel radius must be in [1, 30] for the GPU subtraction");
  const size_t n = size_t(w) * h;
  const int rp = padded_radius(basis.radius), pad = rp - basis.radius;
  const int ks = 2 * rp + 1;  // footprint edge the kernels run with
  ArenaLease lease;
  Arena& ar = *lease;
  size_t need = fields_scratch_bytes(tps, nc, w, h) + 2 * select_scratch_bytes(n, 7) +
                Arena::round(sizeof(float) * 4) +
                Arena::round(sizeof(float) * size_t(n_max + 1) * ks) +
                Arena::round(sizeof(int) * (n_max + 1) * (n_max + 1)) +
                Arena::round(sizeof(float) * size_t(nc) * ks * ks);
  if (want_mask) need += 3 * Arena::round(n);
  ar.reserve(need + 4096);

  float* d_scal = ar.take<float>(4);  // [0] ref fill, [1] diff fill
  // the small uploads (taps, component map, footprints, the field build's)
  // in one copy from pinned staging; the stream is drained before return, so
  // the next call on this thread may reuse the staging
  PinnedStage::local().reset();
  UploadGroup up(ar);
  const std::vector<float> taps = pad_taps(basis.taps_f32(), n_max + 1, basis.ks, pad);
  float* d_taps = up.add(taps.data(), taps.size());
  std::vector<int> comp((n_max + 1) * (n_max + 1), -1);
  for (int c = 0; c < nc; ++c) comp[basis.order[c].first * (n_max + 1) + basis.order[c].second] = c;
  int* d_comp = up.add(comp.data(), comp.size());
  std::vector<float> phi;
  float* d_phi = nullptr;
  if (want_var && rvar) {
    phi = pad_footprints(basis.footprints_f32(), nc, basis.ks, pad);
    d_phi = up.add(phi.data(), phi.size());
  }

  FieldsBuild fb = [&] {
    ProfScope ps("fields", s);
    return build_fields(tps, theta, nc, w, h, ar, s, &up);
  }();
  {
    ProfScope ps("select", s);
    select_strided_median(ref, rmask, n, 7, d_scal, ar, s);
  }

  ModelArgs ma{};
  ma.w = w;
  ma.h = h;
  ma.n_max = n_max;
  ma.nc = nc;
  ma.sci = sci;
  ma.ref = ref;
  ma.rmask = rmask;
  ma.fill = d_scal;
  ma.taps = d_taps;
  ma.comp = d_comp;
  ma.F = fb.dev;
  ma.diff = diff;
  {
    ProfScope ps("model", s);
    dispatch_model(ks, ma, s);
  }

  if (want_var) {
    if (rvar) {
      VarArgs va{};
      va.w = w;
      va.h = h;
      va.nc = nc;
      va.ks = ks;
      va.r = basis.radius;
      va.vr = rvar;
      va.vs = svar;
      va.phi = d_phi;
      va.F = fb.dev;
      va.var = var;
      ProfScope ps("variance", s);
      dispatch_variance(rp, va, s);
    } else {
      hipLaunchKernelGGL(k_var_science_only, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s,
                         svar, var, n);
      DG_LAUNCH_CHECK();
    }
  }

  if (want_mask) {
    ProfScope ps("mask", s);
    uint8_t* eb = ar.take<uint8_t>(n);
    uint8_t* grow = ar.take<uint8_t>(n);
    uint8_t* tmp = ar.take<uint8_t>(n);
    const bool has_grow = rmask != nullptr || saturation > 0.0;
    MaskInitArgs mi{n, sci, ref, smask, rmask, saturation > 0.0 ? 1 : 0,
                    static_cast<float>(saturation), mask, has_grow ? grow : nullptr, eb};
    hipLaunchKernelGGL(k_mask_init, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, mi);
    const int r = basis.radius;
    if (has_grow) {
      hipLaunchKernelGGL(k_dilate_x, dim3((w + 255) / 256, h), dim3(256), 256 + 2 * r, s, grow,
                         tmp, w, h, r);
    }
    hipLaunchKernelGGL(k_mask_finish, dim3((w + 255) / 256, h), dim3(256), 0, s, tmp, diff, mask,
                       eb, w, h, r, has_grow ? 1 : 0);
    DG_LAUNCH_CHECK();
    select_strided_median(diff, mask, n, 7, d_scal + 1, ar, s);
    hipLaunchKernelGGL(k_mask_fill, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, mask, eb,
                       d_scal + 1, diff, var, n);
    DG_LAUNCH_CHECK();
  }
  DG_HIP(hipStreamSynchronize(s));
}

}  // namespace dg
