// Backprojection kernels for gfx950.
//
// Layout: one workgroup of 256 lanes (4 wave64) owns a tile of
// tileX x tileY pixels of the column-major image (lanes run along iy, so a
// wave's loads and stores of v[ix][iy] are contiguous).  Views are streamed in
// chunks: per chunk the workgroup computes, for each view, the window of
// detector cells its tile can reach (from the tile's corner pixels), stages
// only those cells of the filtered rows in LDS, then every lane walks the
// chunk's views in order for its pixel.  The image value lives in a register
// for the whole launch.
//
//   FP64 path   per view v = float(double(v) + contribution), contributions
//               in the reference's evaluation order (this file is compiled
//               with -ffp-contract=off), so IntDiff is bit-exact and the
//               trig families differ only where device and host libm differ.
//   mixed path  FP32 positions, weights and interpolation from FP64-derived
//               per-view constants, Kahan-compensated accumulation.
#include <hip/hip_runtime.h>

#include <cstring>

#include "bp_plan.h"
#include "filter_plan.h"
#include "crmath.h"
#include "mixed_device.h"

// FP64 parity kernels keep the reference's rounding: no contraction.
#pragma clang fp contract(off)

namespace ctsim {
namespace {

constexpr int kBlock = 256;
using dev::kMaxChunk;
using dev::xcd_tile;
// Dynamic LDS of the mixed kernel: 64 KB per workgroup less its static arrays.
constexpr size_t kMixedLdsMax = 62 * 1024;
constexpr size_t kSymLds = 62 * 1024;  // coefficient staging of the symmetric kernel

struct BpArgs {
  int nx, ny, nDet, iDetCenter, nView, fstride;
  int tileX, tileY, tilesY, W, VC;  // W = 0: stage full rows
  double detInc, rotScale, F, sdl;
  const double* r;
  const double* phi;
  const float* xs;
  const float* ys;
  const void* f;
  const void* vc;
  float* image;
  float endScale;  // mixed path: factor applied to the accumulated sum (Trig)
};

__device__ __forceinline__ int nearest_int(double x) {
  return x > 0 ? static_cast<int>(x + 0.5) : static_cast<int>(x - 0.5);
}

// Window read with an index clamp (memory safety only: the window bound makes
// every tap land inside).
template <class T>
__device__ __forceinline__ T wget(const T* w, int lo, int Wv, int i) {
  int k = i - lo;
  k = k < 0 ? 0 : (k >= Wv ? Wv - 1 : k);
  return w[k];
}

// CubicPolyInterpolator::interpolate on a window that holds detectors
// [lo, lo + Wv) (libctsupport/interpolator.cpp:36-75).
__device__ __forceinline__ double cubic_f64(const double* w, int lo, int Wv, int n, double x) {
  const int l = static_cast<int>(floor(x)) - 1;
  const int hi = l + 3;
  if (l < -1) return 0;
  if (l == -1) {
    const double y0 = wget(w, lo, Wv, 0), y1 = wget(w, lo, Wv, 1);
    return y0 + x * (y1 - y0);
  }
  if (hi > n) return 0;
  if (hi == n) {
    const double frac = x - (l + 1);
    const double a = wget(w, lo, Wv, n - 2), b = wget(w, lo, Wv, n - 1);
    return a + frac * (b - a);
  }
  const double xd0 = x - l, xd1 = x - (l + 1), xd2 = x - (l + 2), xd3 = x - (l + 3);
  const double oneSixth = (1. / 6.);
  double y = xd1 * xd2 * xd3 * -oneSixth * wget(w, lo, Wv, l);
  y += xd0 * xd2 * xd3 * 0.5 * wget(w, lo, Wv, l + 1);
  y += xd0 * xd1 * xd3 * -0.5 * wget(w, lo, Wv, l + 2);
  y += xd0 * xd1 * xd2 * oneSixth * wget(w, lo, Wv, l + 3);
  return y;
}

// ------------------------------------------------------------ FP64 path

// Detector position (iDetCenter + L/detInc, or the fixed-point position for
// IntDiff) of one pixel in one view, as the FP64 kernel computes it; used for
// the tile window.
template <int KIND>
__device__ __forceinline__ double pos_f64(const BpArgs& a, const ViewConstF64& c, int ix, int iy) {
  if constexpr (KIND == KIND_IDIFF) {
    long long dx, dy, cs;
    __builtin_memcpy(&dx, &c.a, 8);
    __builtin_memcpy(&dy, &c.b, 8);
    __builtin_memcpy(&cs, &c.c, 8);
    const long long cur = cs + static_cast<long long>(ix) * dx + static_cast<long long>(iy) * dy;
    return static_cast<double>(cur >> 32);
  } else {
    const size_t k = static_cast<size_t>(ix) * a.ny + iy;
    const double r = a.r[k], phi = a.phi[k];
    if constexpr (KIND == KIND_TRIG || KIND == KIND_TABLE) {
      return a.iDetCenter + (r * crm_cos(c.theta - phi)) / a.detInc;
    } else {
      double s, co;
      crm_sincos(c.theta - phi, &s, &co);
      const double rc = r * co, rs = r * s;
      if constexpr (KIND == KIND_EQUIANGULAR) {
        return a.iDetCenter + crm_atan(rc / (a.F + rs)) / a.detInc;
      } else {
        const double U = (a.F + rs) / a.F;
        return a.iDetCenter + (rc / U) * (a.sdl / a.F) / a.detInc;
      }
    }
  }
}

template <int KIND, int INTERP, class TF>
__global__ __launch_bounds__(kBlock) void bp_f64_kernel(BpArgs a) {
  extern __shared__ double s_f[];
  __shared__ int s_lo[kMaxChunk];
  const int t = threadIdx.x;
  const int nTiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, nTiles);
  const int ix0 = (tile / a.tilesY) * a.tileX, iy0 = (tile % a.tilesY) * a.tileY;
  const int ix = ix0 + t / a.tileY, iy = iy0 + t % a.tileY;
  const bool active = (t < a.tileX * a.tileY) && ix < a.nx && iy < a.ny;
  const size_t pix = static_cast<size_t>(ix) * a.ny + iy;
  const int cx1 = min(ix0 + a.tileX - 1, a.nx - 1), cy1 = min(iy0 + a.tileY - 1, a.ny - 1);
  const int Wv = a.W > 0 ? a.W : a.nDet + 8;
  const ViewConstF64* vcs = static_cast<const ViewConstF64*>(a.vc);
  const TF* frows = static_cast<const TF*>(a.f);

  float v = 0.f;
  double r = 0, phi = 0;
  if (active) {
    v = a.image[pix];
    if constexpr (KIND != KIND_IDIFF) {
      r = a.r[pix];
      phi = a.phi[pix];
    }
  }
  const int nDet = a.nDet;
  const int last = nDet - 1;

  for (int v0 = 0; v0 < a.nView; v0 += a.VC) {
    const int nv = min(a.VC, a.nView - v0);
    if (t < nv) {
      int lo = -4;
      if (a.W > 0) {
        const ViewConstF64 c = vcs[v0 + t];
        double p0 = pos_f64<KIND>(a, c, ix0, iy0), p1 = pos_f64<KIND>(a, c, cx1, iy0);
        double p2 = pos_f64<KIND>(a, c, ix0, cy1), p3 = pos_f64<KIND>(a, c, cx1, cy1);
        double pmin = fmin(fmin(p0, p1), fmin(p2, p3));
        // positions outside the representable range: stage anything, the
        // pixels reject them by their own bounds tests
        if (!(pmin > -1e9 && pmin < 1e9)) pmin = -4;
        lo = static_cast<int>(floor(pmin)) - 4;
      }
      s_lo[t] = lo;
    }
    __syncthreads();
    for (int vv = 0; vv < nv; vv++) {
      const int lo = s_lo[vv];
      const TF* frow = frows + static_cast<size_t>(v0 + vv) * a.fstride;
      double* w = s_f + vv * Wv;
      for (int j = t; j < Wv; j += kBlock) {
        const int d = lo + j;
        w[j] = (d >= 0 && d < nDet) ? static_cast<double>(frow[d]) : 0.0;
      }
    }
    __syncthreads();
    if (active) {
      for (int vv = 0; vv < nv; vv++) {
        const ViewConstF64 c = vcs[v0 + vv];
        const double* w = s_f + vv * Wv;
        const int lo = s_lo[vv];
        if constexpr (KIND == KIND_IDIFF) {
          // BackprojectIntDiff::BackprojectView, backprojectors.cpp:579-653
          long long dx, dy, cs;
          __builtin_memcpy(&dx, &c.a, 8);
          __builtin_memcpy(&dy, &c.b, 8);
          __builtin_memcpy(&cs, &c.c, 8);
          const long long cur = cs + static_cast<long long>(ix) * dx + static_cast<long long>(iy) * dy;
          if constexpr (INTERP == INTERP_NEAREST) {
            const int i = static_cast<int>((cur + (1LL << 31)) >> 32);
            if (i >= 0 && i <= last) v = static_cast<float>(static_cast<double>(v) + wget(w, lo, Wv, i));
          } else if constexpr (INTERP == INTERP_LINEAR) {
            const long long i = cur >> 32;
            if (i >= 0 && i <= last) {
              const long long rem = cur & 0xFFFFFFFFLL;
              const double f0 = wget(w, lo, Wv, static_cast<int>(i));
              const double df = i < last ? (wget(w, lo, Wv, static_cast<int>(i) + 1) - f0) * (1. / 4294967296.)
                                         : 0.0;
              const double cval = f0 + (static_cast<double>(rem) * df);
              v = static_cast<float>(static_cast<double>(v) + cval);
            }
          } else {
            const double cval = cubic_f64(w, lo, Wv, nDet, static_cast<double>(cur) / 4294967296.);
            v = static_cast<float>(static_cast<double>(v) + cval);
          }
        } else if constexpr (KIND == KIND_TRIG || KIND == KIND_TABLE) {
          // BackprojectTrig / BackprojectTable::BackprojectView,
          // backprojectors.cpp:365-404, 444-481
          const double L = r * crm_cos(c.theta - phi);
          bool hit = false;
          double cval = 0;
          if constexpr (INTERP == INTERP_NEAREST) {
            const int i = a.iDetCenter + nearest_int(L / a.detInc);
            if (i >= 0 && i < nDet) { hit = true; cval = wget(w, lo, Wv, i); }
          } else if constexpr (INTERP == INTERP_LINEAR) {
            const double p = L / a.detInc;
            const double pf = floor(p);
            const int i = a.iDetCenter + static_cast<int>(pf);
            const double frac = p - pf;
            if (i >= 0 && i < nDet - 1) {
              hit = true;
              cval = (1 - frac) * wget(w, lo, Wv, i) + frac * wget(w, lo, Wv, i + 1);
            }
          } else {
            const double p = a.iDetCenter + (L / a.detInc);
            if (p >= 0 && p < nDet) { hit = true; cval = cubic_f64(w, lo, Wv, nDet, p); }
          }
          if (hit) {
            if constexpr (KIND == KIND_TRIG) cval = a.rotScale * cval;
            v = static_cast<float>(static_cast<double>(v) + cval);
          }
        } else {
          // BackprojectEquiangular / BackprojectEquilinear::BackprojectView,
          // backprojectors.cpp:656-743
          double s, co;
          crm_sincos(c.theta - phi, &s, &co);
          const double rc = r * co, rs = r * s;
          double pos, wgt;
          if constexpr (KIND == KIND_EQUIANGULAR) {
            const double fl = a.F + rs;
            const double gamma = crm_atan(rc / fl);
            pos = gamma / a.detInc;
            wgt = fl * fl + (rc * rc);
          } else {
            const double U = (a.F + rs) / a.F;
            double dp = rc / U;
            dp *= a.sdl / a.F;
            pos = dp / a.detInc;
            wgt = U * U;
          }
          bool hit = false;
          double cval = 0;
          if constexpr (INTERP == INTERP_NEAREST) {
            const int i = a.iDetCenter + nearest_int(pos);
            if (i >= 0 && i < nDet) { hit = true; cval = wget(w, lo, Wv, i) / wgt; }
          } else if constexpr (INTERP == INTERP_LINEAR) {
            const double pf = floor(pos);
            const int i = a.iDetCenter + static_cast<int>(pf);
            const double frac = pos - pf;
            if (i >= 0 && i < nDet - 1) {
              hit = true;
              const double f0 = wget(w, lo, Wv, i), f1 = wget(w, lo, Wv, i + 1);
              cval = (f0 + frac * (f1 - f0)) / wgt;
            }
          } else {
            const double d = a.iDetCenter + pos;
            if (d >= 0 && d < nDet) { hit = true; cval = cubic_f64(w, lo, Wv, nDet, d) / wgt; }
          }
          if (hit) v = static_cast<float>(static_cast<double>(v) + cval);
        }
      }
    }
    __syncthreads();
  }
  if (active) a.image[pix] = v;
}

// CubicPolyInterpolator::interpolate straight from a detector row.
template <class TF>
__device__ __forceinline__ double cubic_direct(const TF* Y, int n, double x) {
  const int l = static_cast<int>(floor(x)) - 1;
  const int hi = l + 3;
  if (l < -1) return 0;
  if (l == -1) return static_cast<double>(Y[0]) + x * (static_cast<double>(Y[1]) - static_cast<double>(Y[0]));
  if (hi > n) return 0;
  if (hi == n) {
    const double frac = x - (l + 1);
    const double a = Y[n - 2], b = Y[n - 1];
    return a + frac * (b - a);
  }
  const double xd0 = x - l, xd1 = x - (l + 1), xd2 = x - (l + 2), xd3 = x - (l + 3);
  const double oneSixth = (1. / 6.);
  double y = xd1 * xd2 * xd3 * -oneSixth * static_cast<double>(Y[l]);
  y += xd0 * xd2 * xd3 * 0.5 * static_cast<double>(Y[l + 1]);
  y += xd0 * xd1 * xd3 * -0.5 * static_cast<double>(Y[l + 2]);
  y += xd0 * xd1 * xd2 * oneSixth * static_cast<double>(Y[l + 3]);
  return y;
}

// BackprojectDiff (backprojectors.cpp:517-570): the detector position walks by
// repeated FP64 addition along x and then y, and the column pointer only
// advances for in-range samples, so one lane owns one column and walks it view
// after view.  Parity path only.
template <int INTERP, class TF>
__global__ __launch_bounds__(64) void bp_diff_kernel(BpArgs a) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= a.nx) return;
  const ViewConstF64* vcs = static_cast<const ViewConstF64*>(a.vc);
  const TF* frows = static_cast<const TF*>(a.f);
  float* col = a.image + static_cast<size_t>(ix) * a.ny;
  const int nDet = a.nDet, last = nDet - 1;
  for (int vw = 0; vw < a.nView; vw++) {
    const ViewConstF64 c = vcs[vw];
    const TF* fr = frows + static_cast<size_t>(vw) * a.fstride;
    auto f = [fr](int i) -> double { return static_cast<double>(fr[i]); };
    double start = c.c;
    for (int k = 0; k < ix; k++) start += c.a;
    double cur = start;
    int out = 0;
    for (int iy = 0; iy < a.ny; iy++, cur += c.b) {
      if constexpr (INTERP == INTERP_NEAREST) {
        const int i = nearest_int(cur);
        if (i >= 0 && i < nDet) { col[out] = static_cast<float>(static_cast<double>(col[out]) + f(i)); out++; }
      } else if constexpr (INTERP == INTERP_LINEAR) {
        const double pf = floor(cur);
        const int i = static_cast<int>(pf);
        const double frac = cur - pf;
        if (i >= 0 && i <= last) {
          const double df = i < last ? f(i + 1) - f(i) : 0.0;
          const double cval = f(i) + (frac * df);
          col[out] = static_cast<float>(static_cast<double>(col[out]) + cval);
          out++;
        }
      } else {
        const double p = cur;
        if (p >= 0 && p < nDet) {
          const double cval = cubic_direct(fr, nDet, p);
          col[out] = static_cast<float>(static_cast<double>(col[out]) + cval);
          out++;
        }
      }
    }
  }
}

// ----------------------------------------------------------- mixed path
#pragma clang fp contract(fast)
using dev::CoefT;
using dev::MixedArgs;

// The mixed path interpolates from per-interval coefficients: on detector
// interval i (position p in [i, i+1), u = p - i) the reference's interpolant
// is the polynomial a0 + u (a1 + u (a2 + u a3)), zero outside the detector,
// with its end rules folded in (CubicPolyInterpolator::interpolate,
// interpolator.cpp:36-75; linear with the IntDiff last-cell rule,
// backprojectors.cpp:635-638).  coef_kernel writes them once per view, in
// FP64 rounded to FP32, so the backprojection inner loop is a window lookup
// and three FMAs.
template <int INTERP, bool LAST_INCL>
__global__ __launch_bounds__(256) void coef_kernel(const float* __restrict__ f, int fstride,
                                                   int nDet, int nView, int blocksPerRow,
                                                   void* __restrict__ out) {
  const int v = blockIdx.x / blocksPerRow;
  const int i = (blockIdx.x - v * blocksPerRow) * 256 + threadIdx.x;
  if (i >= nDet || v >= nView) return;
  const float* row = f + static_cast<size_t>(v) * fstride;
  auto rd = [row](int j) { return row[j]; };
  static_cast<CoefT<INTERP>*>(out)[static_cast<size_t>(v) * nDet + i] =
      dev::interval_coef<INTERP, LAST_INCL>(rd, i, nDet);
}

// One workgroup per (tile, view group).  With a multiple of 8 view groups,
// the groups are dealt to XCDs: XCD k (block ids k, k+8, ...) runs groups
// [k*split/8, (k+1)*split/8) over every tile, so it reads only its own
// views' coefficients, which stay in its L2; otherwise neighbouring tiles
// share an XCD.
template <int GK, int INTERP, int P, int DEG>
__global__ __launch_bounds__(256) void bp_mixed_kernel(MixedArgs a) {
  extern __shared__ float4 s_dyn[];
  __shared__ float4 s_v[dev::kMaxChunk + 2];
  __shared__ int s_lo[dev::kMaxChunk];
  const int nTiles = a.tilesX * a.tilesY;
  int tile, grp;
  if ((a.split & 7) == 0) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    grp = xcd * (a.split >> 3) + local / nTiles;
    tile = local % nTiles;
  } else {
    const int bid = dev::xcd_tile(blockIdx.x, gridDim.x);
    tile = bid % nTiles;
    grp = bid / nTiles;
  }
  dev::bp_mixed_tile<GK, INTERP, P, DEG>(a, tile, grp, reinterpret_cast<CoefT<INTERP>*>(s_dyn), s_v,
                                         s_lo);
}

// The mirror variant: one workgroup per (tile of the mirror domain, view
// group), dealt to XCDs as above.
template <int GK, int INTERP, int P, int DEG, int NM>
__global__ __launch_bounds__(256) void bp_mirror_kernel(MixedArgs a) {
  const int nTiles = a.tilesX * a.tilesY;
  int tile, grp;
  if ((a.split & 7) == 0) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    grp = xcd * (a.split >> 3) + local / nTiles;
    tile = local % nTiles;
  } else {
    const int bid = dev::xcd_tile(blockIdx.x, gridDim.x);
    tile = bid % nTiles;
    grp = bid / nTiles;
  }
  dev::bp_mirror_tile<GK, INTERP, P, DEG, NM>(a, tile, grp);
}

// The rotationally symmetric variant: one workgroup per (quadrant tile, group
// of base views), dealt to XCDs as above.
template <int GK, int INTERP, int P, int DEG, int S, int K>
__global__ __launch_bounds__(256) void bp_sym_kernel(MixedArgs a) {
  extern __shared__ float4 s_dyn[];
  __shared__ float2 s_ab[16];
  __shared__ dev::SymWin s_win[64];
  __shared__ int s_n;
  const int nTiles = a.tilesX * a.tilesY;
  int tile, grp;
  if ((a.split & 7) == 0) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    grp = xcd * (a.split >> 3) + local / nTiles;
    tile = local % nTiles;
  } else {
    const int bid = dev::xcd_tile(blockIdx.x, gridDim.x);
    tile = bid % nTiles;
    grp = bid / nTiles;
  }
  dev::bp_sym_tile<GK, INTERP, P, DEG, S, K>(a, tile, grp, reinterpret_cast<CoefT<INTERP>*>(s_dyn),
                                             s_ab, s_win, &s_n);
}

// The software-pipelined symmetric variant (same mapping).
template <int GK, int INTERP, int P, int DEG, int S, int K>
__global__ __launch_bounds__(256) void bp_sym_pipe_kernel(MixedArgs a) {
  extern __shared__ float4 s_dyn[];
  __shared__ float2 s_ab[64];
  __shared__ dev::SymWin s_win[256];
  const int nTiles = a.tilesX * a.tilesY;
  int tile, grp;
  if ((a.split & 7) == 0) {
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    grp = xcd * (a.split >> 3) + local / nTiles;
    tile = local % nTiles;
  } else {
    const int bid = dev::xcd_tile(blockIdx.x, gridDim.x);
    tile = bid % nTiles;
    grp = bid / nTiles;
  }
  dev::bp_sym_tile_pipe<GK, INTERP, P, DEG, S, K>(a, tile, grp, reinterpret_cast<CoefT<INTERP>*>(s_dyn),
                                                  s_ab, s_win);
}

// image += endScale * (sum of the view groups' partial images), groups added
// in order in FP64 (overwrite: image = endScale * sum).
__global__ void bp_combine_kernel(float* __restrict__ image, const float* __restrict__ partial,
                                  size_t n, int split, double endScale, int overwrite) {
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    double s = 0;
    for (int g = 0; g < split; g++) s += partial[static_cast<size_t>(g) * n + i];
    image[i] = static_cast<float>((overwrite ? 0.0 : static_cast<double>(image[i])) + endScale * s);
  }
}

__global__ void bp_fill_kernel(float* img, size_t n, float v) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) img[i] = v;
}

#pragma clang fp contract(off)
// PostProcessing / ScaleImageByRotIncrement, backprojectors.cpp:330-336.
__global__ void bp_scale_kernel(float* img, size_t n, double s) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) img[i] = static_cast<float>(static_cast<double>(img[i]) * s);
}

template <int KIND, int INTERP, class TF>
void launch_f64(const BpArgs& a, int grid, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((bp_f64_kernel<KIND, INTERP, TF>), dim3(grid), dim3(kBlock), lds, s, a);
}

template <int INTERP, class TF>
void launch_f64_kind(int kind, const BpArgs& a, int grid, size_t lds, hipStream_t s) {
  switch (kind) {
    case KIND_TRIG: launch_f64<KIND_TRIG, INTERP, TF>(a, grid, lds, s); break;
    case KIND_TABLE: launch_f64<KIND_TABLE, INTERP, TF>(a, grid, lds, s); break;
    case KIND_IDIFF: launch_f64<KIND_IDIFF, INTERP, TF>(a, grid, lds, s); break;
    case KIND_EQUIANGULAR: launch_f64<KIND_EQUIANGULAR, INTERP, TF>(a, grid, lds, s); break;
    case KIND_EQUILINEAR: launch_f64<KIND_EQUILINEAR, INTERP, TF>(a, grid, lds, s); break;
  }
}

template <class TF>
void launch_f64_interp(int interp, int kind, const BpArgs& a, int grid, size_t lds,
                       hipStream_t s) {
  switch (interp) {
    case INTERP_NEAREST: launch_f64_kind<INTERP_NEAREST, TF>(kind, a, grid, lds, s); break;
    case INTERP_LINEAR: launch_f64_kind<INTERP_LINEAR, TF>(kind, a, grid, lds, s); break;
    default: launch_f64_kind<INTERP_CUBIC, TF>(kind, a, grid, lds, s); break;
  }
}

template <class TF>
void launch_diff(int interp, const BpArgs& a, int grid, hipStream_t s) {
  switch (interp) {
    case INTERP_NEAREST: hipLaunchKernelGGL((bp_diff_kernel<INTERP_NEAREST, TF>), dim3(grid), dim3(64), 0, s, a); break;
    case INTERP_LINEAR: hipLaunchKernelGGL((bp_diff_kernel<INTERP_LINEAR, TF>), dim3(grid), dim3(64), 0, s, a); break;
    default: hipLaunchKernelGGL((bp_diff_kernel<INTERP_CUBIC, TF>), dim3(grid), dim3(64), 0, s, a); break;
  }
}

template <int GK, int INTERP, int DEG>
void launch_mixed_p(int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  switch (P) {
    case 1: hipLaunchKernelGGL((bp_mixed_kernel<GK, INTERP, 1, DEG>), dim3(grid), dim3(256), lds, s, a); break;
    case 2: hipLaunchKernelGGL((bp_mixed_kernel<GK, INTERP, 2, DEG>), dim3(grid), dim3(256), lds, s, a); break;
    default: hipLaunchKernelGGL((bp_mixed_kernel<GK, INTERP, 4, DEG>), dim3(grid), dim3(256), lds, s, a); break;
  }
}

template <int INTERP>
void launch_mixed(int gk, int deg, int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  if (gk == 0) return launch_mixed_p<0, INTERP, 0>(P, a, grid, lds, s);
  if (gk == 2) return launch_mixed_p<2, INTERP, 0>(P, a, grid, lds, s);
  switch (deg) {
    case 4: launch_mixed_p<1, INTERP, 4>(P, a, grid, lds, s); break;
    case 6: launch_mixed_p<1, INTERP, 6>(P, a, grid, lds, s); break;
    case 8: launch_mixed_p<1, INTERP, 8>(P, a, grid, lds, s); break;
    default: launch_mixed_p<1, INTERP, 0>(P, a, grid, lds, s); break;
  }
}

template <int GK, int INTERP, int DEG, int NM>
void launch_mirror_p(int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  switch (P) {
    case 1: hipLaunchKernelGGL((bp_mirror_kernel<GK, INTERP, 1, DEG, NM>), dim3(grid), dim3(256), lds, s, a); break;
    case 2: hipLaunchKernelGGL((bp_mirror_kernel<GK, INTERP, 2, DEG, NM>), dim3(grid), dim3(256), lds, s, a); break;
    default: hipLaunchKernelGGL((bp_mirror_kernel<GK, INTERP, 4, DEG, NM>), dim3(grid), dim3(256), lds, s, a); break;
  }
}

template <int GK, int INTERP, int DEG>
void launch_mirror_nm(int NM, int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  if (NM == 4) launch_mirror_p<GK, INTERP, DEG, 4>(P, a, grid, lds, s);
  else launch_mirror_p<GK, INTERP, DEG, 2>(P, a, grid, lds, s);
}

template <int INTERP>
void launch_mirror(int gk, int deg, int NM, int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  if (gk == 0) return launch_mirror_nm<0, INTERP, 0>(NM, P, a, grid, lds, s);
  if (gk == 2) return launch_mirror_nm<2, INTERP, 0>(NM, P, a, grid, lds, s);
  switch (deg) {
    case 4: launch_mirror_nm<1, INTERP, 4>(NM, P, a, grid, lds, s); break;
    case 6: launch_mirror_nm<1, INTERP, 6>(NM, P, a, grid, lds, s); break;
    case 8: launch_mirror_nm<1, INTERP, 8>(NM, P, a, grid, lds, s); break;
    default: launch_mirror_nm<1, INTERP, 0>(NM, P, a, grid, lds, s); break;
  }
}

template <int GK, int INTERP, int DEG, int S, int K>
void launch_sym_p(bool pipe, int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  if (pipe) {
    if (P == 1) hipLaunchKernelGGL((bp_sym_pipe_kernel<GK, INTERP, 1, DEG, S, K>), dim3(grid), dim3(256), lds, s, a);
    else hipLaunchKernelGGL((bp_sym_pipe_kernel<GK, INTERP, 2, DEG, S, K>), dim3(grid), dim3(256), lds, s, a);
  } else {
    if (P == 1) hipLaunchKernelGGL((bp_sym_kernel<GK, INTERP, 1, DEG, S, K>), dim3(grid), dim3(256), lds, s, a);
    else hipLaunchKernelGGL((bp_sym_kernel<GK, INTERP, 2, DEG, S, K>), dim3(grid), dim3(256), lds, s, a);
  }
}

template <int GK, int INTERP, int DEG>
void launch_sym_sk(bool pipe, int S, int K, int P, const MixedArgs& a, int grid, size_t lds, hipStream_t s) {
  if (S == 4) launch_sym_p<GK, INTERP, DEG, 4, 1>(pipe, P, a, grid, lds, s);
  else if (K == 1) launch_sym_p<GK, INTERP, DEG, 2, 1>(pipe, P, a, grid, lds, s);
  else launch_sym_p<GK, INTERP, DEG, 2, 2>(pipe, P, a, grid, lds, s);
}

template <int INTERP>
void launch_sym(bool pipe, int gk, int deg, int S, int K, int P, const MixedArgs& a, int grid, size_t lds,
                hipStream_t s) {
  if (gk == 0) return launch_sym_sk<0, INTERP, 0>(pipe, S, K, P, a, grid, lds, s);
  if (gk == 2) return launch_sym_sk<2, INTERP, 0>(pipe, S, K, P, a, grid, lds, s);
  switch (deg) {
    case 4: launch_sym_sk<1, INTERP, 4>(pipe, S, K, P, a, grid, lds, s); break;
    case 6: launch_sym_sk<1, INTERP, 6>(pipe, S, K, P, a, grid, lds, s); break;
    case 8: launch_sym_sk<1, INTERP, 8>(pipe, S, K, P, a, grid, lds, s); break;
    default: launch_sym_sk<1, INTERP, 0>(pipe, S, K, P, a, grid, lds, s); break;
  }
}

// Per-view constants for this launch: reuse the plan's device copy when the
// angles are those of the previous launch, else compute them on the host and
// copy them in stream order (after any kernel still reading the old ones).
template <class VC, class F>
int view_consts(ctsim_ctx* ctx, const BpPlan& p, int kind, const double* h_angles, int nView,
                F compute, const void** out) {
  const bool same = p.cachedKind == kind && static_cast<int>(p.cachedAngles.size()) == nView &&
                    memcmp(p.cachedAngles.data(), h_angles, sizeof(double) * nView) == 0;
  if (!same) {
    const size_t bytes = static_cast<size_t>(nView) * sizeof(VC);
    CTSIM_CHECK(ctx->stage.reserve(bytes, ctx->stream));
    compute(static_cast<VC*>(ctx->stage.ptr));
    if (p.d_vc.bytes < bytes) {
      CTSIM_HIP(hipStreamSynchronize(ctx->stream));
      CTSIM_CHECK(p.d_vc.reserve(bytes));
    }
    CTSIM_CHECK(ctx->stage.upload(p.d_vc, bytes, ctx->stream));
    p.cachedAngles.assign(h_angles, h_angles + nView);
    p.cachedKind = kind;
  }
  *out = p.d_vc.ptr;
  return CTSIM_OK;
}

}  // namespace

int bp_upload_tables(BpPlan& p) {
  if (!p.r.empty()) {
    size_t n = p.r.size();
    CTSIM_CHECK(p.d_rphi.reserve(2 * n * sizeof(double)));
    CTSIM_HIP(hipMemcpy(p.d_rphi.ptr, p.r.data(), n * sizeof(double), hipMemcpyHostToDevice));
    CTSIM_HIP(hipMemcpy(p.d_rphi.as<double>() + n, p.phi.data(), n * sizeof(double),
                        hipMemcpyHostToDevice));
  }
  CTSIM_CHECK(p.d_xy.reserve((p.xs.size() + p.ys.size()) * sizeof(float)));
  CTSIM_HIP(hipMemcpy(p.d_xy.ptr, p.xs.data(), p.xs.size() * sizeof(float), hipMemcpyHostToDevice));
  CTSIM_HIP(hipMemcpy(p.d_xy.as<float>() + p.xs.size(), p.ys.data(), p.ys.size() * sizeof(float),
                      hipMemcpyHostToDevice));
  return CTSIM_OK;
}

int bp_launch(ctsim_ctx* ctx, const BpPlan& p, int precision, const void* d_filtered,
              const double* h_angles, int nView, float* d_image) {
  if (nView <= 0) return CTSIM_OK;
  const hipStream_t s = ctx->stream;
  BpArgs a{};
  a.nx = p.nx; a.ny = p.ny; a.nDet = p.nDet; a.iDetCenter = p.iDetCenter; a.nView = nView;
  a.fstride = p.nDet * p.interpFactor;
  a.tileX = p.tileX; a.tileY = p.tileY;
  a.tilesY = (p.ny + p.tileY - 1) / p.tileY;
  const int tilesX = (p.nx + p.tileX - 1) / p.tileX;
  const int grid = tilesX * a.tilesY;
  a.detInc = p.detInc; a.rotScale = p.rotScale; a.F = p.F; a.sdl = p.sdl;
  a.r = p.d_rphi.as<double>();
  a.phi = p.r.empty() ? nullptr : p.d_rphi.as<double>() + p.r.size();
  a.xs = p.d_xy.as<float>();
  a.ys = p.d_xy.as<float>() + p.xs.size();
  a.f = d_filtered;
  a.image = d_image;
  a.endScale = 1.f;

  // The FP64 index path serves the F64 precision, and also the mixed
  // precision where FP32 positions would change discrete decisions: nearest
  // interpolation (a rounding tie picks a different cell) and Diff (its
  // column walk decides which pixel a sample lands in).  The mixed variant
  // reads float filtered views.
  // A mixed launch whose full-row staging cannot fit one view in LDS (very
  // long rows with the source close to the image) also takes the index path.
  const bool f64 = precision == CTSIM_PREC_F64 || !bp_mixed_tiled(p);
  const bool inF32 = precision == CTSIM_PREC_MIXED;
  if (f64) {
    CTSIM_CHECK(view_consts<ViewConstF64>(
        ctx, p, 0, h_angles, nView,
        [&](ViewConstF64* o) { bp_view_consts_f64(p, h_angles, nView, o); }, &a.vc));
    if (p.kind == KIND_DIFF) {
      const int g = (p.nx + 63) / 64;
      if (inF32) launch_diff<float>(p.interp, a, g, s);
      else launch_diff<double>(p.interp, a, g, s);
      CTSIM_HIP(hipGetLastError());
      return CTSIM_OK;
    }
    a.W = p.windowF64;
    const int Wv = a.W > 0 ? a.W : p.nDet + 8;
    const size_t budget = 48 * 1024;
    int vc = static_cast<int>(budget / (Wv * sizeof(double)));
    if (vc < 1) vc = 1;
    if (vc > kMaxChunk) vc = kMaxChunk;
    a.VC = vc;
    const size_t lds = static_cast<size_t>(vc) * Wv * sizeof(double);
    if (lds > 64 * 1024) return fail(CTSIM_ERR_UNSUPPORTED, "backprojector: detector row too long for LDS");
    ctx->prof_begin();
    if (inF32) launch_f64_interp<float>(p.interp, p.kind, a, grid, lds, s);
    else launch_f64_interp<double>(p.interp, p.kind, a, grid, lds, s);
    ctx->prof_end();
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }

  const ViewConstF32* vc32 = nullptr;
  CTSIM_CHECK(bp_view_consts_mixed(ctx, p, h_angles, nView, &vc32));
  return bp_launch_mixed(ctx, p, p.mixed, static_cast<const float*>(d_filtered), vc32, nView,
                         d_image);
}

// The symmetric kernel applies to a square centred grid, views with a
// rotational symmetry (of the angles last uploaded, which must be these
// nView), P <= 2 and windows small enough that one base view's 4*S windows
// fit the LDS it stages into.
bool bp_sym_ok(const BpPlan& p, const MixedShape& m, int nView) {
  if (static_cast<int>(p.cachedAngles.size()) != nView || p.cachedKind != 1) return false;
  return bp_sym_ok_for(p, m, p.cachedSym);
}

bool bp_sym_ok_for(const BpPlan& p, const MixedShape& m, const SymViews& sv) {
  if (m.sym == 3 || !p.symImage || sv.S == 0 || m.P > 2 || m.W <= 0) return false;
  const size_t elem = p.interp == INTERP_CUBIC ? sizeof(float4) : sizeof(float2);
  if (m.sym == 2) return m.W <= dev::kPipeRow;  // the pipelined body stages rows of kPipeRow cells, two per lane
  return static_cast<size_t>(4 * sv.S) * m.W * elem <= sym_lds(m);
}

int bp_mirror_chunk(const BpPlan& p, const MixedShape& m, int NM, int nView) {
  const bool cubic = p.interp == INTERP_CUBIC;
  const int Wv = m.W > 0 ? m.W : p.nDet + 4;
  // one buffer resource spans every coefficient row
  if (static_cast<double>(nView) * p.nDet * (cubic ? 16 : 8) >= 2147483648.0) return 0;
  const int upw = dev::mirror_window_units(Wv, cubic);
  // a power of two up to 32 (a 64-view block of the kernel then takes an
  // even number of chunks) whose windows, copied in whole rounds of 256
  // units, fit one buffer
  int c = 1;
  while (2 * c <= m.chunk && 2 * c <= 32) c *= 2;
  while (c > 0 && (c * NM * upw + 255) / 256 * 256 * 16 > dev::kMirBufBytes) c /= 2;
  return c;
}

size_t sym_lds(const MixedShape& m) {
  return m.symKB > 0 && static_cast<size_t>(m.symKB) * 1024 < kSymLds ? static_cast<size_t>(m.symKB) * 1024
                                                                      : kSymLds;
}

bool bp_mixed_tiled(const BpPlan& p) {
  const size_t mixedView = static_cast<size_t>(p.mixed.W > 0 ? p.mixed.W : p.nDet + 4) *
                           (p.interp == INTERP_CUBIC ? sizeof(float4) : sizeof(float2));
  return p.kind != KIND_DIFF && p.interp != INTERP_NEAREST && mixedView <= kMixedLdsMax;
}

// The mixed path: interval coefficients of every view, then the tiled
// kernel, then (with view groups) the ordered combine.
int bp_launch_mixed(ctsim_ctx* ctx, const BpPlan& p, const MixedShape& m, const float* d_filtered,
                    const ViewConstF32* d_vc, int nView, float* d_image) {
  CTSIM_CHECK(bp_mixed_coef(ctx, p, d_filtered, nView));
  const double endScale = p.kind == KIND_TRIG ? p.rotScale : 1.0;
  return bp_mixed_tiles(ctx, p, m, ctx->scratch2.ptr, d_vc, nView, d_image, endScale, false);
}

// Interval coefficients of nView filtered float rows into ctx->scratch2.
int bp_mixed_coef(ctsim_ctx* ctx, const BpPlan& p, const float* d_filtered, int nView) {
  const size_t elem = p.interp == INTERP_CUBIC ? sizeof(float4) : sizeof(float2);
  CTSIM_CHECK(ctx->scratch2.reserve(static_cast<size_t>(nView) * p.nDet * elem));
  return bp_mixed_coef_into(ctx, p, d_filtered, nView, ctx->scratch2.ptr);
}

int bp_mixed_coef_into(ctsim_ctx* ctx, const BpPlan& p, const float* d_filtered, int nView,
                       void* d_coef) {
  const hipStream_t s = ctx->stream;
  const bool cubic = p.interp == INTERP_CUBIC;
  const bool lastIncl = p.kind == KIND_IDIFF;
  const int perRow = (p.nDet + 255) / 256;
  const int cgrid = nView * perRow;
  const int fstride = p.nDet * p.interpFactor;
  if (cubic)
    hipLaunchKernelGGL((coef_kernel<INTERP_CUBIC, false>), dim3(cgrid), dim3(256), 0, s, d_filtered,
                       fstride, p.nDet, nView, perRow, d_coef);
  else if (lastIncl)
    hipLaunchKernelGGL((coef_kernel<INTERP_LINEAR, true>), dim3(cgrid), dim3(256), 0, s, d_filtered,
                       fstride, p.nDet, nView, perRow, d_coef);
  else
    hipLaunchKernelGGL((coef_kernel<INTERP_LINEAR, false>), dim3(cgrid), dim3(256), 0, s, d_filtered,
                       fstride, p.nDet, nView, perRow, d_coef);
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

// The tiled kernel (and, with view groups, the ordered combine) over
// coefficients `coef`: image += endScale * sum, or image = endScale * sum
// when overwrite.
int bp_mixed_tiles(ctsim_ctx* ctx, const BpPlan& p, const MixedShape& m, const void* coef,
                   const ViewConstF32* d_vc, int nView, float* d_image, double endScale,
                   bool overwrite, const SymViews* views, float* partial) {
  const hipStream_t s = ctx->stream;
  const bool cubic = p.interp == INTERP_CUBIC;
  const size_t elem = cubic ? sizeof(float4) : sizeof(float2);
  MixedArgs a{};
  a.nx = p.nx; a.ny = p.ny; a.nDet = p.nDet; a.nView = nView;
  a.lanesY = m.lanesY;
  const int P = m.P, tx = 256 / m.lanesY, ty = m.lanesY * P;
  a.tilesX = (p.nx + tx - 1) / tx;
  a.tilesY = (p.ny + ty - 1) / ty;
  a.W = m.W > 0 ? m.W : p.nDet + 4;
  a.split = m.split < 1 ? 1 : (m.split > nView ? nView : m.split);
  a.viewsPerGroup = (nView + a.split - 1) / a.split;
  a.split = (nView + a.viewsPerGroup - 1) / a.viewsPerGroup;
  a.chunk = m.chunk < 1 ? 1 : (m.chunk > kMaxChunk ? kMaxChunk : m.chunk);
  a.center = static_cast<float>(p.iDetCenter);
  a.F = static_cast<float>(p.F);
  a.twoF = static_cast<float>(2 * p.F);
  a.endScale = static_cast<float>(endScale);
  a.overwrite = overwrite ? 1 : 0;
  const int gk = p.geom == GEOM_PARALLEL ? 0 : (p.geom == GEOM_EQUIANGULAR ? 1 : 2);
  a.posScale = gk == 2 ? static_cast<float>(p.sdl / p.detInc) : static_cast<float>(1.0 / p.detInc);
  const int deg = gk == 1 ? p.atanDeg : 0;
  if (deg > 0) for (int k = 0; k <= deg; k++) a.atan[k] = p.atanCoef[k];
  else a.atan[0] = static_cast<float>(1.0 / p.detInc);
  a.xs = p.d_xy.as<float>();
  a.ys = p.d_xy.as<float>() + p.xs.size();
  a.vc = d_vc;
  a.coef = coef;
  a.image = d_image;
  const size_t npix = static_cast<size_t>(p.nx) * p.ny;
  const SymViews sv = views ? *views : p.cachedSym;
  const bool sym = m.sym && (views ? bp_sym_ok_for(p, m, sv) : bp_sym_ok(p, m, nView));
  // the mirror body: the mirror partners of these views (the cached ones
  // must be of these nView angles)
  const bool known = views || (static_cast<int>(p.cachedAngles.size()) == nView && p.cachedKind == 1);
  const int mirNM = m.sym == 3 && known && sv.NM >= 2 && bp_mirror_chunk(p, m, sv.NM, nView) > 0 ? sv.NM : 1;
  a.domX = p.nx;
  a.domY = p.ny;
  if (mirNM > 1) {
    a.domX = (p.nx + 1) / 2;
    a.domY = mirNM == 4 ? (p.ny + 1) / 2 : p.ny;
    a.mapO = sv.mapO;
    a.tilesX = (a.domX + tx - 1) / tx;
    a.tilesY = (a.domY + ty - 1) / ty;
    a.chunk = bp_mirror_chunk(p, m, mirNM, nView);
  }
  if (sym) {
    // tiles of the quadrant; groups and chunks of base views
    a.half = p.nx / 2;
    a.Q = nView / sv.S;
    a.dirf = sv.dir;
    a.tilesX = (a.half + tx - 1) / tx;
    a.tilesY = (a.half + ty - 1) / ty;
    a.split = m.split < 1 ? 1 : (m.split > a.Q ? a.Q : m.split);
    a.viewsPerGroup = (a.Q + a.split - 1) / a.split;
    a.split = (a.Q + a.viewsPerGroup - 1) / a.viewsPerGroup;
    a.capW = m.sym == 2 ? 4 * sv.S * dev::kPipeRow : static_cast<int>(sym_lds(m) / elem);
  }
  // (the mirror kernel stages into static LDS buffers of its own)
  const size_t lds = sym ? static_cast<size_t>(a.capW) * elem
                         : (mirNM > 1 ? 0 : static_cast<size_t>(a.chunk) * a.W * elem);
  if (lds > kMixedLdsMax) return fail(CTSIM_ERR_UNSUPPORTED, "backprojector: detector row too long for LDS");
  if (a.split > 1) {
    if (partial) {
      a.partial = partial;
    } else {
      CTSIM_CHECK(ctx->scratch.reserve(static_cast<size_t>(a.split) * npix * sizeof(float)));
      a.partial = ctx->scratch.as<float>();
    }
  }
  const int grid = a.tilesX * a.tilesY * a.split;
  ctx->prof_begin();
  if (sym) {
    const bool pipe = m.sym == 2;
    if (cubic) launch_sym<INTERP_CUBIC>(pipe, gk, deg, sv.S, sv.K, P, a, grid, lds, s);
    else launch_sym<INTERP_LINEAR>(pipe, gk, deg, sv.S, sv.K, P, a, grid, lds, s);
  } else if (mirNM > 1) {
    if (cubic) launch_mirror<INTERP_CUBIC>(gk, deg, mirNM, P, a, grid, lds, s);
    else launch_mirror<INTERP_LINEAR>(gk, deg, mirNM, P, a, grid, lds, s);
  } else if (cubic) {
    launch_mixed<INTERP_CUBIC>(gk, deg, P, a, grid, lds, s);
  } else {
    launch_mixed<INTERP_LINEAR>(gk, deg, P, a, grid, lds, s);
  }
  ctx->prof_end();
  if (a.split > 1) {
    int g = static_cast<int>((npix + 255) / 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(bp_combine_kernel, dim3(g), dim3(256), 0, s, d_image, a.partial, npix, a.split,
                       endScale, a.overwrite);
  }
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

// Per-view FP32 constants of a mixed launch on the device (cached by angles).
int bp_view_consts_mixed(ctsim_ctx* ctx, const BpPlan& p, const double* h_angles, int nView,
                         const ViewConstF32** out) {
  const void* vc = nullptr;
  CTSIM_CHECK(view_consts<ViewConstF32>(
      ctx, p, 1, h_angles, nView,
      [&](ViewConstF32* o) {
        const MirrorViews mv = bp_view_consts_f32(p, h_angles, nView, o);
        p.cachedSym = bp_sym_views(h_angles, nView);
        p.cachedSym.NM = mv.NM;
        p.cachedSym.mapO = mv.mapO;
      },
      &vc));
  *out = static_cast<const ViewConstF32*>(vc);
  return CTSIM_OK;
}

// Times candidate launch shapes of the mixed kernel on the given views (into
// a scratch image) and keeps the fastest in the plan.
int bp_autotune(ctsim_ctx* ctx, BpPlan& p, const float* d_filtered, const double* h_angles,
                int nView, double* bestMs) {
  if (nView <= 0) return fail(CTSIM_ERR_INVALID, "autotune: no views");
  const bool f64 = p.kind == KIND_DIFF || p.interp == INTERP_NEAREST;
  if (f64) {  // these run on the FP64 index path; nothing to tune
    if (bestMs) *bestMs = 0;
    return CTSIM_OK;
  }
  const ViewConstF32* vc = nullptr;
  CTSIM_CHECK(bp_view_consts_mixed(ctx, p, h_angles, nView, &vc));
  DevBuf img;
  CTSIM_CHECK(img.reserve(static_cast<size_t>(p.nx) * p.ny * sizeof(float)));
  hipEvent_t e0, e1;
  CTSIM_HIP(hipEventCreate(&e0));
  CTSIM_HIP(hipEventCreate(&e1));
  const int elem = p.interp == INTERP_CUBIC ? 16 : 8;
  MixedShape best = p.mixed;
  float bestT = 1e30f;
  int rc = CTSIM_OK;
  // candidates: the plain kernel over lanesY x P x split x chunk, then the
  // symmetric kernel (when the views and the image allow it) over
  // lanesY x P x split (of the base views) x base views per chunk
  for (int sym = 0; sym <= 3 && rc == CTSIM_OK; sym++) {
    if (sym == 3 && p.cachedSym.NM < 2) continue;  // no mirror partners
    for (int lanesY : {8, 16, 32}) {
      for (int P : {1, 2, 4}) {
        if (lanesY * P > 2 * p.ny && P > 1) continue;
        const int tx = 256 / lanesY, ty = lanesY * P;
        const int W = bp_window_bound(p, tx, ty);
        const int Wv = W > 0 ? W : p.nDet + 4;
        MixedShape m;
        m.lanesY = lanesY; m.P = P; m.W = W; m.sym = sym;
        const bool rot = sym == 1 || sym == 2;
        if (rot && !bp_sym_ok(p, m, nView)) continue;  // not even in 62 KB
        const int views = rot ? nView / p.cachedSym.S : nView;
        const int nm = sym == 3 ? p.cachedSym.NM : 1;
        for (int split : {1, 2, 4, 8, 16, 32}) {
          if (split > 1 && views / split < (rot ? 4 : 8)) continue;
          if (split == 32 && sym != 3) continue;  // the mirror kernel's quarter domain has few tiles
          for (int chunk : {1, 2, 3, 4, 8, 16, 32}) {
            if (sym == 3) {  // chunks that fit the mirror kernel's staging buffer
              m.chunk = chunk;
              if (chunk == 3 || chunk > 16 || bp_mirror_chunk(p, m, nm, nView) != chunk) continue;
            } else if (chunk == 1 || chunk == 3 ||
                       (rot ? chunk > 16 : (chunk < 4 || static_cast<size_t>(chunk) * Wv * elem > kMixedLdsMax))) {
              continue;
            }
           for (int symKB : {16, 24, 32, 48, 62}) {
            if (sym != 1 && symKB != 62) continue;  // only the chunked body stages into symKB
            if (sym == 2 && chunk != 2) continue;   // the pipelined body has no chunk
            m.symKB = symKB;
            if (rot && !bp_sym_ok(p, m, nView)) continue;
            m.split = split; m.chunk = chunk;
            float ms = 0, tot = 0;
            for (int rep = 0; rep < 3 && rc == CTSIM_OK; rep++) {
              if ((rc = bp_launch_mixed(ctx, p, m, d_filtered, vc, nView, img.as<float>())) != CTSIM_OK) break;
              if (rep == 0) continue;  // warm-up
              (void)hipEventRecord(e0, ctx->stream);
              if ((rc = bp_launch_mixed(ctx, p, m, d_filtered, vc, nView, img.as<float>())) != CTSIM_OK) break;
              (void)hipEventRecord(e1, ctx->stream);
              (void)hipEventSynchronize(e1);
              (void)hipEventElapsedTime(&ms, e0, e1);
              tot += ms;
            }
            if (rc != CTSIM_OK) break;
            if (tot / 2 < bestT) { bestT = tot / 2; best = m; }
           }
           if (rc != CTSIM_OK) break;
          }
          if (rc != CTSIM_OK) break;
        }
        if (rc != CTSIM_OK) break;
      }
      if (rc != CTSIM_OK) break;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != CTSIM_OK) return rc;
  p.mixed = best;
  p.mixedTuned = true;
  if (bestMs) *bestMs = bestT;
  return CTSIM_OK;
}

int bp_launch_post(ctsim_ctx* ctx, const BpPlan& p, float* d_image) {
  if (p.kind == KIND_TRIG) return CTSIM_OK;  // Trig scales per view (backprojectors.cpp:324-328)
  const size_t n = static_cast<size_t>(p.nx) * p.ny;
  int grid = static_cast<int>((n + 255) / 256);
  if (grid > 4096) grid = 4096;
  hipLaunchKernelGGL(bp_scale_kernel, dim3(grid), dim3(256), 0, ctx->stream, d_image, n, p.rotScale);
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

}  // namespace ctsim

using namespace ctsim;

extern "C" {

int ctsim_backprojector_create(ctsim_ctx* ctx, ctsim_backprojector** out, int geometry, int nDet,
                               double detInc, int nView, double viewDiameter, double focalLength,
                               double sourceDetectorLength, const char* backprojName,
                               const char* interpName, int interpFactor, int nx, int ny,
                               const double* roi) {
  if (!ctx || !out) return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_create: NULL argument");
  *out = nullptr;
  CTSIM_CHECK(ctx->activate());
  auto* bp = new ctsim_backprojector();
  int rc = bp_plan_init(bp->plan, geometry, nDet, detInc, nView, viewDiameter, focalLength,
                        sourceDetectorLength, backprojName, interpName, interpFactor, nx, ny, roi);
  if (rc == CTSIM_OK) rc = bp_upload_tables(bp->plan);
  if (rc != CTSIM_OK) {
    delete bp;
    return rc;
  }
  *out = bp;
  return CTSIM_OK;
}

void ctsim_backprojector_destroy(ctsim_backprojector* bp) { delete bp; }

int ctsim_backproject_views(ctsim_ctx* ctx, const ctsim_backprojector* bp, int precision,
                            const void* d_filtered, const double* h_angles, int nView,
                            float* d_image) {
  if (!ctx || !bp || (!d_filtered && nView > 0) || (!h_angles && nView > 0) || !d_image)
    return fail(CTSIM_ERR_INVALID, "ctsim_backproject_views: NULL argument");
  if (precision != CTSIM_PREC_F64 && precision != CTSIM_PREC_MIXED)
    return fail(CTSIM_ERR_INVALID, "ctsim_backproject_views: unknown precision");
  if (nView < 0) return fail(CTSIM_ERR_INVALID, "ctsim_backproject_views: nView < 0");
  CTSIM_CHECK(ctx->activate());
  return bp_launch(ctx, bp->plan, precision, d_filtered, h_angles, nView, d_image);
}

int ctsim_backproject_postprocess(ctsim_ctx* ctx, const ctsim_backprojector* bp, float* d_image) {
  if (!ctx || !bp || !d_image)
    return fail(CTSIM_ERR_INVALID, "ctsim_backproject_postprocess: NULL argument");
  CTSIM_CHECK(ctx->activate());
  return bp_launch_post(ctx, bp->plan, d_image);
}

int ctsim_backprojector_autotune(ctsim_ctx* ctx, ctsim_backprojector* bp, const float* d_filtered,
                                 const double* h_angles, int nView, int* shape7, double* bestMs) {
  if (!ctx || !bp || !d_filtered || !h_angles)
    return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_autotune: NULL argument");
  CTSIM_CHECK(ctx->activate());
  CTSIM_CHECK(bp_autotune(ctx, bp->plan, d_filtered, h_angles, nView, bestMs));
  if (shape7) {
    const MixedShape& m = bp->plan.mixed;
    shape7[0] = m.lanesY; shape7[1] = m.P; shape7[2] = m.split; shape7[3] = m.chunk; shape7[4] = m.W;
    shape7[5] = m.sym; shape7[6] = m.sym == 1 || m.sym == 2 ? static_cast<int>(sym_lds(m) / 1024) : 0;
  }
  return CTSIM_OK;
}

int ctsim_backprojector_set_shape(ctsim_backprojector* bp, const int* shape6) {
  if (!bp || !shape6) return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_set_shape: NULL argument");
  const int lanesY = shape6[0], P = shape6[1], split = shape6[2], chunk = shape6[3], sym = shape6[4],
            symKB = shape6[5];
  if (!(lanesY == 8 || lanesY == 16 || lanesY == 32 || lanesY == 64) || !(P == 1 || P == 2 || P == 4) ||
      split < 1 || split > 64 || chunk < 1 || chunk > kMaxChunk || sym < 0 || sym > 3 || symKB < 0 ||
      symKB > 62)
    return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_set_shape: shape out of range");
  MixedShape m;
  m.lanesY = lanesY; m.P = P; m.split = split; m.chunk = chunk; m.sym = sym; m.symKB = symKB;
  m.W = bp_window_bound(bp->plan, 256 / lanesY, lanesY * P);
  const int Wv = m.W > 0 ? m.W : bp->plan.nDet + 4;
  if (static_cast<size_t>(chunk) * Wv * (bp->plan.interp == INTERP_CUBIC ? 16 : 8) > kMixedLdsMax)
    return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_set_shape: staging exceeds 64 KB of LDS");
  bp->plan.mixed = m;
  return CTSIM_OK;
}

int ctsim_fbp_slice(ctsim_ctx* ctx, ctsim_procsignal* ps, const ctsim_backprojector* bp,
                    int precision, const float* d_sino, const double* h_angles, int nView,
                    float* d_image) {
  if (!ctx || !ps || !bp || !d_image || (nView > 0 && (!d_sino || !h_angles)))
    return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice: NULL argument");
  if (precision != CTSIM_PREC_F64 && precision != CTSIM_PREC_MIXED)
    return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice: unknown precision");
  if (nView < 0) return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice: nView < 0");
  FilterPlan& f = ps->plan;
  const BpPlan& p = bp->plan;
  if (f.n != p.nDet || f.n * f.preinterp != p.nDet * p.interpFactor)
    return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice: filter and backprojector detector counts differ");
  CTSIM_CHECK(ctx->activate());
  const hipStream_t s = ctx->stream;
  const size_t npix = static_cast<size_t>(p.nx) * p.ny;
  const bool mixed = precision == CTSIM_PREC_MIXED;
  if (!mixed || !bp_mixed_tiled(p) || p.interpFactor != 1 || nView == 0) {
    // The separate steps (pjrec: zeroed image, filterSignal, BackprojectView
    // per view, PostProcessing).
    int g = static_cast<int>((npix + 255) / 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(bp_fill_kernel, dim3(g), dim3(256), 0, s, d_image, npix, 0.f);
    CTSIM_HIP(hipGetLastError());
    if (nView == 0) return CTSIM_OK;
    const int stride = f.n * f.preinterp;
    CTSIM_CHECK(ctx->scratch3.reserve(static_cast<size_t>(nView) * stride *
                                      (mixed ? sizeof(float) : sizeof(double))));
    CTSIM_CHECK(filter_launch(ctx, f, precision, d_sino, nView, ctx->scratch3.ptr, stride));
    CTSIM_CHECK(bp_launch(ctx, p, precision, ctx->scratch3.ptr, h_angles, nView, d_image));
    return bp_launch_post(ctx, p, d_image);
  }
  // Mixed and tiled: the filter writes interval coefficients (FFT methods
  // straight from LDS; convolution through float rows), then the tiled
  // kernel overwrites the image with rotScale * sum.
  const bool cubic = p.interp == INTERP_CUBIC;
  const size_t elem = cubic ? sizeof(float4) : sizeof(float2);
  if (filter_can_coef(f)) {
    CTSIM_CHECK(ctx->scratch2.reserve(static_cast<size_t>(nView) * p.nDet * elem));
    CTSIM_CHECK(filter_launch_coef(ctx, f, d_sino, nView, ctx->scratch2.ptr, p.interp,
                                   p.kind == KIND_IDIFF));
  } else {
    CTSIM_CHECK(ctx->scratch3.reserve(static_cast<size_t>(nView) * f.n * sizeof(float)));
    CTSIM_CHECK(filter_launch(ctx, f, precision, d_sino, nView, ctx->scratch3.ptr, f.n));
    CTSIM_CHECK(bp_mixed_coef(ctx, p, ctx->scratch3.as<float>(), nView));
  }
  const ViewConstF32* vc = nullptr;
  CTSIM_CHECK(bp_view_consts_mixed(ctx, p, h_angles, nView, &vc));
  return bp_mixed_tiles(ctx, p, p.mixed, ctx->scratch2.ptr, vc, nView, d_image, p.rotScale, true);
}

int ctsim_backprojector_mirror_members(const ctsim_backprojector* bp, const double* h_angles, int nView) {
  if (!bp || !h_angles || nView < 1)
    return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_mirror_members: NULL argument or no views");
  std::vector<ViewConstF32> vc(nView);
  return bp_view_consts_f32(bp->plan, h_angles, nView, vc.data()).NM;
}

int ctsim_backprojector_info(const ctsim_backprojector* bp, double* o) {
  if (!bp || !o) return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_info: NULL argument");
  const BpPlan& p = bp->plan;
  o[0] = p.xMin; o[1] = p.xMax; o[2] = p.yMin; o[3] = p.yMax;
  o[4] = p.xInc; o[5] = p.yInc; o[6] = p.rotScale; o[7] = p.iDetCenter;
  return CTSIM_OK;
}

}  // extern "C"
