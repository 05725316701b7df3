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

// FP64 parity kernels keep the reference's rounding: no contraction.
#pragma clang fp contract(off)

namespace ctsim {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxChunk = 64;

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

// Remap block ids so that neighbouring tiles run on the same XCD.  The
// dispatcher deals workgroups round-robin over the 8 XCDs, so XCD k receives
// block ids k, k+8, ...; give it the contiguous tile range
// [k*q + min(k, r), ...) where n = 8q + r (a bijection on [0, n)).
__device__ __forceinline__ int xcd_tile(int bid, int n) {
  const int q = n >> 3, r = n & 7;
  const int k = bid & 7, local = bid >> 3;
  return k * q + min(k, r) + local;
}

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
      return a.iDetCenter + (r * cos(c.theta - phi)) / a.detInc;
    } else {
      double s, co;
      sincos(c.theta - phi, &s, &co);
      const double rc = r * co, rs = r * s;
      if constexpr (KIND == KIND_EQUIANGULAR) {
        return a.iDetCenter + atan(rc / (a.F + rs)) / a.detInc;
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
          const double L = r * cos(c.theta - phi);
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
          sincos(c.theta - phi, &s, &co);
          const double rc = r * co, rs = r * s;
          double pos, wgt;
          if constexpr (KIND == KIND_EQUIANGULAR) {
            const double fl = a.F + rs;
            const double gamma = atan(rc / fl);
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

// Same interpolation rules in FP32 (mixed path).
__device__ __forceinline__ float cubic_f32(const float* w, int lo, int Wv, int n, float x) {
  const float fl = floorf(x);
  const int l = static_cast<int>(fl) - 1;
  const int hi = l + 3;
  if (l < -1 || hi > n) return 0.f;
  if (l == -1) {
    const float y0 = wget(w, lo, Wv, 0), y1 = wget(w, lo, Wv, 1);
    return y0 + x * (y1 - y0);
  }
  if (hi == n) {
    const float frac = x - fl;
    const float a = wget(w, lo, Wv, n - 2), b = wget(w, lo, Wv, n - 1);
    return a + frac * (b - a);
  }
  // Lagrange basis on nodes l..l+3 with t = x - (l+1) in [0,1).
  const float t = x - fl;
  const float tm1 = t - 1.f, tp1 = t + 1.f, tm2 = t - 2.f;
  const float c0 = -(1.f / 6.f) * t * tm1 * tm2;
  const float c1 = 0.5f * tp1 * tm1 * tm2;
  const float c2 = -0.5f * tp1 * t * tm2;
  const float c3 = (1.f / 6.f) * tp1 * t * tm1;
  const float* p = w + (l - lo);
  const int k = l - lo;
  if (k < 0 || k + 3 >= Wv) {
    return c0 * wget(w, lo, Wv, l) + c1 * wget(w, lo, Wv, l + 1) + c2 * wget(w, lo, Wv, l + 2) +
           c3 * wget(w, lo, Wv, l + 3);
  }
  return c0 * p[0] + c1 * p[1] + c2 * p[2] + c3 * p[3];
}

// atan for FP32: odd polynomial t*P(t^2) on [0,1] (weighted least-squares fit
// of atan(t)/t in t^2, max abs error 1.3e-7), with atan(t) = pi/2 - atan(1/t)
// beyond 1.
template <bool SMALL_T>
__device__ __forceinline__ float atan_f32(float t) {
  if constexpr (SMALL_T) {  // |t| <= 1 is guaranteed by the geometry
    const float z = t * t;
    float p = 0.0028499013278633356f;
    p = p * z - 0.01606869138777256f;
    p = p * z + 0.042691655457019806f;
    p = p * z - 0.07504309713840485f;
    p = p * z + 0.10640943795442581f;
    p = p * z - 0.14203648269176483f;
    p = p * z + 0.1999261975288391f;
    p = p * z - 0.3333307206630707f;
    p = p * z + 1.0f;
    return t * p;
  }
  const float at = fabsf(t);
  const bool inv = at > 1.f;
  const float u = inv ? __builtin_amdgcn_rcpf(at) : at;
  const float z = u * u;
  float p = 0.0028499013278633356f;
  p = p * z - 0.01606869138777256f;
  p = p * z + 0.042691655457019806f;
  p = p * z - 0.07504309713840485f;
  p = p * z + 0.10640943795442581f;
  p = p * z - 0.14203648269176483f;
  p = p * z + 0.1999261975288391f;
  p = p * z - 0.3333307206630707f;
  p = p * z + 1.0f;
  float a = u * p;
  a = inv ? 1.5707963267948966f - a : a;
  return copysignf(a, t);
}


// Detector position and weight of one pixel in one view (mixed path).
// ca, sa are cos/detInc, sin/detInc (parallel) or cos, sin (fan).
template <int GK, bool SMALL_T>
__device__ __forceinline__ void pos_f32(float center, float F, float posScale, float ca, float sa,
                                        float x, float y, float& pos, float& wgt) {
  if constexpr (GK == 0) {
    pos = fmaf(y, sa, fmaf(x, ca, center));
    wgt = 1.f;
  } else {
    const float rc = fmaf(y, sa, x * ca);
    const float fl = fmaf(x, sa, fmaf(-y, ca, F));  // F + r sin(beta - phi)
    const float q = __builtin_amdgcn_rcpf(fl);
    if constexpr (GK == 1) {
      pos = fmaf(atan_f32<SMALL_T>(rc * q), posScale, center);
      wgt = __builtin_amdgcn_rcpf(fmaf(fl, fl, rc * rc));
    } else {
      pos = fmaf(rc * q, posScale, center);
      const float u = F * q;
      wgt = u * u;
    }
  }
}

// Cubic through y0..y3 at nodes -1, 0, 1, 2, evaluated at t in [0, 1), in
// Horner form (the same polynomial as the reference's Lagrange form).
__device__ __forceinline__ float cubic_horner(float y0, float y1, float y2, float y3, float t) {
  const float a2 = fmaf(0.5f, y0 + y2, -y1);
  const float a3 = fmaf(1.f / 6.f, y3 - y0, 0.5f * (y1 - y2));
  const float a1 = fmaf(-1.f / 6.f, y3, fmaf(-0.5f, y1, fmaf(-1.f / 3.f, y0, y2)));
  return fmaf(t, fmaf(t, fmaf(t, a3, a2), a1), y1);
}

template <int GK, int INTERP, bool LAST_INCL, bool SMALL_T>
__global__ __launch_bounds__(kBlock) void bp_f32_kernel(BpArgs a) {
  extern __shared__ float s_g[];
  __shared__ int s_lo2[kMaxChunk];
  __shared__ int s_in[kMaxChunk];
  __shared__ float2 s_cs[kMaxChunk];
  const int t = threadIdx.x;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int ix0 = (tile / a.tilesY) * a.tileX, iy0 = (tile % a.tilesY) * a.tileY;
  const int ix = ix0 + t / a.tileY, iy = iy0 + t % a.tileY;
  const bool active = (t < a.tileX * a.tileY) && ix < a.nx && iy < a.ny;
  const size_t pix = static_cast<size_t>(ix) * a.ny + iy;
  const int cx1 = min(ix0 + a.tileX - 1, a.nx - 1), cy1 = min(iy0 + a.tileY - 1, a.ny - 1);
  const int Wv = a.W > 0 ? a.W : a.nDet + 8;
  const ViewConstF32* vcs = static_cast<const ViewConstF32*>(a.vc);
  const float* frows = static_cast<const float*>(a.f);
  const float x = a.xs[active ? ix : ix0], y = a.ys[active ? iy : iy0];
  const int nDet = a.nDet;
  const float center = static_cast<float>(a.iDetCenter);
  const float F = static_cast<float>(a.F);
  const float posScale = GK == 1 ? static_cast<float>(1.0 / a.detInc)
                                 : static_cast<float>(a.sdl / a.detInc);
  float sum = 0.f, comp = 0.f;

  for (int v0 = 0; v0 < a.nView; v0 += a.VC) {
    const int nv = min(a.VC, a.nView - v0);
    if (t < nv) {
      const ViewConstF32 c = vcs[v0 + t];
      int lo = -4, interior = 0;
      if (a.W > 0) {
        float p0, p1, p2, p3, wd;
        pos_f32<GK, SMALL_T>(center, F, posScale, c.ca, c.sa, a.xs[ix0], a.ys[iy0], p0, wd);
        pos_f32<GK, SMALL_T>(center, F, posScale, c.ca, c.sa, a.xs[cx1], a.ys[iy0], p1, wd);
        pos_f32<GK, SMALL_T>(center, F, posScale, c.ca, c.sa, a.xs[ix0], a.ys[cy1], p2, wd);
        pos_f32<GK, SMALL_T>(center, F, posScale, c.ca, c.sa, a.xs[cx1], a.ys[cy1], p3, wd);
        float pmin = fminf(fminf(p0, p1), fminf(p2, p3));
        float pmax = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
        if (!(pmin > -1e9f && pmin < 1e9f)) pmin = -4.f;
        lo = static_cast<int>(floorf(pmin)) - 4;
        // every tap of every pixel is an interior Lagrange tap
        interior = (pmin >= 2.f && pmax <= static_cast<float>(nDet - 4) &&
                    pmax - pmin <= static_cast<float>(Wv - 10)) ? 1 : 0;
      }
      s_lo2[t] = lo;
      s_in[t] = interior;
      s_cs[t] = make_float2(c.ca, c.sa);
    }
    __syncthreads();
    for (int vv = 0; vv < nv; vv++) {
      const int lo = s_lo2[vv];
      const float* frow = frows + static_cast<size_t>(v0 + vv) * a.fstride;
      float* w = s_g + vv * Wv;
      for (int j = t; j < Wv; j += kBlock) {
        const int d = lo + j;
        w[j] = (d >= 0 && d < nDet) ? frow[d] : 0.f;
      }
    }
    __syncthreads();
    if (active) {
      for (int vv = 0; vv < nv; vv++) {
        const float2 cs = s_cs[vv];
        const float* w = s_g + vv * Wv;
        const int lo = __builtin_amdgcn_readfirstlane(s_lo2[vv]);
        float pos, wgt;
        pos_f32<GK, SMALL_T>(center, F, posScale, cs.x, cs.y, x, y, pos, wgt);
        float cval = 0.f;
        if constexpr (INTERP == INTERP_NEAREST) {
          const int i = static_cast<int>(floorf(pos + 0.5f));
          if (i >= 0 && i < nDet) cval = wget(w, lo, Wv, i);
        } else if constexpr (INTERP == INTERP_LINEAR) {
          const float pf = floorf(pos);
          const int i = static_cast<int>(pf);
          const float fr = pos - pf;
          const int hiOk = LAST_INCL ? nDet - 1 : nDet - 2;
          if (i >= 0 && i <= hiOk) {
            const float f0 = wget(w, lo, Wv, i);
            const float f1 = (i < nDet - 1) ? wget(w, lo, Wv, i + 1) : f0;
            cval = fmaf(fr, f1 - f0, f0);
          }
        } else {
          if (__builtin_amdgcn_readfirstlane(s_in[vv])) {
            const float pf = floorf(pos);
            int k = static_cast<int>(pf) - 1 - lo;
            k = min(max(k, 0), Wv - 4);
            cval = cubic_horner(w[k], w[k + 1], w[k + 2], w[k + 3], pos - pf);
          } else {
            cval = cubic_f32(w, lo, Wv, nDet, pos);
          }
        }
        // Kahan-compensated accumulation across views
        const float yk = cval * wgt - comp;
        const float tk = sum + yk;
        comp = (tk - sum) - yk;
        sum = tk;
      }
    }
    __syncthreads();
  }
  if (active) a.image[pix] = a.image[pix] + a.endScale * (sum - comp);
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

template <int GK, int INTERP>
void launch_f32(bool lastIncl, bool smallT, const BpArgs& a, int grid, size_t lds, hipStream_t s) {
  if (GK == 1 && smallT) {
    if (lastIncl)
      hipLaunchKernelGGL((bp_f32_kernel<GK, INTERP, true, true>), dim3(grid), dim3(kBlock), lds, s, a);
    else
      hipLaunchKernelGGL((bp_f32_kernel<GK, INTERP, false, true>), dim3(grid), dim3(kBlock), lds, s, a);
  } else {
    if (lastIncl)
      hipLaunchKernelGGL((bp_f32_kernel<GK, INTERP, true, false>), dim3(grid), dim3(kBlock), lds, s, a);
    else
      hipLaunchKernelGGL((bp_f32_kernel<GK, INTERP, false, false>), dim3(grid), dim3(kBlock), lds, s, a);
  }
}

template <int GK>
void launch_f32_interp(int interp, bool lastIncl, bool smallT, const BpArgs& a, int grid,
                       size_t lds, hipStream_t s) {
  switch (interp) {
    case INTERP_NEAREST: launch_f32<GK, INTERP_NEAREST>(lastIncl, smallT, a, grid, lds, s); break;
    case INTERP_LINEAR: launch_f32<GK, INTERP_LINEAR>(lastIncl, smallT, a, grid, lds, s); break;
    default: launch_f32<GK, INTERP_CUBIC>(lastIncl, smallT, a, grid, lds, s); break;
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
  const bool f64 = precision == CTSIM_PREC_F64 || p.kind == KIND_DIFF || p.interp == INTERP_NEAREST;
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
    if (inF32) launch_f64_interp<float>(p.interp, p.kind, a, grid, lds, s);
    else launch_f64_interp<double>(p.interp, p.kind, a, grid, lds, s);
    CTSIM_HIP(hipGetLastError());
    return CTSIM_OK;
  }

  CTSIM_CHECK(view_consts<ViewConstF32>(
      ctx, p, 1, h_angles, nView,
      [&](ViewConstF32* o) { bp_view_consts_f32(p, h_angles, nView, o); }, &a.vc));
  a.W = p.windowF32;
  if (p.kind == KIND_TRIG) a.endScale = static_cast<float>(p.rotScale);
  const int Wv = a.W > 0 ? a.W : p.nDet + 8;
  const size_t budget = 32 * 1024;
  int vc = static_cast<int>(budget / (Wv * sizeof(float)));
  if (vc < 1) vc = 1;
  if (vc > kMaxChunk) vc = kMaxChunk;
  a.VC = vc;
  const size_t lds = static_cast<size_t>(vc) * Wv * sizeof(float);
  if (lds > 64 * 1024) return fail(CTSIM_ERR_UNSUPPORTED, "backprojector: detector row too long for LDS");
  const bool lastIncl = p.kind == KIND_IDIFF;
  const int gk = p.geom == GEOM_PARALLEL ? 0 : (p.geom == GEOM_EQUIANGULAR ? 1 : 2);
  if (gk == 0) launch_f32_interp<0>(p.interp, lastIncl, p.smallTan, a, grid, lds, s);
  else if (gk == 1) launch_f32_interp<1>(p.interp, lastIncl, p.smallTan, a, grid, lds, s);
  else launch_f32_interp<2>(p.interp, lastIncl, p.smallTan, a, grid, lds, s);
  CTSIM_HIP(hipGetLastError());
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

int ctsim_backprojector_info(const ctsim_backprojector* bp, double* o) {
  if (!bp || !o) return fail(CTSIM_ERR_INVALID, "ctsim_backprojector_info: NULL argument");
  const BpPlan& p = bp->plan;
  o[0] = p.xMin; o[1] = p.xMax; o[2] = p.yMin; o[3] = p.yMax;
  o[4] = p.xInc; o[5] = p.yInc; o[6] = p.rotScale; o[7] = p.iDetCenter;
  return CTSIM_OK;
}

}  // extern "C"
