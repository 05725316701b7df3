// Fan-beam to parallel rebinning (Projections::interpolateToParallel,
// libctsim/projections.cpp:1250-1328, with ParallelRaysums, 1339-1469).
//
// Every fan sample is a parallel ray (theta, t).  The reference sorts all
// samples by (t, theta); detector g of the new sinogram is the g-th run of
// nView samples (one fan detector, all views, increasing theta).  It then
// interpolates each run linearly to evenly spaced view angles, and each new
// view linearly to evenly spaced detector positions, storing floats after
// each pass.  The sort depends only on the geometry and the view angles, so
// the plan does it once on the host (the reference's expressions, glibc
// trig, no contraction); the two interpolation passes run on the device, one
// lane per output sample, in FP64 with the reference's expression order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "ctsim_internal.h"

#pragma clang fp contract(off)

namespace ctsim {
namespace {

struct RebinPlan {
  int geom = 0, nDet = 0, nView = 0;
  double newRotInc = 0, newDetStart = 0, newDetInc = 0;
  std::vector<double> newAngles;
  DevBuf d_theta;   // [nDet][nView] sorted theta of each run
  DevBuf d_src;     // [nDet][nView] source sample (v * nDet + d) of each run entry
  DevBuf d_t;       // [nDet] t of each run
  DevBuf d_angles;  // [nView] new view angles
  DevBuf d_pos;     // [nDet] new detector positions
  DevBuf d_mid;     // [nView][nDet] float after the theta pass
};

// LinearInterpolator<double>::interpolate with abscissae and values clamped
// outside (include/interpolator.h:203-262); the search hint the reference
// passes does not change the result for increasing abscissae.
template <class YF>
__device__ __forceinline__ double lin_interp(const double* X, const YF& Y, int n, double dX) {
  int iLower = -1, iUpper = n;
  while (iUpper - iLower > 1) {
    const int iMiddle = (iUpper + iLower) >> 1;
    if (dX >= X[iMiddle]) iLower = iMiddle;
    else iUpper = iMiddle;
  }
  if (dX == X[0]) return Y(0);
  if (dX < X[0]) return Y(0);
  if (dX == X[n - 1]) return Y(n - 1);
  if (dX > X[n - 1]) return Y(n - 1);
  const double y0 = Y(iLower), y1 = Y(iUpper);
  return y0 + (y1 - y0) * ((dX - X[iLower]) / (X[iUpper] - X[iLower]));
}

// Pass 1: run g at new view v (projections.cpp:1285-1298).
__global__ __launch_bounds__(256) void rebin_theta_kernel(const float* __restrict__ in,
                                                          const double* __restrict__ theta,
                                                          const int* __restrict__ src,
                                                          const double* __restrict__ angles,
                                                          float* __restrict__ mid, int nDet,
                                                          int nView) {
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= static_cast<size_t>(nDet) * nView) return;
  const int g = static_cast<int>(k / nView), v = static_cast<int>(k % nView);
  const double* X = theta + static_cast<size_t>(g) * nView;
  const int* S = src + static_cast<size_t>(g) * nView;
  auto Y = [in, S](int i) { return static_cast<double>(in[S[i]]); };
  mid[static_cast<size_t>(v) * nDet + g] = static_cast<float>(lin_interp(X, Y, nView, angles[v]));
}

// Pass 2: new view v at new detector position d (projections.cpp:1300-1325).
__global__ __launch_bounds__(256) void rebin_t_kernel(const float* __restrict__ mid,
                                                      const double* __restrict__ t,
                                                      const double* __restrict__ pos,
                                                      float* __restrict__ out, int nDet, int nView) {
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= static_cast<size_t>(nDet) * nView) return;
  const int v = static_cast<int>(k / nDet), d = static_cast<int>(k % nDet);
  const float* row = mid + static_cast<size_t>(v) * nDet;
  auto Y = [row](int i) { return static_cast<double>(row[i]); };
  out[k] = static_cast<float>(lin_interp(t, Y, nDet, pos[d]));
}

}  // namespace
}  // namespace ctsim

using namespace ctsim;

struct ctsim_rebin {
  RebinPlan p;
};

extern "C" {

int ctsim_rebin_create(ctsim_ctx* ctx, ctsim_rebin** out, int geometry, int nDet, int nView,
                       double detStart, double detInc, double focalLength,
                       double sourceDetectorLength, double viewDiameter, double rotStart,
                       const double* h_angles) {
  if (!ctx || !out || !h_angles) return fail(CTSIM_ERR_INVALID, "ctsim_rebin_create: NULL argument");
  if (geometry != GEOM_EQUIANGULAR && geometry != GEOM_EQUILINEAR)
    return fail(CTSIM_ERR_INVALID, "ctsim_rebin_create: fan-beam geometry expected "
                                   "(parallel projections are returned unchanged by the reference)");
  if (nDet < 2 || nView < 2) return fail(CTSIM_ERR_INVALID, "ctsim_rebin_create: too few samples");
  *out = nullptr;
  CTSIM_CHECK(ctx->activate());
  auto* r = new ctsim_rebin();
  RebinPlan& p = r->p;
  p.geom = geometry;
  p.nDet = nDet;
  p.nView = nView;
  p.newRotInc = kTwoPi / nView;
  p.newDetStart = -viewDiameter / 2;
  p.newDetInc = viewDiameter / nDet;
  if (nDet % 2 == 0) p.newDetInc = viewDiameter / (nDet - 1);
  // ParallelRaysums coordinates, view-major (projections.cpp:1352-1379)
  const size_t nc = static_cast<size_t>(nView) * nDet;
  std::vector<double> th(nc), tt(nc);
  for (int iV = 0; iV < nView; iV++) {
    const double dViewAngle = h_angles[iV];
    double dDetPos = detStart;
    for (int iD = 0; iD < nDet; iD++) {
      const size_t k = static_cast<size_t>(iV) * nDet + iD;
      double theta, t;
      if (geometry == GEOM_EQUILINEAR) {
        const double dFanAngle = atan(dDetPos / sourceDetectorLength);
        theta = dViewAngle + dFanAngle;
        t = focalLength * sin(dFanAngle);
      } else {
        theta = dViewAngle + dDetPos;
        t = focalLength * sin(dDetPos);
      }
      while (theta < 0.) theta += kTwoPi;  // normalizeAngle, mathfuncs.cpp:72-81
      while (theta >= kTwoPi) theta -= kTwoPi;
      th[k] = theta;
      tt[k] = t;
      dDetPos += detInc;
    }
  }
  // compareByT (include/projections.h:50-52); exact ties keep view-major order
  std::vector<int> idx(nc);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (tt[a] != tt[b]) return tt[a] < tt[b];
    if (th[a] != th[b]) return th[a] < th[b];
    return a < b;
  });
  std::vector<double> theta(nc), tRun(nDet), pos(nDet);
  for (size_t k = 0; k < nc; k++) theta[k] = th[idx[k]];
  for (int g = 0; g < nDet; g++) tRun[g] = tt[idx[static_cast<size_t>(g) * nView]];
  p.newAngles.resize(nView);
  double a = rotStart;
  for (int v = 0; v < nView; v++, a += p.newRotInc) p.newAngles[v] = a;
  double x = p.newDetStart;
  for (int d = 0; d < nDet; d++, x += p.newDetInc) pos[d] = x;
  int rc = CTSIM_OK;
  auto up = [&](DevBuf& d, const void* src, size_t n) {
    if (rc != CTSIM_OK) return;
    if ((rc = d.reserve(n)) != CTSIM_OK) return;
    if (hipMemcpy(d.ptr, src, n, hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(CTSIM_ERR_HIP, "ctsim_rebin_create: upload failed");
  };
  up(p.d_theta, theta.data(), nc * sizeof(double));
  up(p.d_src, idx.data(), nc * sizeof(int));
  up(p.d_t, tRun.data(), nDet * sizeof(double));
  up(p.d_angles, p.newAngles.data(), nView * sizeof(double));
  up(p.d_pos, pos.data(), nDet * sizeof(double));
  if (rc == CTSIM_OK) rc = p.d_mid.reserve(nc * sizeof(float));
  if (rc != CTSIM_OK) {
    delete r;
    return rc;
  }
  *out = r;
  return CTSIM_OK;
}

void ctsim_rebin_destroy(ctsim_rebin* r) { delete r; }

int ctsim_rebin_info(const ctsim_rebin* r, double* out4, double* h_newAngles) {
  if (!r || !out4) return fail(CTSIM_ERR_INVALID, "ctsim_rebin_info: NULL argument");
  out4[0] = 0;  // rotStart of the new projections (projections.cpp:1272)
  out4[1] = r->p.newRotInc;
  out4[2] = r->p.newDetStart;
  out4[3] = r->p.newDetInc;
  if (h_newAngles) std::copy(r->p.newAngles.begin(), r->p.newAngles.end(), h_newAngles);
  return CTSIM_OK;
}

int ctsim_rebin_run(ctsim_ctx* ctx, ctsim_rebin* r, const float* d_in, float* d_out) {
  if (!ctx || !r || !d_in || !d_out) return fail(CTSIM_ERR_INVALID, "ctsim_rebin_run: NULL argument");
  CTSIM_CHECK(ctx->activate());
  RebinPlan& p = r->p;
  const size_t n = static_cast<size_t>(p.nDet) * p.nView;
  const unsigned grid = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(rebin_theta_kernel, dim3(grid), dim3(256), 0, ctx->stream, d_in,
                     p.d_theta.as<double>(), p.d_src.as<int>(), p.d_angles.as<double>(),
                     p.d_mid.as<float>(), p.nDet, p.nView);
  hipLaunchKernelGGL(rebin_t_kernel, dim3(grid), dim3(256), 0, ctx->stream, p.d_mid.as<float>(),
                     p.d_t.as<double>(), p.d_pos.as<double>(), d_out, p.nDet, p.nView);
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

}  // extern "C"
