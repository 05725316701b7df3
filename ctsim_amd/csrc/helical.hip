// Helical 180LI interpolation and half-scan feathering on the device
// (Projections::Helical180LI_Equiangular / HalfScanFeather,
// libctsim/projections.cpp:231-439).
//
// The reference scatters every acquired sample into the interpolated view it
// maps to (its own view below pi + fan, the conjugate ray's view above) and
// accumulates into float storage in loop order.  Here one lane owns one
// output sample and gathers its sources: the direct sample of the same view
// and detector, then the conjugate samples of the mirrored detector whose
// rounded view lands on this one, in increasing source view, which is the
// reference's accumulation order, so the float sums are bit-identical.
// Fan angles per detector come from a host table (the reference's integer
// detector centre; equilinear: atan(t / sdl), a derived geometry the
// reference does not implement).  Arithmetic is FP64 without contraction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "ctsim_internal.h"

#pragma clang fp contract(off)

namespace ctsim {
namespace {

struct HelixArgs {
  int nDet, nView, L, offsetView;
  double dbeta, fan;
  const double* gam;  // fan angle per detector
  const float* in;    // [nView][nDet]
  float* out;         // [L+1][nDet]
};

__device__ __forceinline__ int nearest_i(double x) {
  return x > 0 ? static_cast<int>(x + 0.5) : static_cast<int>(x - 0.5);
}

// Weight of an acquired sample (beta, gamma) in the reference's region chain
// (projections.cpp:306-352); false outside every region.
__device__ __forceinline__ bool region_weight(double beta, double gamma, double fan, double& w) {
  const double PI = kPi;
  if (!(beta > fan - 2 * gamma && beta < 2 * PI + fan - 2 * gamma)) return false;
  if (beta > fan - 2 * gamma && beta <= 2 * fan)
    w = (beta + 2 * gamma - fan) / (PI + 2 * gamma);
  else if (beta > 2 * fan && beta <= PI - 2 * gamma)
    w = (beta + 2 * gamma - fan) / (PI + 2 * gamma);
  else if (beta > PI - 2 * gamma && beta <= PI + fan)
    w = (beta + 2 * gamma - fan) / (PI + 2 * gamma);
  else if (beta > PI + fan && beta <= PI + 2 * fan - 2 * gamma)
    w = (2 * PI - beta - 2 * gamma + fan) / (PI - 2 * gamma);
  else if (beta > PI + 2 * fan - 2 * gamma && beta <= 2 * PI)
    w = (2 * PI - beta - 2 * gamma + fan) / (PI - 2 * gamma);
  else if (beta > 2 * PI && beta <= 2 * PI + fan - 2 * gamma)
    w = (2 * PI - beta - 2 * gamma + fan) / (PI - 2 * gamma);
  else
    return false;
  return true;
}

__global__ __launch_bounds__(256) void helical_kernel(HelixArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= static_cast<size_t>(a.L + 1) * a.nDet) return;
  const int nv = static_cast<int>(k / a.nDet), nd = static_cast<int>(k % a.nDet);
  const double PI = kPi, dbeta = a.dbeta, fan = a.fan;
  float acc = 0.f;
  // direct sample: acquired below pi + fan, kept in place
  {
    const int iView = nv;
    const double beta = iView * dbeta;
    const double gamma = a.gam[nd];
    double w;
    if (beta < PI + fan && iView + a.offsetView < a.nView && region_weight(beta, gamma, fan, w))
      acc += w * a.in[static_cast<size_t>(iView + a.offsetView) * a.nDet + nd];
  }
  // conjugate samples of the mirrored detector, in increasing source view
  {
    const int iDet = a.nDet - 1 - nd;
    const double gamma = a.gam[iDet];
    const double x0 = nv - (2 * gamma - PI) / dbeta;
    const int c0 = static_cast<int>(floor(x0));
    for (int iView = c0 - 2; iView <= c0 + 2; iView++) {
      if (iView < 0 || iView > 2 * a.L) continue;
      const double beta = iView * dbeta;
      if (beta < PI + fan) continue;
      if (nearest_i(((iView * dbeta) + 2 * gamma - PI) / dbeta) != nv) continue;
      if (iView + a.offsetView >= a.nView) continue;
      double w;
      if (region_weight(beta, gamma, fan, w))
        acc += w * a.in[static_cast<size_t>(iView + a.offsetView) * a.nDet + iDet];
    }
  }
  a.out[k] = acc;
}

struct FeatherArgs {
  int nDet, nView;
  double dbeta, fan, scale;
  const double* gam;
  const float* in;
  float* out;
  int* error;
};

// x of the Crawford-King feathering weight at the conjugate sample
// (projections.cpp:405-418); false outside every region.
__device__ __forceinline__ bool feather_x(double beta1, double gamma1, double fan, double& x) {
  const double PI = kPi;
  if (beta1 <= (fan - 2 * gamma1))
    x = beta1 / (fan - 2 * gamma1);
  else if ((fan - 2 * gamma1 <= beta1) && beta1 <= PI - 2 * gamma1)
    x = 1;
  else if ((PI - 2 * gamma1 <= beta1) && (beta1 <= PI + fan))
    x = (PI + fan - beta1) / (fan + 2 * gamma1);
  else
    return false;
  return true;
}

__global__ __launch_bounds__(256) void feather_kernel(FeatherArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= static_cast<size_t>(a.nView) * a.nDet) return;
  const int iv = static_cast<int>(k / a.nDet), id = static_cast<int>(k % a.nDet);
  const double PI = kPi, dbeta = a.dbeta, fan = a.fan;
  float v = a.in[k];
  const double gamma = a.gam[id];
  if (iv * dbeta >= PI - 2 * gamma) {
    // redundant sample: keeps w2 of itself
    const int iView1 = nearest_i(((iv * dbeta) + 2 * gamma - PI) / dbeta);
    double x;
    if (!feather_x(iView1 * dbeta, -gamma, fan, x)) {
      atomicOr(a.error, 1);
    } else {
      const double w1 = (3 * x - 2 * x * x) * x;
      const double w2 = 1 - w1;
      v *= w2;
    }
  } else {
    // possibly the conjugate of redundant samples of the mirrored detector:
    // the last one in the reference's loop order wins (a copy, then w1)
    const int iDet2 = a.nDet - 1 - id;
    const double gamma2 = a.gam[iDet2];
    const double x0 = iv - (2 * gamma2 - PI) / dbeta;
    const int c0 = static_cast<int>(floor(x0));
    for (int iView2 = c0 - 2; iView2 <= c0 + 2; iView2++) {
      if (iView2 < 0 || iView2 >= a.nView) continue;
      if (!(iView2 * dbeta >= PI - 2 * gamma2)) continue;
      if (nearest_i(((iView2 * dbeta) + 2 * gamma2 - PI) / dbeta) != iv) continue;
      double x;
      if (!feather_x(iv * dbeta, -gamma2, fan, x)) {
        atomicOr(a.error, 1);
        continue;
      }
      const double w1 = (3 * x - 2 * x * x) * x;
      v = a.in[static_cast<size_t>(iView2) * a.nDet + iDet2];
      v *= w1;
    }
  }
  v *= a.scale;
  a.out[k] = v;
}

// The detector angles of a scan, kept in the context: consecutive calls on
// one scan (a slice position after another) upload them once.
int gamma_table(ctsim_ctx* ctx, int geom, int nDet, double detInc, double sdl, const double** out) {
  if (!(ctx->helixGam.ptr && ctx->helixGeom == geom && ctx->helixNDet == nDet && ctx->helixDetInc == detInc &&
        ctx->helixSdl == sdl)) {
    ctx->helixGeom = -1;  // invalid until the upload below has finished
    std::vector<double> g(nDet);
    for (int i = 0; i < nDet; i++)
      g[i] = geom == GEOM_EQUILINEAR ? atan((i - (nDet - 1) / 2) * detInc / sdl)
                                     : (i - (nDet - 1) / 2) * detInc;
    // earlier launches on the stream may still read the old table
    CTSIM_HIP(hipStreamSynchronize(ctx->stream));
    CTSIM_CHECK(ctx->helixGam.reserve(nDet * sizeof(double)));
    CTSIM_HIP(hipMemcpyAsync(ctx->helixGam.ptr, g.data(), nDet * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
    // the host vector dies at return: wait for the copy
    CTSIM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->helixGeom = geom;
    ctx->helixNDet = nDet;
    ctx->helixDetInc = detInc;
    ctx->helixSdl = sdl;
  }
  *out = ctx->helixGam.as<double>();
  return CTSIM_OK;
}

}  // namespace
}  // namespace ctsim

using namespace ctsim;

extern "C" {

int ctsim_helical180li(ctsim_ctx* ctx, int geometry, int nDet, int nView, double rotInc,
                       double detInc, double fanBeamAngle, double sourceDetectorLength,
                       int interpView, const float* d_in, float* d_out, int outCapacity,
                       int* nViewOut, double* h_anglesOut) {
  if (!ctx || !d_in || !d_out || !nViewOut)
    return fail(CTSIM_ERR_INVALID, "ctsim_helical180li: NULL argument");
  if (geometry != GEOM_EQUIANGULAR && geometry != GEOM_EQUILINEAR)
    return fail(CTSIM_ERR_UNSUPPORTED, "Helical 180LI is implemented for fan-beam geometries only");
  if (nDet < 1 || nView < 1 || !(rotInc > 0))
    return fail(CTSIM_ERR_INVALID, "ctsim_helical180li: empty or invalid scan");
  const double PI = kPi, dbeta = rotInc, fan = fanBeamAngle;
  // projections.cpp:240-266
  if (nView < static_cast<int>((2 * (PI + fan)) / dbeta) - 1)
    return fail(CTSIM_ERR_INVALID, "Data set does not include 360 +2*FanBeamAngle views");
  int offsetView = 0;
  if (interpView < 0) {
    interpView = static_cast<int>((PI + fan) / dbeta);
  } else {
    if (interpView * dbeta < PI + fan || interpView * dbeta + PI + fan > nView * dbeta)
      return fail(CTSIM_ERR_INVALID, "There isn't PI+fanAngle of data on either side of the "
                                     "requested interpolation view");
    offsetView = interpView - static_cast<int>((PI + fan) / dbeta);
  }
  const int L = static_cast<int>((PI + fan) / dbeta);
  if (L + 1 > outCapacity) return fail(CTSIM_ERR_INVALID, "ctsim_helical180li: output too small");
  CTSIM_CHECK(ctx->activate());
  const double* gam = nullptr;
  CTSIM_CHECK(gamma_table(ctx, geometry, nDet, detInc, sourceDetectorLength, &gam));
  HelixArgs a{nDet, nView, L, offsetView, dbeta, fan, gam, d_in, d_out};
  const size_t n = static_cast<size_t>(L + 1) * nDet;
  hipLaunchKernelGGL(helical_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, a);
  CTSIM_HIP(hipGetLastError());
  *nViewOut = L + 1;
  if (h_anglesOut)
    for (int i = 0; i <= L; i++) h_anglesOut[i] = (i + offsetView) * dbeta;
  return CTSIM_OK;
}

int ctsim_halfscan_feather(ctsim_ctx* ctx, int geometry, int nDet, int nView, double rotInc,
                           double detInc, double fanBeamAngle, double sourceDetectorLength,
                           const float* d_in, float* d_out) {
  if (!ctx || !d_in || !d_out) return fail(CTSIM_ERR_INVALID, "ctsim_halfscan_feather: NULL argument");
  if (d_in == d_out) return fail(CTSIM_ERR_INVALID, "ctsim_halfscan_feather: needs separate buffers");
  if (geometry != GEOM_EQUIANGULAR && geometry != GEOM_EQUILINEAR)
    return fail(CTSIM_ERR_UNSUPPORTED, "HalfScanFeather is implemented for fan-beam geometries only");
  const double PI = kPi;
  if (nView != static_cast<int>((PI + fanBeamAngle) / rotInc) + 1)
    return fail(CTSIM_ERR_INVALID, "Data set does seem have enough data to be a halfscan data set");
  CTSIM_CHECK(ctx->activate());
  const double* gam = nullptr;
  CTSIM_CHECK(gamma_table(ctx, geometry, nDet, detInc, sourceDetectorLength, &gam));
  CTSIM_CHECK(ctx->helixErr.reserve(sizeof(int)));
  int zero = 0;
  CTSIM_HIP(hipMemcpyAsync(ctx->helixErr.ptr, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  FeatherArgs a{nDet, nView, rotInc, fanBeamAngle, nView * rotInc / PI, gam, d_in, d_out,
                ctx->helixErr.as<int>()};
  const size_t n = static_cast<size_t>(nView) * nDet;
  hipLaunchKernelGGL(feather_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, a);
  CTSIM_HIP(hipGetLastError());
  int flag = 0;
  CTSIM_HIP(hipMemcpyAsync(&flag, ctx->helixErr.ptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  CTSIM_HIP(hipStreamSynchronize(ctx->stream));
  if (flag) return fail(CTSIM_ERR_INVALID, "HalfScanFeather: sample outside every weighting region");
  return CTSIM_OK;
}

}  // extern "C"
