// Phantom rasterisation on the device (Phantom::convertToImagefile,
// libctsim/phantom.cpp:463-538; the phm2if tool).
//
// Each pixel is the average over nsample x nsample sample points of the
// summed attenuation of the elements containing them.  One lane owns one
// pixel and walks the elements in the reference's list order (newest first),
// adding each element's double sum of sample attenuations into the float
// pixel, as the reference does; sample coordinates follow the reference's
// repeated additions (column and row starts from host tables built the same
// way), so the image is bit-identical.
#include <hip/hip_runtime.h>

#include <vector>

#include "scan_plan.h"

#pragma clang fp contract(off)

namespace ctsim {
namespace {

struct RasterArgs {
  const ElemGPU* elems;
  int nElem, nx, ny, nsample;
  double kxinc, kyinc, kxofs, kyofs, factor;
  const double* xs;  // column starts x, by repeated addition
  const double* ys;  // row starts y
  float* img;        // [nx][ny] column-major
};

// PhantomElement::isPointInside with PHM_COORD, phantom.cpp:913-975.
__device__ __forceinline__ bool point_inside(const ElemGPU& e, double x, double y) {
  const double xt = x * e.p2o[0] + y * e.p2o[2] + e.p2o[4];
  const double yt = x * e.p2o[1] + y * e.p2o[3] + e.p2o[5];
  x = xt;
  y = yt;
  switch (e.type) {
    case PE_RECT:
      return !(x > 1. || x < -1. || y > 1. || y < -1.);
    case PE_TRI:
      return !(y < 0. || y > 1. - x || y > 1. + x);
    case PE_ELLIPSE:
      if (x > 1. || x < -1. || y > 1. || y < -1.) return false;
      return !(x * x + y * y > 1.);
    case PE_SEGMENT:
      if (x > 1. || x < -1. || y > 0.) return false;
      x *= e.u;
      y *= e.v;
      return !(x * x + (y - e.v) * (y - e.v) > e.u * e.u + e.v * e.v);
    case PE_SECTOR:
      if (x > 1. || x < -1. || y > 1.) return false;
      if (y > 1. - x || y > 1. + x) return false;
      x *= e.u;
      y *= e.v;
      return !(x * x + (y - e.v) * (y - e.v) > e.u * e.u + e.v * e.v);
  }
  return false;
}

__global__ __launch_bounds__(256) void raster_kernel(RasterArgs a) {
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= static_cast<size_t>(a.nx) * a.ny) return;
  const int ix = static_cast<int>(k / a.ny), iy = static_cast<int>(k % a.ny);
  const double x = a.xs[ix], y = a.ys[iy];
  float v = 0.f;
  for (int el = a.nElem - 1; el >= 0; el--) {
    const ElemGPU& e = a.elems[el];
    double dAtten = 0;
    double xi = x + a.kxofs;
    for (int kx = 0; kx < a.nsample; kx++, xi += a.kxinc) {
      double yi = y + a.kyofs;
      for (int ky = 0; ky < a.nsample; ky++, yi += a.kyinc)
        if (point_inside(e, xi, yi)) dAtten += e.atten;
    }
    v += dAtten;
  }
  if (a.nsample > 1) v *= a.factor;
  a.img[k] = v;
}

}  // namespace
}  // namespace ctsim

using namespace ctsim;

extern "C" int ctsim_phantom_rasterize(ctsim_ctx* ctx, ctsim_phantom* phm, int nx, int ny,
                                       double viewRatio, int nsample, float* d_image,
                                       double* ext4) {
  if (!ctx || !phm || !d_image) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_rasterize: NULL argument");
  if (nx < 2 || ny < 2) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_rasterize: image smaller than 2x2");
  CTSIM_CHECK(ctx->activate());
  PhantomPlan& p = phm->plan;
  if (p.dirty || !p.d_elems.ptr) {
    CTSIM_CHECK(p.d_elems.reserve((p.elems.empty() ? 1 : p.elems.size()) * sizeof(ElemGPU)));
    if (!p.elems.empty())
      CTSIM_HIP(hipMemcpy(p.d_elems.ptr, p.elems.data(), p.elems.size() * sizeof(ElemGPU),
                          hipMemcpyHostToDevice));
    p.dirty = false;
  }
  if (nsample < 1) nsample = 1;
  // phantom.cpp:474-500
  const double dx = p.xmax - p.xmin, dy = p.ymax - p.ymin;
  const double xcent = p.xmin + dx / 2, ycent = p.ymin + dy / 2;
  const double maxAxis = dx > dy ? dx : dy;
  const double dHalflen = viewRatio * ((kSqrt2 * maxAxis) / kSqrt2 / 2);
  const double xmin = xcent - dHalflen, xmax = xcent + dHalflen;
  const double ymin = ycent - dHalflen, ymax = ycent + dHalflen;
  const double xinc = (xmax - xmin) / nx, yinc = (ymax - ymin) / ny;
  std::vector<double> tab(static_cast<size_t>(nx) + ny);
  double x = xmin;
  for (int ix = 0; ix < nx; ix++, x += xinc) tab[ix] = x;
  double y = ymin;
  for (int iy = 0; iy < ny; iy++, y += yinc) tab[nx + iy] = y;
  DevBuf d_tab;
  CTSIM_CHECK(d_tab.reserve(tab.size() * sizeof(double)));
  CTSIM_HIP(hipMemcpy(d_tab.ptr, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  RasterArgs a{};
  a.elems = p.d_elems.as<ElemGPU>();
  a.nElem = static_cast<int>(p.elems.size());
  a.nx = nx;
  a.ny = ny;
  a.nsample = nsample;
  a.kxinc = xinc / nsample;
  a.kyinc = yinc / nsample;
  a.kxofs = a.kxinc / 2;
  a.kyofs = a.kyinc / 2;
  a.factor = 1.0 / static_cast<double>(nsample * nsample);
  a.xs = d_tab.as<double>();
  a.ys = d_tab.as<double>() + nx;
  a.img = d_image;
  const size_t n = static_cast<size_t>(nx) * ny;
  hipLaunchKernelGGL(raster_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, a);
  CTSIM_HIP(hipGetLastError());
  CTSIM_HIP(hipStreamSynchronize(ctx->stream));  // the coordinate table is released at return
  if (ext4) {
    ext4[0] = xmin; ext4[1] = xmax; ext4[2] = ymin; ext4[3] = ymax;
  }
  return CTSIM_OK;
}
