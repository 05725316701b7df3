// The library's correctly rounded FP64 cos/sin/atan/atan2 (crmath.h)
// evaluated on device arrays, so the device build of the routines can be
// checked against the host build of the same source.
#include <hip/hip_runtime.h>

#include "crmath.h"
#include "ctsim_internal.h"

namespace {

__global__ __launch_bounds__(256) void crmath_eval_kernel(int fn, int n, const double* __restrict__ x,
                                                          const double* __restrict__ y,
                                                          double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = x[i];
  double r;
  if (fn == 0) r = crm_cos(v);
  else if (fn == 1) r = crm_sin(v);
  else if (fn == 2) r = crm_atan(v);
  else r = crm_atan2(y[i], v);
  out[i] = r;
}

}  // namespace

extern "C" int ctsim_crmath_eval(ctsim_ctx* ctx, int fn, int n, const double* d_x, const double* d_y,
                                 double* d_out) {
  using namespace ctsim;
  if (!ctx || !d_x || !d_out || (fn == 3 && !d_y))
    return fail(CTSIM_ERR_INVALID, "ctsim_crmath_eval: NULL argument");
  if (fn < 0 || fn > 3 || n < 0) return fail(CTSIM_ERR_INVALID, "ctsim_crmath_eval: bad function or count");
  if (n == 0) return CTSIM_OK;
  CTSIM_CHECK(ctx->activate());
  crmath_eval_kernel<<<(n + 255) / 256, 256, 0, ctx->stream>>>(fn, n, d_x, d_y, d_out);
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}
