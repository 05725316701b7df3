// ctsim_fbp_slice_static: ctsim_fbp_slice at MIXED precision with every
// scratch buffer in one fixed device array, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving
// loops).
//
// `scratch` is a file-scope device array.  Host code never passes its name:
// in host code the name is a host-side stand-in, not a device address.  Its
// address on each device is taken once, with that device current, by
// hipGetSymbolAddress when a context is created on the device
// (static_scratch_setup), and that per-device address is what the kernels
// receive.  Every call on one device shares the device's copy: calls that
// overlap on a device race.  Nothing in the array needs clearing between
// calls: each call writes every byte it later reads (per-view constants,
// interval coefficients, filtered rows, and view-group partial images that
// the kernels store in full before the combine reads them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "bp_plan.h"
#include "filter_plan.h"

namespace ctsim {
namespace {

constexpr size_t kStaticScratchBytes = CTSIM_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
constexpr int kStaticMaxViews = 16384;

__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

std::mutex g_mutex;                   // guards g_scratchPtr and g_vcHost
void* g_scratchPtr[kMaxDevices] = {};  // device address of `scratch` per device
// Host staging of the per-view constants (fixed capacity; no allocation).
ViewConstF32 g_vcHost[kStaticMaxViews];

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

}  // namespace

// Takes the address of `scratch` on `device` (current) once.
int static_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(CTSIM_ERR_UNSUPPORTED, "static scratch: device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratchPtr[device]) return CTSIM_OK;
  void* ptr = nullptr;
  CTSIM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratchPtr[device] = ptr;
  return CTSIM_OK;
}

}  // namespace ctsim

using namespace ctsim;

extern "C" int ctsim_fbp_slice_static(ctsim_ctx* ctx, ctsim_procsignal* ps,
                                      const ctsim_backprojector* bp, const float* d_sino,
                                      const double* h_angles, int nView, float* d_image) {
  if (!ctx || !ps || !bp || !d_image || !d_sino || !h_angles)
    return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice_static: NULL argument");
  if (nView < 1) return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice_static: nView < 1");
  if (nView > kStaticMaxViews)
    return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice_static: more than 16384 views");
  FilterPlan& f = ps->plan;
  const BpPlan& p = bp->plan;
  if (f.n != p.nDet || f.n * f.preinterp != p.nDet * p.interpFactor)
    return fail(CTSIM_ERR_INVALID,
                "ctsim_fbp_slice_static: filter and backprojector detector counts differ");
  if (!bp_mixed_tiled(p) || p.interpFactor != 1)
    return fail(CTSIM_ERR_UNSUPPORTED,
                "ctsim_fbp_slice_static: needs the tiled mixed path (linear or cubic "
                "interpolation, no preinterpolation)");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) base = g_scratchPtr[ctx->device];
  }
  if (!base) return fail(CTSIM_ERR_INVALID, "ctsim_fbp_slice_static: static scratch not set up");
  // Layout: view constants | interval coefficients | filtered rows and
  // split projection values (convolution filters) | view-group partial
  // images.
  const bool cubic = p.interp == INTERP_CUBIC;
  const size_t elem = cubic ? sizeof(float4) : sizeof(float2);
  const bool coefFilter = filter_can_coef(f);
  const MixedShape& m = p.mixed;
  const int split = std::max(1, std::min(m.split, nView));
  const size_t npix = static_cast<size_t>(p.nx) * p.ny;
  const size_t vcBytes = align256(sizeof(ViewConstF32) * nView);
  const size_t coefBytes = align256(static_cast<size_t>(nView) * p.nDet * elem);
  const size_t rowBytes = coefFilter ? 0 : align256(static_cast<size_t>(nView) * f.n * sizeof(float));
  const size_t workBytes = coefFilter ? 0 : align256(filter_work_bytes(f, CTSIM_PREC_MIXED, nView));
  const size_t partBytes = split > 1 ? align256(static_cast<size_t>(split) * npix * sizeof(float)) : 0;
  const size_t need = vcBytes + coefBytes + rowBytes + workBytes + partBytes;
  if (need > kStaticScratchBytes)
    return fail(CTSIM_ERR_INVALID,
                "ctsim_fbp_slice_static: this slice needs " + std::to_string(need >> 20) +
                    " MiB of scratch; the static array holds " +
                    std::to_string(kStaticScratchBytes >> 20) + " MiB");
  unsigned char* b = static_cast<unsigned char*>(base);
  ViewConstF32* d_vc = reinterpret_cast<ViewConstF32*>(b);
  void* d_coef = b + vcBytes;
  float* d_rows = reinterpret_cast<float*>(b + vcBytes + coefBytes);
  void* d_work = workBytes ? b + vcBytes + coefBytes + rowBytes : nullptr;
  float* d_part = partBytes ? reinterpret_cast<float*>(b + vcBytes + coefBytes + rowBytes + workBytes) : nullptr;
  CTSIM_CHECK(ctx->activate());
  SymViews sv;
  {
    // the copy from pageable memory has read g_vcHost when it returns
    std::lock_guard<std::mutex> lock(g_mutex);
    const MirrorViews mv = bp_view_consts_f32(p, h_angles, nView, g_vcHost);
    sv = bp_sym_views(h_angles, nView);
    sv.NM = mv.NM;
    sv.mapO = mv.mapO;
    CTSIM_HIP(hipMemcpyAsync(d_vc, g_vcHost, sizeof(ViewConstF32) * nView, hipMemcpyHostToDevice,
                             ctx->stream));
  }
  if (coefFilter) {
    CTSIM_CHECK(filter_launch_coef(ctx, f, d_sino, nView, d_coef, p.interp, p.kind == KIND_IDIFF));
  } else {
    CTSIM_CHECK(filter_launch(ctx, f, CTSIM_PREC_MIXED, d_sino, nView, d_rows, f.n, d_work));
    CTSIM_CHECK(bp_mixed_coef_into(ctx, p, d_rows, nView, d_coef));
  }
  return bp_mixed_tiles(ctx, p, m, d_coef, d_vc, nView, d_image, p.rotScale, true, &sv, d_part);
}
