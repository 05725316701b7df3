// Context, error state and device-buffer helpers of libctsim_amd.
#include <cstdio>
#include <cstring>

#include "ctsim_internal.h"

namespace ctsim {

static thread_local std::string t_lastError;

void set_error(const std::string& msg) { t_lastError = msg; }

int fail(int code, const std::string& msg) {
  t_lastError = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
           what, file, line);
  t_lastError = buf;
  return CTSIM_ERR_HIP;
}

int DevBuf::reserve(size_t n) {
  if (n <= bytes && ptr) return CTSIM_OK;
  // kernels still queued may read the old allocation
  if (ptr) CTSIM_HIP(hipDeviceSynchronize());
  release();
  size_t want = n < 256 ? 256 : n;
  CTSIM_HIP(hipMalloc(&ptr, want));
  bytes = want;
  return CTSIM_OK;
}

void DevBuf::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int PinnedStage::reserve(size_t n, hipStream_t s) {
  if (!done) CTSIM_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
  // The previous upload out of this area must have finished before the host
  // writes into it again.
  CTSIM_HIP(hipEventSynchronize(done));
  if (n <= bytes && ptr) return CTSIM_OK;
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  size_t want = n < 4096 ? 4096 : n;
  CTSIM_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
  bytes = want;
  (void)s;
  return CTSIM_OK;
}

int PinnedStage::upload(DevBuf& dst, size_t n, hipStream_t s) {
  CTSIM_CHECK(dst.reserve(n));
  CTSIM_HIP(hipMemcpyAsync(dst.ptr, ptr, n, hipMemcpyHostToDevice, s));
  CTSIM_HIP(hipEventRecord(done, s));
  return CTSIM_OK;
}

void PinnedStage::release() {
  if (done) {
    (void)hipEventSynchronize(done);
    (void)hipEventDestroy(done);
  }
  if (ptr) (void)hipHostFree(ptr);
  ptr = nullptr;
  done = nullptr;
  bytes = 0;
}

}  // namespace ctsim

int ctsim_ctx::activate() const {
  CTSIM_HIP(hipSetDevice(device));
  return CTSIM_OK;
}

void ctsim_ctx::prof_begin_on(hipStream_t s) {
  if (!profOn) return;
  if (profEvents.size() < 2 * (profUsed + 1)) {
    // timing-only events: without the system-scope fence a record is a
    // timestamp, not a cache writeback and invalidate between the kernels
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreateWithFlags(&a, hipEventDisableSystemFence) != hipSuccess ||
        hipEventCreateWithFlags(&b, hipEventDisableSystemFence) != hipSuccess)
      return;
    profEvents.push_back(a);
    profEvents.push_back(b);
  }
  (void)hipEventRecord(profEvents[2 * profUsed], s);
}

void ctsim_ctx::prof_end_on(hipStream_t s) {
  if (!profOn || profEvents.size() < 2 * (profUsed + 1)) return;
  (void)hipEventRecord(profEvents[2 * profUsed + 1], s);
  profUsed++;
}

ctsim_ctx::~ctsim_ctx() {
  for (hipEvent_t e : profEvents) (void)hipEventDestroy(e);
}

using namespace ctsim;

extern "C" {

const char* ctsim_last_error(void) { return t_lastError.c_str(); }

const char* ctsim_version(void) { return "ctsim_amd 0.1 (gfx950)"; }

int ctsim_ctx_create(ctsim_ctx** out, int device, void* stream) {
  if (!out) return fail(CTSIM_ERR_INVALID, "ctsim_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  CTSIM_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(CTSIM_ERR_INVALID, "ctsim_ctx_create: no such device");
  auto* c = new ctsim_ctx();
  c->device = device;
  int rc = c->activate();
  if (rc) { delete c; return rc; }
  // the device address of the no-allocation entry point's static scratch;
  // a failure here only disables that entry point, which then says so
  (void)static_scratch_setup(device);
  // NULL is HIP's null stream, so work is ordered with other null-stream work
  c->stream = static_cast<hipStream_t>(stream);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->numCUs = prop.multiProcessorCount;
  *out = c;
  return CTSIM_OK;
}

void ctsim_ctx_destroy(ctsim_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
}

int ctsim_ctx_set_stream(ctsim_ctx* ctx, void* stream) {
  if (!ctx) return fail(CTSIM_ERR_INVALID, "ctsim_ctx_set_stream: ctx is NULL");
  CTSIM_CHECK(ctx->activate());
  ctx->stream = static_cast<hipStream_t>(stream);
  return CTSIM_OK;
}

void* ctsim_ctx_stream(const ctsim_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int ctsim_ctx_profile(ctsim_ctx* ctx, int enable) {
  if (!ctx) return fail(CTSIM_ERR_INVALID, "ctsim_ctx_profile: ctx is NULL");
  ctx->profOn = enable != 0;
  ctx->profUsed = 0;
  return CTSIM_OK;
}

int ctsim_ctx_profile_read(ctsim_ctx* ctx, double* totalMs, int* count) {
  if (!ctx || !totalMs || !count) return fail(CTSIM_ERR_INVALID, "ctsim_ctx_profile_read: NULL argument");
  CTSIM_CHECK(ctx->activate());
  double tot = 0;
  for (size_t k = 0; k < ctx->profUsed; k++) {
    CTSIM_HIP(hipEventSynchronize(ctx->profEvents[2 * k + 1]));
    float ms = 0;
    CTSIM_HIP(hipEventElapsedTime(&ms, ctx->profEvents[2 * k], ctx->profEvents[2 * k + 1]));
    tot += ms;
  }
  *totalMs = tot;
  *count = static_cast<int>(ctx->profUsed);
  ctx->profUsed = 0;
  return CTSIM_OK;
}

int ctsim_ctx_synchronize(ctsim_ctx* ctx) {
  if (!ctx) return fail(CTSIM_ERR_INVALID, "ctsim_ctx_synchronize: ctx is NULL");
  CTSIM_CHECK(ctx->activate());
  CTSIM_HIP(hipStreamSynchronize(ctx->stream));
  return CTSIM_OK;
}

}  // extern "C"
