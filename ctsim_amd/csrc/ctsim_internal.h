// Internal declarations shared by the host and device sources of libctsim_amd.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "ctsim_hip.h"

namespace ctsim {

// LDS a workgroup may allocate on gfx950 (160 KB per CU, all of it
// addressable by one workgroup).
constexpr size_t kLdsPerGroup = 160 * 1024;

// ---------------------------------------------------------------- constants
// include/ctsupport.h:167-177
constexpr double kPi = 3.14159265358979323846;
constexpr double kHalfPi = 1.57079632679489661923;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr double kSqrt2 = 1.414213562373095049;
constexpr double kFEps = 1.0E-6;
constexpr double kDEps = 1.0E-10;

// Geometry ids, libctsim/scanner.cpp:29-33.
enum Geometry { GEOM_PARALLEL = 0, GEOM_EQUIANGULAR = 1, GEOM_EQUILINEAR = 2, GEOM_LINOGRAM = 3 };
// Backprojector ids, libctsim/backprojectors.cpp:28-32.
enum Family { BP_TRIG = 0, BP_TABLE = 1, BP_DIFF = 2, BP_IDIFF = 3 };
// Interpolation ids, libctsim/backprojectors.cpp:53-56.
enum Interp { INTERP_NEAREST = 0, INTERP_LINEAR = 1, INTERP_CUBIC = 2 };

// ------------------------------------------------------------ error state
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define CTSIM_HIP(call)                                                   \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return ::ctsim::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

#define CTSIM_CHECK(expr)          \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != CTSIM_OK) return rc_; \
  } while (0)

// Takes the device address of the static scratch array of
// ctsim_fbp_slice_static on `device` (the current device), once per device.
int static_scratch_setup(int device);

// ------------------------------------------------------------ device buffer
// A growable device allocation owned by a context or a plan.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t n);
  void release();
  ~DevBuf() { release(); }
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host staging area whose previous upload is fenced by an event.
struct PinnedStage {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;
  int reserve(size_t n, hipStream_t s);
  int upload(DevBuf& dst, size_t n, hipStream_t s);  // copies ptr[0..n) -> dst, records done
  void release();
  ~PinnedStage() { release(); }
};

}  // namespace ctsim

// The context: one device, one stream, its scratch.
struct ctsim_ctx {
  int device = 0;
  hipStream_t stream = nullptr;  // nullptr is the null stream
  ctsim::PinnedStage stage;
  ctsim::DevBuf viewConsts;   // per-view constants of the current launch
  ctsim::DevBuf scratch;      // kernel scratch (partials, counters)
  ctsim::DevBuf scratch2;     // interval coefficients of the mixed path
  ctsim::DevBuf scratch3;     // filtered views of ctsim_fbp_slice
  // helical interpolation and half-scan feathering: the detector angles of
  // the last scan (rebuilt when geometry, nDet, detInc or SDL change) and the
  // feather's error word
  ctsim::DevBuf helixGam, helixErr;
  int helixGeom = -1, helixNDet = 0;
  double helixDetInc = 0, helixSdl = 0;
  int numCUs = 256;
  int activate() const;       // hipSetDevice(device)
  // Kernel timing: while enabled, the main backprojection kernel of each
  // launch is bracketed by a pair of events (see ctsim_ctx_profile).
  bool profOn = false;
  std::vector<hipEvent_t> profEvents;  // start, stop pairs
  size_t profUsed = 0;                 // pairs recorded since the last read
  void prof_begin() { prof_begin_on(stream); }
  void prof_end() { prof_end_on(stream); }
  void prof_begin_on(hipStream_t s);
  void prof_end_on(hipStream_t s);
  ~ctsim_ctx();
};
