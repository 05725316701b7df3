/* ctsim_hip.h -- C ABI of libctsim_amd.so, the MI355X (gfx950) implementation
 * of CTSim's projection / filtering / backprojection hot path.
 *
 * Conventions
 *   - Every entry point returns an int status: CTSIM_OK (0) or a negative
 *     CTSIM_ERR_* code; ctsim_last_error() returns the message of the last
 *     failure on the calling thread.  Nothing throws across this boundary.
 *   - "d_" pointers are device memory (hipMalloc'd or equivalent) and are
 *     read/written asynchronously on the context's stream; "h_" pointers are
 *     host memory read during the call.
 *   - Sinograms are row-major [nView][nDet] (one DetectorArray per row,
 *     include/scanner.h:41-62).  Images are column-major [nx][ny]: pixel
 *     (ix, iy) is element ix*ny + iy, the layout of the reference's
 *     v[ix][iy] (libctsim/array2dfile.cpp:290-298).
 *   - Names (geometry, filter, method, backprojector, interpolation) are the
 *     reference's own strings, matched case-insensitively, including its
 *     quirks ("shepp" selects the Abs-Sinc filter, libctsim/filter.cpp:35-59;
 *     the table method is spelled "fouier-table", libctsim/procsignal.cpp:45).
 *
 * Precision
 *   CTSIM_PREC_F64    the drop-in path: FP64 arithmetic with the reference's
 *                     evaluation order and float32 storage, view-ordered
 *                     float32 image updates.  Filtered views are double.
 *   CTSIM_PREC_MIXED  FP32 backprojection with Kahan-compensated view
 *                     accumulation, FP32 LDS FFT / split-BF16 MFMA Toeplitz
 *                     filtering.  Filtered views are float.  FP64 is kept
 *                     for host geometry and trig tables.
 */
#ifndef CTSIM_HIP_H
#define CTSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CTSIM_OK 0
#define CTSIM_ERR_INVALID (-1)     /* bad argument or name (reference: fail()) */
#define CTSIM_ERR_HIP (-2)         /* HIP runtime error */
#define CTSIM_ERR_UNSUPPORTED (-3) /* combination the reference cannot run either */

#define CTSIM_PREC_F64 0
#define CTSIM_PREC_MIXED 1

typedef struct ctsim_ctx ctsim_ctx;
typedef struct ctsim_phantom ctsim_phantom;
typedef struct ctsim_scanner ctsim_scanner;
typedef struct ctsim_procsignal ctsim_procsignal;
typedef struct ctsim_backprojector ctsim_backprojector;

/* ------------------------------------------------------------- context */

/* Bind a context to a device and a stream (NULL: the context creates its own
 * default-priority stream).  No counterpart in the reference, which is
 * single-threaded CPU code; it carries the device state the classes below
 * need. */
int ctsim_ctx_create(ctsim_ctx** out, int device, void* stream);
void ctsim_ctx_destroy(ctsim_ctx* ctx);
int ctsim_ctx_set_stream(ctsim_ctx* ctx, void* stream);
void* ctsim_ctx_stream(const ctsim_ctx* ctx);
int ctsim_ctx_synchronize(ctsim_ctx* ctx);
/* Message of the last failure on this thread (reference: fail()/failMessage(),
 * e.g. libctsim/scanner.cpp:93-106). */
const char* ctsim_last_error(void);
/* Library build identification. */
const char* ctsim_version(void);

/* ------------------------------------------------------- backprojection */

/* Replaces Backprojector::Backprojector -> initBackprojector and the
 * Backproject / BackprojectTable / BackprojectDiff constructors
 * (libctsim/backprojectors.cpp:100-180, 264-319, 413-429, 491-502).
 * Geometry values come from the Projections (include/projections.h:146-160):
 * detInc, nView (for rotScale), viewDiameter (phmLen = viewDiameter/sqrt2),
 * focalLength and sourceDetectorLength.  roi is NULL or the four values of a
 * ReconstructionROI in its member order {xMin, yMin, xMax, yMax}
 * (include/reconstruct.h:36-41).  Equilinear and equiangular geometries
 * select their fan-beam backprojectors whatever backprojName says
 * (backprojectors.cpp:160-163). */
int ctsim_backprojector_create(ctsim_ctx* ctx, ctsim_backprojector** out, int geometry,
                               int nDet, double detInc, int nView, double viewDiameter,
                               double focalLength, double sourceDetectorLength,
                               const char* backprojName, const char* interpName,
                               int interpFactor, int nx, int ny, const double* roi);
void ctsim_backprojector_destroy(ctsim_backprojector* bp);

/* Replaces the view loop of Reconstructor::reconstructView over
 * Backprojector::BackprojectView (libctsim/reconstruct.cpp:164-187,
 * backprojectors.cpp:110-114 and the family bodies 365-743): adds views
 * [0, nView) of d_filtered (row stride nDet*interpFactor; double for F64,
 * float for MIXED) at angles h_angles into d_image, in view order. */
int ctsim_backproject_views(ctsim_ctx* ctx, const ctsim_backprojector* bp, int precision,
                            const void* d_filtered, const double* h_angles, int nView,
                            float* d_image);

/* Replaces Backprojector::PostProcessing (backprojectors.cpp:116-121,
 * 324-336, 435-442, 508-515): scales by rotScale except for Trig, which
 * scaled per view. */
int ctsim_backproject_postprocess(ctsim_ctx* ctx, const ctsim_backprojector* bp,
                                  float* d_image);

/* Geometry the backprojector derived: xMin, xMax, yMin, yMax, xInc, yInc,
 * rotScale, iDetCenter (8 doubles). */
int ctsim_backprojector_info(const ctsim_backprojector* bp, double* out8);

#ifdef __cplusplus
}
#endif
#endif /* CTSIM_HIP_H */
