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
typedef struct ctsim_batch ctsim_batch;
typedef struct ctsim_rebin ctsim_rebin;

/* ------------------------------------------------------------- context */

/* Bind a context to a device and a stream (a hipStream_t; NULL is the null
 * stream).  Every call on the context is ordered on that stream.  No
 * counterpart in the reference, which is single-threaded CPU code; it
 * carries the device state the classes below need. */
int ctsim_ctx_create(ctsim_ctx** out, int device, void* stream);
void ctsim_ctx_destroy(ctsim_ctx* ctx);
int ctsim_ctx_set_stream(ctsim_ctx* ctx, void* stream);
void* ctsim_ctx_stream(const ctsim_ctx* ctx);
int ctsim_ctx_synchronize(ctsim_ctx* ctx);
/* No counterpart in the reference (measurement).  While enabled, the main
 * backprojection kernel of every launch on this context is bracketed by HIP
 * events on the context's stream (timing-only events, without the
 * system-scope fence; each record still adds a few microseconds between
 * kernels, so time throughput with profiling off); ctsim_ctx_profile_read
 * waits for them and returns the summed kernel time and the number of
 * launches since the last read (or since enabling). */
int ctsim_ctx_profile(ctsim_ctx* ctx, int enable);
int ctsim_ctx_profile_read(ctsim_ctx* ctx, double* totalMs, int* count);
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

/* One whole slice: d_image = the reconstruction of nView rows of d_sino
 * (the reference's sequence for one slice: a zeroed ImageFile, filterSignal
 * of every view, BackprojectView of every view, PostProcessing;
 * reconstruct.cpp:156-239).  No single counterpart in the reference.  The
 * mixed precision fuses the steps (FFT filters write the backprojector's
 * interval coefficients straight from LDS; the backprojection overwrites the
 * image with the post-scale applied); F64, nearest interpolation and Diff run
 * the separate steps.  The filter's point count must equal the
 * backprojector's detector count. */
int ctsim_fbp_slice(ctsim_ctx* ctx, ctsim_procsignal* ps, const ctsim_backprojector* bp,
                    int precision, const float* d_sino, const double* h_angles, int nView,
                    float* d_image);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ctsim_fbp_slice at CTSIM_PREC_MIXED with every scratch
 * buffer (per-view constants, interval coefficients, filtered rows and
 * BF16-split projection values of convolution filters, view-group partial
 * images) in one static device array
 * of CTSIM_STATIC_SCRATCH_BYTES per device, whose address is taken when a
 * context is created on the device.  Calls no hipMalloc/hipFree or other
 * allocating routine; the caller passes the inputs and the output.  Same
 * result as ctsim_fbp_slice(..., CTSIM_PREC_MIXED, ...), bit for bit.
 * Needs the tiled mixed path (linear or cubic interpolation, no
 * preinterpolation) and at most 16384 views; a slice whose scratch exceeds
 * the array is rejected with CTSIM_ERR_INVALID before any launch.  Every
 * call on one device shares that device's array: calls that overlap on a
 * device race. */
#define CTSIM_STATIC_SCRATCH_BYTES (64u << 20)
int ctsim_fbp_slice_static(ctsim_ctx* ctx, ctsim_procsignal* ps, const ctsim_backprojector* bp,
                           const float* d_sino, const double* h_angles, int nView,
                           float* d_image);

/* Geometry the backprojector derived: xMin, xMax, yMin, yMax, xInc, yInc,
 * rotScale, iDetCenter (8 doubles). */
int ctsim_backprojector_info(const ctsim_backprojector* bp, double* out8);
/* No counterpart in the reference (introspection of the MIXED mirror kernel).
 * The mirror members these view angles allow on this backprojector's grid:
 * 4 (every view has partners under x -> -x, y -> -y and both), 2 (under one
 * of x -> -x and the point mirror) or 1 (none: the plain kernel runs).  Host
 * computation only. */
int ctsim_backprojector_mirror_members(const ctsim_backprojector* bp, const double* h_angles,
                                       int nView);
/* No counterpart in the reference (launch tuning of the MIXED kernel).
 * Times candidate shapes (lanes along y, rows per lane, view groups, views
 * per LDS chunk, plain or rotationally symmetric kernel and the latter's LDS
 * staging size) on nView float rows of d_filtered into a scratch image and
 * keeps the fastest for later MIXED launches of this plan; shape7 (optional)
 * receives lanesY, P, split, chunk, window, sym, staging KB and bestMs
 * (optional) its time. */
int ctsim_backprojector_autotune(ctsim_ctx* ctx, ctsim_backprojector* bp, const float* d_filtered,
                                 const double* h_angles, int nView, int* shape7, double* bestMs);
/* Sets the MIXED launch shape explicitly: lanesY (8, 16, 32, 64), P (1, 2,
 * 4), split (1..64), chunk (1..64), sym (0; 1: the symmetric kernel where the
 * image is square and centred, the views come in 2 or 4 rotated copies, and
 * P <= 2; 2: its software-pipelined form, windows of at most 128 cells;
 * 3: the mirror kernel where the grid is centred and every view has partner
 * views under the grid's mirrors, 2 or 4 updates per position, else the
 * plain kernel), staging KB of the chunked symmetric kernel (0..62, 0: 62). */
int ctsim_backprojector_set_shape(ctsim_backprojector* bp, const int* shape6);

/* ------------------------------------------------------------ filtering */

/* Replaces ProcessSignal::ProcessSignal -> init (libctsim/procsignal.cpp:82-442)
 * with the same arguments (include/procsignal.h:66-69): the filter is
 * generated on the host in FP64 exactly as the reference generates it
 * (SignalFilter, libctsim/filter.cpp:110-584).  methodName may be
 * "convolution", "fourier", "fouier-table", or "fft"/"rfftw" (the reference's
 * FFTW real-to-halfcomplex path, procsignal.cpp:412-440, 600-613) or "fftw"
 * (its complex FFTW path, 614-626). */
int ctsim_procsignal_create(ctsim_ctx* ctx, ctsim_procsignal** out, const char* filterName,
                            const char* methodName, double bw, double signalIncrement, int n,
                            double param, const char* domainName, const char* generationName,
                            int zeropad, int preinterpolationFactor, int geometry,
                            double focalLength, double sourceDetectorLength);
void ctsim_procsignal_destroy(ctsim_procsignal* ps);
/* getNFilterPoints / getFilterMin / getFilterMax / getFilterIncrement
 * (include/procsignal.h:80-84). */
int ctsim_procsignal_info(const ctsim_procsignal* ps, int* nFilterPoints, double* minMaxInc3);
/* getFilter (include/procsignal.h:85-86): copies up to capacity values,
 * returns the number of filter points. */
int ctsim_procsignal_get_filter(const ctsim_procsignal* ps, double* h_out, int capacity);
/* Replaces ProcessSignal::filterSignal (procsignal.cpp:544-629) over nView
 * rows of d_in (float [nView][n]) into d_out ([nView][n]; double for F64,
 * float for MIXED). */
int ctsim_filter_views(ctsim_ctx* ctx, ctsim_procsignal* ps, int precision, const float* d_in,
                       int nView, void* d_out);

/* ------------------------------------------------------- phantom, scanner */

/* Replaces Phantom::Phantom / createFromPhantom (libctsim/phantom.cpp:58-170):
 * builtinName is NULL/"" (empty phantom) or "herman", "shepp-logan",
 * "unit-pulse". */
int ctsim_phantom_create(ctsim_phantom** out, const char* builtinName);
void ctsim_phantom_destroy(ctsim_phantom* phm);
/* Replaces Phantom::addPElem (phantom.cpp:236-256): type is "rectangle",
 * "triangle", "ellipse", "sector" or "segment", rotDeg in degrees, as in the
 * phantom file format (phantom.cpp:184-213). */
int ctsim_phantom_add_element(ctsim_phantom* phm, const char* type, double cx, double cy,
                              double u, double v, double rotDeg, double atten);
/* xmin(), xmax(), ymin(), ymax() (include/phantom.h:196-199) into ext4;
 * returns the number of elements. */
int ctsim_phantom_extent(const ctsim_phantom* phm, double* ext4);
/* Element `index` of Phantom::listPElem() (the reference's newest-first list
 * order, phantom.cpp:247), as PhantomElement reports it (include/phantom.h:
 * 41-128): type (1 rectangle, 2 triangle, 3 ellipse, 4 sector, 5 segment),
 * cx, cy, u, v, rot (radians), atten, xmin, ymin, xmax, ymax (outline
 * bounding box), nOutlinePoints, and the rotation in degrees as it was
 * defined (13 doubles).  Phantom::print and printDefinitions
 * (phantom.cpp:272-322) read it. */
int ctsim_phantom_element(const ctsim_phantom* phm, int index, double* out13);

/* Replaces Scanner::Scanner (libctsim/scanner.cpp:86-224; include/scanner.h:75-77),
 * rotLen in radians.  nSample < 1 is rejected (the reference divides by
 * zero). */
int ctsim_scanner_create(ctsim_scanner** out, const ctsim_phantom* phm, const char* geometry,
                         int nDet, int nView, int offsetView, int nSample, double rotLen,
                         double focalLengthRatio, double centerDetectorRatio, double viewRatio,
                         double scanRatio);
void ctsim_scanner_destroy(ctsim_scanner* sc);
/* Scanner accessors (include/scanner.h:92-108): detInc, detStart, rotInc,
 * viewDiameter, focalLength, sourceDetectorLength, fanBeamAngle, phmLen,
 * detLen, centerDetectorLength, scanDiameter, geometry (12 doubles). */
int ctsim_scanner_info(const ctsim_scanner* sc, double* out12);
/* Replaces Scanner::collectProjections(proj, phm, iStartView, iNumViews,
 * iOffsetView, iStorageOffset, trace) (scanner.cpp:296-419): writes rows
 * [iStorageOffset, iStorageOffset + iNumViews) of d_sino (float [*][nDet])
 * and, if h_angles is not NULL, the same entries of h_angles (the
 * DetectorArray view angles). */
int ctsim_collect_projections(ctsim_ctx* ctx, const ctsim_scanner* sc, ctsim_phantom* phm,
                              int precision, int iStartView, int iNumViews, int iOffsetView,
                              int iStorageOffset, float* d_sino, double* h_angles);

/* ------------------------------------------------------ batched slices */

/* One slice of a batch: the Projections header fields a Reconstructor reads
 * (reconstruct.cpp:50-103), the reconstruction options (pjrec's, with
 * interpolation factor 1) and the view angles (host, nView doubles; read at
 * ctsim_batch_create). */
typedef struct ctsim_slice_desc {
  int geometry; /* 0 parallel, 1 equiangular, 2 equilinear */
  int nDet;
  int nView;
  double detInc;
  double viewDiameter;
  double focalLength;
  double sourceDetectorLength;
  const char* filter;
  const char* method;
  double filterParam;
  const char* generation;
  int zeropad;
  const char* backproj;
  const char* interp; /* "linear" or "cubic" */
  int nx;
  int ny;
  const double* angles;
} ctsim_slice_desc;

/* No single counterpart in the reference (it reconstructs one slice per
 * pjrec run, tools/pjrec.cpp:342-347): a batch plan for the MIXED path of
 * nSlices slices of any geometry and size.  Filtering and backprojection of
 * every slice run in one persistent kernel (or a filter launch and a
 * backprojection launch) captured in a HIP graph.  Convolution filters run as
 * FFT convolutions in the batch. */
int ctsim_batch_create(ctsim_ctx* ctx, ctsim_batch** out, const ctsim_slice_desc* slices,
                       int nSlices);
void ctsim_batch_destroy(ctsim_batch* b);
/* Reconstructs every slice: d_sinos[k] float [nView][nDet] (device) into
 * d_images[k] float [nx][ny] (device, overwritten, post-scaled as after
 * PostProcessing).  Host arrays of device pointers; asynchronous on the
 * context's stream. */
int ctsim_batch_run(ctsim_ctx* ctx, ctsim_batch* b, const float* const* d_sinos,
                    float* const* d_images);
/* Waits for the batch and reports a backprojection item that gave up
 * waiting for its slice's filtering (never expected). */
int ctsim_batch_status(ctsim_ctx* ctx, ctsim_batch* b);
/* Launch shape: P rows per lane (1, 2, 4), lanesY (8, 16, 32), LDS bytes per
 * view chunk, fused (1) or two-phase (0), sym (0; 1, 2 or 4: slices with a
 * square centred image whose views come in 2 or 4 rotated copies run the
 * rotationally symmetric body, when P <= 2 or in two-phase mode, where they
 * take P = 2 beside a P = 4 launch of the others, with each quadrant tile's
 * views in that many groups whose partial images are added after the
 * backprojection; + 8: in two-phase mode the other slices whose views have
 * mirror partners run the mirror body). */
int ctsim_batch_set_shape(ctsim_ctx* ctx, ctsim_batch* b, int P, int lanesY, int chunkBytes,
                          int fused, int sym);
/* Times candidate shapes on these inputs and keeps the fastest; shape5 gets
 * P, lanesY, chunkBytes, fused, sym. */
int ctsim_batch_autotune(ctsim_ctx* ctx, ctsim_batch* b, const float* const* d_sinos,
                         float* const* d_images, int* shape5, double* bestMs);
/* out4: slices, pixel-view updates per run, fused, persistent grid size. */
int ctsim_batch_info(const ctsim_batch* b, double* out4);

/* ---------------------------------------------------------------- helical */

/* Replaces Projections::Helical180LI (projections.cpp:200-358) on the device:
 * d_in float [nView][nDet] of a fan-beam scan (rotInc, detInc, fan angle as
 * in the Projections header) -> d_out float [L+1][nDet], L = int((pi + fan) /
 * rotInc), *nViewOut = L+1, h_anglesOut (optional, L+1 doubles) the new view
 * angles.  interpView < 0 picks the reference's default view.  geometry 1
 * (equiangular) follows the reference bit for bit; geometry 2 (equilinear,
 * which the reference rejects at projections.cpp:214-219) uses the same
 * rules with gamma = atan(t / sourceDetectorLength).  Samples for which the
 * reference would index past its arrays are dropped. */
int ctsim_helical180li(ctsim_ctx* ctx, int geometry, int nDet, int nView, double rotInc,
                       double detInc, double fanBeamAngle, double sourceDetectorLength,
                       int interpView, const float* d_in, float* d_out, int outCapacity,
                       int* nViewOut, double* h_anglesOut);
/* Replaces Projections::HalfScanFeather (projections.cpp:369-439), reading
 * d_in and writing d_out (separate buffers), both float [nView][nDet]. */
int ctsim_halfscan_feather(ctsim_ctx* ctx, int geometry, int nDet, int nView, double rotInc,
                           double detInc, double fanBeamAngle, double sourceDetectorLength,
                           const float* d_in, float* d_out);

/* -------------------------------------------------------------- rebinning */

/* Replaces Projections::interpolateToParallel (projections.cpp:1250-1328,
 * ParallelRaysums 1339-1469) for a fan-beam sinogram (geometry 1 or 2) with
 * the Projections header fields given and view angles h_angles (host): the
 * plan sorts the (theta, t) coordinates of every sample once; runs rebin
 * float [nView][nDet] rows on the device.  Exact (t, theta) ties, which the
 * reference's unstable sort leaves unordered, keep view-major order. */
int ctsim_rebin_create(ctsim_ctx* ctx, ctsim_rebin** out, int geometry, int nDet, int nView,
                       double detStart, double detInc, double focalLength,
                       double sourceDetectorLength, double viewDiameter, double rotStart,
                       const double* h_angles);
void ctsim_rebin_destroy(ctsim_rebin* r);
/* out4: rotStart (0), rotInc, detStart, detInc of the parallel projections;
 * h_newAngles (optional, nView doubles): their view angles. */
int ctsim_rebin_info(const ctsim_rebin* r, double* out4, double* h_newAngles);
int ctsim_rebin_run(ctsim_ctx* ctx, ctsim_rebin* r, const float* d_in, float* d_out);

/* ------------------------------------------------------------ rasteriser */

/* Replaces Phantom::convertToImagefile (phantom.cpp:463-538, the phm2if
 * tool): d_image float [nx][ny] column-major (nx, ny >= 2), nsample x
 * nsample samples per pixel; ext4 (optional) receives the axis extent
 * xmin, xmax, ymin, ymax.  Synchronises the context's stream. */
int ctsim_phantom_rasterize(ctsim_ctx* ctx, ctsim_phantom* phm, int nx, int ny, double viewRatio,
                            int nsample, float* d_image, double* ext4);

/* ------------------------------------------------------------- statistics */

/* Replaces ImageFile::comparativeStatistics (imagefile.cpp:185-240) on
 * device images: d_ref is the reference, both float [nx][ny] column-major.
 * Writes d, r, e to h_dre[3] and synchronises the context's stream. */
int ctsim_image_stats(ctsim_ctx* ctx, const float* d_ref, const float* d_cmp, int nx, int ny,
                      double* h_dre);

/* -------------------------------------------------- FP64 elementary functions */

/* No counterpart in the reference, which calls the C library: evaluates the
 * correctly rounded FP64 routines the FP64 kernels use in place of the C
 * library's (fn 0 cos, 1 sin, 2 atan, 3 atan2(d_y[i], d_x[i])) on n device
 * doubles into d_out, for checking the device build of the routines against
 * their host build and a higher-precision reference. */
int ctsim_crmath_eval(ctsim_ctx* ctx, int fn, int n, const double* d_x, const double* d_y,
                      double* d_out);

#ifdef __cplusplus
}
#endif
#endif /* CTSIM_HIP_H */
