/* CPU oracle for the CTSim projection / filtering / backprojection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms, used as the checker by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py.  The product (ctsim_amd/) never links,
 * loads or calls it.
 *
 * The four driver entry points take exactly the arguments of the reference
 * harness (oracle/ref_harness.cpp) so the two can be compared call for call:
 *   orc_pipeline    phantom -> Scanner::collectProjections -> per view
 *                   ProcessSignal::filterSignal + Backprojector::BackprojectView
 *                   -> PostProcessing
 *   orc_filter      ProcessSignal ctor + filterSignal on caller data
 *   orc_backproject Backprojector on caller-supplied filtered views
 *   orc_stats       ImageFile::comparativeStatistics
 * Images are column-major: pixel (ix, iy) lives at [ix * ny + iy], which is
 * the reference's v[ix][iy] (array2dfile.cpp:290-298).
 *
 * Parity status (pinned by tests/test_oracle_*.py):
 *   every function here is checked bit-for-bit against the reference built
 *   from source (oracle/_ref) and against the committed golden vectors in
 *   tests/golden/, and independently against closed forms (analytic ellipse
 *   and rectangle Radon transforms, filter frequency responses, DFT
 *   identities, Simpson exactness on cubics, Lagrange exactness on cubics).
 */
#ifndef CTSIM_ORACLE_H
#define CTSIM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int orc_pipeline(const int* types, const double* params, int nElem, const char* builtin,
                 const char* geometry, int nDet, int nView, int offsetView, int nSample,
                 double rotLen, double focalRatio, double centerRatio, double viewRatio,
                 double scanRatio, const char* filterName, double filterParam,
                 const char* methodName, int zeropad, const char* generationName,
                 const char* interpName, int interpFactor, const char* backprojName, int nx,
                 int ny, float* sino, double* angles, double* filtered, float* image,
                 float* imageRecon, double* info, double* timings);

int orc_filter(const char* filterName, const char* methodName, double bw, double inc, int n,
               double param, const char* domainName, const char* generationName, int zeropad,
               int preinterp, int geometry, double focalLength, double sdl, const float* in,
               int nView, double* out, double* filterOut, int filterCap, int* nFilterPoints,
               double* filterMinMaxInc);

int orc_backproject(const int* types, const double* params, int nElem, const char* builtin,
                    const char* geometry, int nDet, int nView, int offsetView, int nSample,
                    double rotLen, double focalRatio, double centerRatio, double viewRatio,
                    double scanRatio, const char* backprojName, const char* interpName,
                    int interpFactor, int nx, int ny, const double* roi, const double* filtered,
                    const double* angles, int postProcess, float* image);

int orc_stats(int nx, int ny, const float* refImage, const float* cmpImage, double* dre);

int orc_fbp_sample(const int* types, const double* params, int nElem, const char* builtin,
                   const char* geometry, int nDet, int nView, int nSample, double rotLen,
                   double focalRatio, double centerRatio, const char* filterName,
                   const char* methodName, int zeropad, const char* interpName,
                   const char* backprojName, int nx, int ny, const float* sino,
                   const double* angles, int nUse, float* image, double* timings);

const char* orc_last_error(void);

/* Building blocks exposed for closed-form checks. */
double orc_frequency_response(int filterID, double bw, double u, double param);
double orc_spatial_response_analytic(int filterID, double bw, double x, double param);
double orc_spatial_response_calc(int filterID, double bw, double x, double param, int n);
double orc_integrate_simpson(double xmin, double xmax, const double* y, int np);
double orc_cubic_interpolate(const double* y, int n, double x);
int orc_add_zeropad_factor(int n, int zeropad);
/* Ray sum of one segment (x1,y1)-(x2,y2) through a phantom. */
double orc_raysum_line(const int* types, const double* params, int nElem, const char* builtin,
                       double x1, double y1, double x2, double y2);
/* Phantom extent after construction: xmin, xmax, ymin, ymax. */
int orc_phantom_extent(const int* types, const double* params, int nElem, const char* builtin,
                       double* ext);

/* Helical 180LI interpolation and half-scan feathering
 * (Projections::Helical180LI_Equiangular / HalfScanFeather,
 * libctsim/projections.cpp:231-439); geom 1 equiangular, 2 equilinear. */
int orc_helical180li(int geom, int nDet, int nView, double rotInc, double detInc,
                     double fanAngle, double sdl, int interpView, const float* in, float* out,
                     double* anglesOut);
int orc_halfscan_feather(int geom, int nDet, int nView, double rotInc, double detInc,
                         double fanAngle, double sdl, float* data);

/* Fan-beam to parallel rebinning (Projections::interpolateToParallel,
 * libctsim/projections.cpp:1250-1328). */
int orc_interpolate_to_parallel(int geom, int nDet, int nView, double detStart, double detInc,
                                double focalLength, double sdl, double viewDiameter,
                                double rotStart, const double* angles, const float* in,
                                float* out, double* anglesOut, double* info6);

/* Phantom::convertToImagefile (libctsim/phantom.cpp:463-538). */
int orc_rasterize(const int* types, const double* params, int nElem, const char* builtin,
                  int nx, int ny, double viewRatio, int nsample, float* img, double* ext4);

#ifdef __cplusplus
}
#endif
#endif
