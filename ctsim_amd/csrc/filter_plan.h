// ProcessSignal plan: the filter the reference generates for one
// (filter, method, generation, geometry, detector) combination, and the
// device tables the filtering kernels use.
#pragma once

#include "ctsim_internal.h"

namespace ctsim {

// Filter method ids, libctsim/procsignal.cpp:34-41 ("fftw"/"rfftw" exist when
// the reference is built with FFTW).
enum FilterMethod {
  METHOD_CONVOLUTION = 0,
  METHOD_FOURIER = 1,
  METHOD_FOURIER_TABLE = 2,
  METHOD_FFT = 3,
  METHOD_FFTW = 4,
  METHOD_RFFTW = 5,
};

struct FilterPlan {
  int filterId = 0, method = 0, generation = 0, geom = 0;
  int n = 0, nfp = 0, nOut = 0, preinterp = 1, zeropad = 0;
  double inc = 0, bw = 0, F = 0, sdl = 0, fmin = 0, fmax = 0, finc = 0, param = 0;
  std::vector<double> filter;     // m_adFilter, nfp points
  std::vector<double> preweight;  // fan-beam input weights (empty for parallel)
  // Spectral multipliers for k = 0..N/2 applied to the real and imaginary
  // parts of the unnormalised forward transform (1/N folded in).
  std::vector<double> hre, him;
  // Exact-DFT twiddle tables (fourier / fouier-table), index i*j.
  std::vector<double> tfc, tfs, tbc, tbs;
  bool tablesBuilt = false;

  DevBuf d_filter, d_weight, d_weight32, d_hmul64, d_hmul32, d_tw64, d_tw32, d_dft;
  DevBuf d_toeplitz;  // bf16 hi, mid, lo of g[q] = dx * filter[2n-2-q], mixed convolution
};

// Host side (filter_host.cpp).
int filter_plan_init(FilterPlan& p, const char* filterName, const char* methodName, double bw,
                     double signalInc, int n, double param, const char* domainName,
                     const char* generationName, int zeropad, int preinterp, int geometry,
                     double focalLength, double sdl);
void filter_build_dft_tables(FilterPlan& p);
bool filter_is_fourier(const FilterPlan& p);

// Device side (filter.hip).
int filter_upload(FilterPlan& p);
// d_work: device scratch of filter_work_bytes() for the mixed convolution
// (its split projection values), or nullptr to take the context's scratch.
int filter_launch(ctsim_ctx* ctx, FilterPlan& p, int precision, const float* d_in, int nView,
                  void* d_out, int outStride, void* d_work = nullptr);
size_t filter_work_bytes(const FilterPlan& p, int precision, int nView);
// Mixed FFT filtering straight into backprojection interval coefficients
// (float4 cubic / float2 linear, [nView][n]); CTSIM_ERR_UNSUPPORTED for
// convolution filters or non-power-of-two transforms.
int filter_launch_coef(ctsim_ctx* ctx, FilterPlan& p, const float* d_in, int nView, void* d_coef,
                       int interp, bool lastIncl);
// True when filter_launch_coef supports the plan: an FFT method with a
// power-of-two transform of at most 4096 points and no pre-interpolation.
bool filter_can_coef(const FilterPlan& p);

}  // namespace ctsim

struct ctsim_procsignal {
  ctsim::FilterPlan plan;
};
