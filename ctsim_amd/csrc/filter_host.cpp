// Host side of the ProcessSignal plan: filter generation exactly as the
// reference's SignalFilter / ProcessSignal constructors do it.  Compiled with
// g++ -O2 -ffp-contract=off so the kernels come out bit-identical.
#include <strings.h>

#include <cmath>
#include <complex>

#include "filter_plan.h"

namespace ctsim {

namespace {

// Name tables, libctsim/filter.cpp:46-59 (index = id; "shepp" sits at the
// Abs-Sinc id 4 and "abs_sinc" at the Shepp id 5) and procsignal.cpp:42-49,
// 66-69; filter.cpp:80-83.
const char* const kFilterNames[] = {"abs_bandlimit", "abs_hamming", "abs_hanning", "abs_cosine",
                                    "shepp",         "abs_sinc",    "bandlimit",   "sinc",
                                    "hamming",       "hanning",     "cosine",      "triangle"};
const char* const kMethodNames[] = {"convolution", "fourier", "fouier-table", "fft", "fftw", "rfftw"};
const char* const kGenerationNames[] = {"direct", "inverse-fourier"};
const char* const kDomainNames[] = {"frequency", "spatial"};

enum {
  F_ABS_BANDLIMIT = 0, F_ABS_G_HAMMING = 1, F_ABS_HANNING = 2, F_ABS_COSINE = 3,
  F_ABS_SINC = 4, F_SHEPP = 5, F_BANDLIMIT = 6, F_SINC = 7, F_G_HAMMING = 8,
  F_HANNING = 9, F_COSINE = 10, F_TRIANGLE = 11
};

int lookup(const char* name, const char* const* table, int n) {
  if (!name) return -1;
  for (int i = 0; i < n; i++)
    if (strcasecmp(name, table[i]) == 0) return i;
  return -1;
}

// SignalFilter::sinc (filter.h:120-124), integral_abscos (filter.h:163-164)
double sinc(double x) { return fabs(x) > kFEps ? sin(x) / x : 1.0; }
double sinc(double x, double mult) { return fabs(x) > kFEps ? sin(x * mult) / x : 1.0; }
double integral_abscos(double u, double w) {
  return fabs(u) > kFEps ? (cos(u * w) - 1) / (u * u) + w / u * sin(u * w) : (w * w / 2);
}

// SignalFilter::frequencyResponse, filter.cpp:403-475
double frequency_response(int id, double bw, double u, double param) {
  const double au = fabs(u), abw = fabs(bw);
  const bool out = au >= (abw / 2) + kFEps;
  switch (id) {
    case F_BANDLIMIT: return out ? 0. : 1;
    case F_ABS_BANDLIMIT: return out ? 0. : au;
    case F_TRIANGLE: return out ? 0 : 1 - au / abw;
    case F_COSINE: return out ? 0 : cos(kPi * au / abw);
    case F_ABS_COSINE: return out ? 0 : au * cos(kPi * au / abw);
    case F_SINC: return abw * sinc(kPi * abw * au, 1.);
    case F_ABS_SINC: return out ? 0 : au * abw * sinc(kPi * abw * au, 1.);
    case F_HANNING: param = 0.5; [[fallthrough]];
    case F_G_HAMMING: return out ? 0 : param + (1 - param) * cos(kTwoPi * au / abw);
    case F_ABS_HANNING: param = 0.5; [[fallthrough]];
    case F_ABS_G_HAMMING: return out ? 0 : au * (param + (1 - param) * cos(kTwoPi * au / abw));
    default: return 0;  // reference: warning, 0 (filter.cpp:468-471)
  }
}

bool have_analytic(int id) {  // filter.cpp:489-512
  switch (id) {
    case F_BANDLIMIT: case F_TRIANGLE: case F_COSINE: case F_G_HAMMING: case F_HANNING:
    case F_ABS_BANDLIMIT: case F_ABS_COSINE: case F_ABS_G_HAMMING: case F_ABS_HANNING:
    case F_SHEPP: case F_SINC:
      return true;
    default:
      return false;
  }
}

// SignalFilter::spatialResponseAnalytic, filter.cpp:525-584
double spatial_analytic(int id, double bw, double x, double param) {
  const double u = kTwoPi * x, w = bw / 2, b = kPi / bw, b2 = kTwoPi / bw;
  double t;
  switch (id) {
    case F_BANDLIMIT: return bw * sinc(u * w, 1.0);
    case F_TRIANGLE: t = sinc(u * w, 1.0); return bw * t * t;
    case F_COSINE: return sinc(b - u, w) + sinc(b + u, w);
    case F_HANNING: param = 0.5; [[fallthrough]];
    case F_G_HAMMING:
      return 2 * param * sin(u * w) / u + (1 - param) * (sinc(b2 - u, w) + sinc(b2 + u, w));
    case F_ABS_BANDLIMIT: return 2 * integral_abscos(u, w);
    case F_ABS_COSINE: return integral_abscos(b - u, w) + integral_abscos(b + u, w);
    case F_ABS_HANNING: param = 0.5; [[fallthrough]];
    case F_ABS_G_HAMMING:
      return 2 * param * integral_abscos(u, w) +
             (1 - param) * (integral_abscos(u - b2, w) + integral_abscos(u + b2, w));
    case F_SHEPP:
      if (fabs(u) < 1E-7) return 4. / (kPi * bw * bw);
      return fabs((2 / bw) * sin(u * w)) * sinc(u * w, 1.) * sinc(u * w, 1.);
    case F_SINC: return fabs(x) < bw / 2 ? 1. : 0.;
    default: return 0;
  }
}

// integrateSimpson, libctsupport/mathfuncs.cpp:36-60
double integrate_simpson(double xmin, double xmax, const double* y, int np) {
  if (np < 2) return 0.;
  if (np == 2) return (xmax - xmin) * (y[0] + y[1]) / 2;
  double area = 0;
  const int nDiv = (np - 1) / 2;
  const double width = (xmax - xmin) / static_cast<double>(np - 1);
  for (int i = 1; i <= nDiv; i++) {
    const int xr = 2 * i, xl = xr - 2, xm = xr - 1;
    area += (width / 3.0) * (y[xl] + 4.0 * y[xm] + y[xr]);
  }
  if ((np & 1) == 0) area += width * (y[np - 2] + y[np - 1]) / 2;
  return area;
}

// SignalFilter::spatialResponseCalc, filter.cpp:358-381 (N_INTEGRAL = 500)
double spatial_calc(int id, double bw, double x, double param, int n) {
  const double zmin = 0, zmax = id == F_TRIANGLE ? bw : bw / 2;
  const double zinc = (zmax - zmin) / (n - 1);
  std::vector<double> q(n);
  double z = zmin;
  for (int i = 0; i < n; i++, z += zinc) q[i] = frequency_response(id, bw, z, param) * cos(kTwoPi * z * x);
  return 2 * integrate_simpson(zmin, zmax, q.data(), n);
}

// SignalFilter::init + createFrequencyFilter/createSpatialFilter,
// filter.cpp:158-229
std::vector<double> signal_filter(int id, double fmin, double fmax, int n, double bw, double param,
                                  bool spatial) {
  std::vector<double> out(n);
  const double inc = (fmax - fmin) / (n - 1);
  if (!spatial) {
    double x = fmin;
    for (int i = 0; i < n; x += inc, i++) out[i] = frequency_response(id, bw, x, param);
    return out;
  }
  if (id == F_SHEPP) {
    const double a = 2 * bw, c = -4. / (a * a);
    const int center = (n - 1) / 2;
    out[center] = 4. / (a * a);
    for (int i = 1; i <= center; i++) out[center + i] = out[center - i] = c / (4 * (i * i) - 1);
    return out;
  }
  double x = fmin;
  for (int i = 0; i < n; i++, x += inc)
    out[i] = have_analytic(id) ? spatial_analytic(id, bw, x, param) : spatial_calc(id, bw, x, param, 500);
  return out;
}

// Fourier::shuffleNaturalToFourierOrder / shuffleFourierToNaturalOrder,
// include/fourier.h:49-101
void shuffle_natural_to_fourier(std::vector<double>& v) {
  const int n = static_cast<int>(v.size());
  std::vector<double> t(n);
  if (n % 2) {
    const int h = (n - 1) / 2;
    t[0] = v[h];
    for (int i = 0; i < h; i++) t[i + 1] = v[i + 1 + h];
    for (int i = 0; i < h; i++) t[i + h + 1] = v[i];
  } else {
    const int h = n / 2;
    t[0] = v[h];
    for (int i = 0; i < h - 1; i++) t[i + 1] = v[i + h + 1];
    for (int i = 0; i < h; i++) t[i + h] = v[i];
  }
  v.swap(t);
}
void shuffle_fourier_to_natural(std::vector<double>& v) {
  const int n = static_cast<int>(v.size());
  std::vector<double> t(n);
  if (n % 2) {
    const int h = (n - 1) / 2;
    t[h] = v[0];
    for (int i = 0; i < h; i++) t[i + 1 + h] = v[i + 1];
    for (int i = 0; i < h; i++) t[i] = v[i + h + 1];
  } else {
    const int h = n / 2;
    t[h] = v[0];
    for (int i = 0; i < h; i++) t[i] = v[i + h];
    for (int i = 0; i < h - 1; i++) t[i + h + 1] = v[i + 1];
  }
  v.swap(t);
}

// ProcessSignal::finiteFourierTransform (double in, complex out),
// procsignal.cpp:706-728; direction < 0 is FORWARD (normalised by 1/n).
void dft(const std::vector<double>& in, std::vector<double>& re, std::vector<double>& im, int dir) {
  const int n = static_cast<int>(in.size());
  dir = dir < 0 ? -1 : 1;
  const double ainc = dir * 2 * kPi / n;
  re.assign(n, 0);
  im.assign(n, 0);
  for (int i = 0; i < n; i++) {
    double sr = 0, si = 0;
    for (int j = 0; j < n; j++) {
      const double a = i * j * ainc;
      sr += in[j] * cos(a);
      si += in[j] * sin(a);
    }
    if (dir < 0) {
      sr /= n;
      si /= n;
    }
    re[i] = sr;
    im[i] = si;
  }
}

}  // namespace

// ProcessSignal ctor + init, procsignal.cpp:82-442
int filter_plan_init(FilterPlan& p, const char* filterName, const char* methodName, double bw,
                     double signalInc, int n, double param, const char* domainName,
                     const char* generationName, int zeropad, int preinterp, int geometry,
                     double focalLength, double sdl) {
  p.method = lookup(methodName, kMethodNames, 6);
  if (p.method < 0)
    return fail(CTSIM_ERR_INVALID, std::string("Invalid filter method name ") + (methodName ? methodName : "(null)"));
  p.generation = lookup(generationName, kGenerationNames, 2);
  if (p.generation < 0)
    return fail(CTSIM_ERR_INVALID, std::string("Invalid frequency filter name ") + (generationName ? generationName : "(null)"));
  p.filterId = lookup(filterName, kFilterNames, 12);
  if (p.filterId < 0)
    return fail(CTSIM_ERR_INVALID, std::string("Invalid Filter name ") + (filterName ? filterName : "(null)"));
  if (lookup(domainName, kDomainNames, 2) < 0)
    return fail(CTSIM_ERR_INVALID, std::string("Invalid domain name ") + (domainName ? domainName : "(null)"));
  if (n < 2) return fail(CTSIM_ERR_INVALID, "ProcessSignal: at least 2 signal points required");
  if (preinterp != 1)
    return fail(CTSIM_ERR_UNSUPPORTED,
                "ProcessSignal: preinterpolation factor other than 1 is not supported");
  p.geom = geometry;
  p.F = focalLength;
  p.sdl = sdl;
  p.bw = bw;
  p.n = n;
  p.inc = signalInc;
  p.param = param;
  p.zeropad = zeropad;
  p.preinterp = preinterp;
  if (geometry == GEOM_EQUILINEAR) {  // procsignal.cpp:152-158
    const double sc = sdl / focalLength;
    p.inc /= sc;
    p.bw *= sc;
  }
  // The reference remaps "fft" to its FFTW real transform when built with
  // FFTW (procsignal.cpp:160-168); this build provides that transform itself.
  if (p.method == METHOD_FFT) p.method = METHOD_RFFTW;
  const bool freq = p.method != METHOD_CONVOLUTION;
  if (!freq) {
    p.nfp = 2 * (n - 1) + 1;
    if (p.generation == 0) {
      p.fmin = -p.inc * (n - 1);
      p.fmax = p.inc * (n - 1);
      p.finc = (p.fmax - p.fmin) / (p.nfp - 1);
      p.filter = signal_filter(p.filterId, p.fmin, p.fmax, p.nfp, p.bw, param, true);
    } else {
      p.fmin = -1. / (2 * p.inc);
      p.fmax = 1. / (2 * p.inc);
      p.finc = (p.fmax - p.fmin) / (p.nfp - 1);
      std::vector<double> fr = signal_filter(p.filterId, p.fmin, p.fmax, p.nfp, p.bw, param, false);
      shuffle_natural_to_fourier(fr);
      std::vector<double> re, im;
      dft(fr, re, im, -1);
      p.filter = re;
      shuffle_fourier_to_natural(p.filter);
      for (int i = 0; i < p.nfp; i++) p.filter[i] /= p.inc;
    }
    if (geometry == GEOM_EQUILINEAR) {
      for (int i = 0; i < p.nfp; i++) p.filter[i] *= 0.5;
    } else if (geometry == GEOM_EQUIANGULAR) {
      for (int i = 0; i < p.nfp; i++) {
        const int k = i - ((p.nfp - 1) / 2);
        const double s = 1 / sinc(k * p.inc);
        const double sc = 0.5 * s * s;
        p.filter[i] *= sc;
      }
    }
  } else if (p.generation == 0) {
    // addZeropadFactor, procsignal.cpp:867-881
    p.nfp = n;
    if (zeropad > 0) {
      const double l = log(n) / log(2);
      p.nfp = 1 << (static_cast<int>(floor(l)) + zeropad);
    }
    p.nOut = p.nfp * preinterp;
    p.fmin = -1. / (2 * p.inc);
    p.fmax = 1. / (2 * p.inc);
    if (p.nfp % 2) {
      p.finc = (p.fmax - p.fmin) / (p.nfp - 1);
    } else {
      p.finc = (p.fmax - p.fmin) / p.nfp;
      p.fmax -= p.finc;
    }
    p.filter = signal_filter(p.filterId, p.fmin, p.fmax, p.nfp, p.bw, param, false);
    if (geometry == GEOM_EQUILINEAR) {
      for (int i = 0; i < p.nfp; i++) p.filter[i] *= 0.5;
    } else if (geometry == GEOM_EQUIANGULAR) {
      for (int i = 0; i < p.nfp; i++) {
        const int k = i - ((p.nfp - 1) / 2);
        const double s = 1 / sinc(k * p.inc);
        const double sc = 0.5 * s * s;
        p.filter[i] *= sc;
      }
    }
    shuffle_natural_to_fourier(p.filter);
  } else {
    const int nsp = 2 * (n - 1) + 1;
    p.fmin = -p.inc * (n - 1);
    p.fmax = p.inc * (n - 1);
    p.finc = (p.fmax - p.fmin) / (nsp - 1);
    p.nfp = nsp;
    if (zeropad > 0) {
      const double l = log(nsp) / log(2);
      int np2 = static_cast<int>(floor(l));
      if (l != floor(l)) np2++;
      np2 += zeropad - 1;
      p.nfp = 1 << np2;
    }
    p.nOut = p.nfp * preinterp;
    std::vector<double> sp = signal_filter(p.filterId, p.fmin, p.fmax, nsp, p.bw, param, true);
    if (geometry == GEOM_EQUILINEAR) {
      for (int i = 0; i < nsp; i++) sp[i] *= 0.5;
    } else if (geometry == GEOM_EQUIANGULAR) {
      for (int i = 0; i < nsp; i++) {
        const int k = i - ((nsp - 1) / 2);
        double s = sin(k * p.inc);
        if (fabs(s) < 1E-7) s = 1;
        else s = (k * p.inc) / s;
        const double sc = 0.5 * s * s;
        sp[i] *= sc;
      }
    }
    sp.resize(std::max(p.nfp, nsp), 0.0);
    for (int i = nsp; i < p.nfp; i++) sp[i] = 0;
    sp.resize(p.nfp);
    std::vector<double> re, im;
    dft(sp, re, im, 1);
    p.filter.resize(p.nfp);
    for (int i = 0; i < p.nfp; i++) p.filter[i] = std::abs(std::complex<double>(re[i], im[i])) * p.inc;
  }

  // Fan-beam pre-weighting of the input, procsignal.cpp:551-561.
  p.preweight.clear();
  if (geometry == GEOM_EQUILINEAR) {
    p.preweight.resize(n);
    for (int i = 0; i < n; i++) {
      const int k = i - (n / 2);
      p.preweight[i] = p.F / sqrt(p.F * p.F + k * k * p.inc * p.inc);
    }
  } else if (geometry == GEOM_EQUIANGULAR) {
    p.preweight.resize(n);
    for (int i = 0; i < n; i++) {
      const int k = i - (n / 2);
      p.preweight[i] = p.F * cos(k * p.inc);
    }
  }

  // Spectral multipliers of the Fourier methods (k = 0..N/2, 1/N folded in).
  if (freq) {
    const int N = p.nfp;
    p.hre.assign(N / 2 + 1, 0.0);
    p.him.assign(N / 2 + 1, 0.0);
    for (int k = 0; k <= N / 2; k++) {
      const double hk = p.filter[k], hn = p.filter[(N - k) % N];
      if (p.method == METHOD_RFFTW) {
        // halfcomplex product, procsignal.cpp:604-605: the real part of
        // frequency k meets filter[k], its imaginary part filter[N-k]
        p.hre[k] = hk / N;
        p.him[k] = hn / N;
      } else {
        // complex product then real part of the inverse: the filter acts
        // through its even part
        p.hre[k] = p.him[k] = 0.5 * (hk + hn) / N;
      }
    }
  }
  return CTSIM_OK;
}

bool filter_is_fourier(const FilterPlan& p) {
  return p.method == METHOD_FOURIER || p.method == METHOD_FOURIER_TABLE;
}

// Twiddle tables of the exact DFT.  "fourier" evaluates cos/sin(i*j*inc) with
// inc = -/+ 2 pi / N (procsignal.cpp:706-774); "fouier-table" reads tables
// built by repeated angle addition (procsignal.cpp:399-410).  Both are
// indexed by i*j.
void filter_build_dft_tables(FilterPlan& p) {
  if (p.tablesBuilt) return;
  const int N = p.nfp;
  const size_t nk = static_cast<size_t>(N - 1) * (N - 1) + 1;
  p.tfc.resize(nk);
  p.tfs.resize(nk);
  p.tbc.resize(nk);
  p.tbs.resize(nk);
  if (p.method == METHOD_FOURIER) {
    const double incF = -1 * 2 * kPi / N;
    const double incB = 1 * 2 * kPi / N;
    for (size_t k = 0; k < nk; k++) {
      const int ki = static_cast<int>(k);
      const double aF = ki * incF, aB = ki * incB;
      p.tfc[k] = cos(aF);
      p.tfs[k] = sin(aF);
      p.tbc[k] = cos(aB);
      p.tbs[k] = sin(aB);
    }
  } else {
    // nFourier = max(nfp, nOut)^2 + 1 entries; only the first (N-1)^2+1 are read.
    const double ainc = (2. * kPi) / N;
    double a = 0;
    for (size_t k = 0; k < nk; k++) {
      const double c = cos(a), s = sin(a);
      // forward: sumImag -= x * sinTable (procsignal.cpp:788-791)
      p.tfc[k] = c;
      p.tfs[k] = -s;
      p.tbc[k] = c;
      p.tbs[k] = s;
      a += ainc;
    }
  }
  p.tablesBuilt = true;
}

}  // namespace ctsim
