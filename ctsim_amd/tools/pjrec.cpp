// pjrec: reconstruct an image from a .pj projection file on the GPU.
// Command line and defaults follow the reference's tools/pjrec.cpp:31-47,
// 127-137 (filter abs_bandlimit, convolution, direct generation, zeropad 1,
// linear interpolation, idiff backprojection).  --precision mixed selects
// the fast path; the default f64 path reproduces the reference's image.
#include <getopt.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "ctsim/ctsim.h"

namespace {

enum { O_INTERP, O_PREINTERP, O_FILTER, O_FILTER_METHOD, O_ZEROPAD, O_FILTER_GENERATION,
       O_FILTER_PARAM, O_BACKPROJ, O_TRACE, O_DEBUG, O_VERBOSE, O_HELP, O_VERSION, O_PRECISION,
       O_DEVICE };

const option kOptions[] = {
    {"interp", 1, nullptr, O_INTERP},
    {"preinterpolation-factor", 1, nullptr, O_PREINTERP},
    {"filter", 1, nullptr, O_FILTER},
    {"filter-method", 1, nullptr, O_FILTER_METHOD},
    {"zeropad", 1, nullptr, O_ZEROPAD},
    {"filter-generation", 1, nullptr, O_FILTER_GENERATION},
    {"filter-param", 1, nullptr, O_FILTER_PARAM},
    {"backproj", 1, nullptr, O_BACKPROJ},
    {"trace", 1, nullptr, O_TRACE},
    {"debug", 0, nullptr, O_DEBUG},
    {"verbose", 0, nullptr, O_VERBOSE},
    {"help", 0, nullptr, O_HELP},
    {"version", 0, nullptr, O_VERSION},
    {"precision", 1, nullptr, O_PRECISION},
    {"device", 1, nullptr, O_DEVICE},
    {nullptr, 0, nullptr, 0}};

void usage(const char* prog) {
  std::cout << "usage: " << prog << " raysum-file image-file nx-image ny-image [OPTIONS]\n"
            << "Image reconstruction from raysum projections on an MI355X GPU\n\n"
            << "  --interp nearest|linear|cubic          (default linear)\n"
            << "  --preinterpolation-factor n            (default 1)\n"
            << "  --filter abs_bandlimit|abs_sinc|abs_cosine|abs_hamming|abs_hanning|shepp|\n"
            << "           bandlimit|sinc|cosine|triangle|hamming|hanning (default abs_bandlimit)\n"
            << "  --filter-method convolution|fourier|fourier_table|fft|fftw|rfftw\n"
            << "  --zeropad n                            (default 1)\n"
            << "  --filter-generation direct|inverse_fourier\n"
            << "  --filter-param alpha                   Hamming alpha\n"
            << "  --backproj trig|table|diff|idiff       (default idiff)\n"
            << "  --trace none|console\n"
            << "  --precision f64|mixed                  (default f64)\n"
            << "  --device n                             GPU ordinal (default 0)\n"
            << "  --verbose --debug --version --help\n";
}

bool to_int(const char* s, int& v) {
  char* e = nullptr;
  v = static_cast<int>(strtol(s, &e, 10));
  return e == s + strlen(s);
}

bool to_double(const char* s, double& v) {
  char* e = nullptr;
  v = strtod(s, &e);
  return e == s + strlen(s);
}

}  // namespace

int main(int argc, char* const argv[]) {
  const auto t0 = std::chrono::steady_clock::now();
  std::string filter = "abs_bandlimit", method = "convolution", generation = "direct";
  std::string interp = "linear", backproj = "idiff";
  int zeropad = 1, preinterp = 1, trace = Trace::TRACE_NONE;
  double filterParam = -1;
  bool verbose = false;
  while (true) {
    const int c = getopt_long(argc, argv, "", kOptions, nullptr);
    if (c == -1) break;
    switch (c) {
      case O_INTERP: interp = optarg; break;
      case O_PREINTERP: if (!to_int(optarg, preinterp)) { usage(argv[0]); return 1; } break;
      case O_FILTER: filter = optarg; break;
      case O_FILTER_METHOD: method = optarg; break;
      case O_FILTER_GENERATION: generation = optarg; break;
      case O_FILTER_PARAM: if (!to_double(optarg, filterParam)) { usage(argv[0]); return 1; } break;
      case O_ZEROPAD: if (!to_int(optarg, zeropad)) { usage(argv[0]); return 1; } break;
      case O_BACKPROJ: backproj = optarg; break;
      case O_VERBOSE: verbose = true; break;
      case O_DEBUG: break;
      case O_TRACE:
        if ((trace = Trace::convertTraceNameToID(optarg)) == Trace::TRACE_INVALID) {
          usage(argv[0]);
          return 1;
        }
        break;
      case O_PRECISION:
        if (!strcmp(optarg, "mixed")) ctsim_amd::setPrecision(CTSIM_PREC_MIXED);
        else if (!strcmp(optarg, "f64")) ctsim_amd::setPrecision(CTSIM_PREC_F64);
        else { usage(argv[0]); return 1; }
        break;
      case O_DEVICE: {
        int d;
        if (!to_int(optarg, d)) { usage(argv[0]); return 1; }
        ctsim_amd::setDevice(d);
        break;
      }
      case O_VERSION: std::cout << "Version " << ctsim_version() << "\n"; return 0;
      case O_HELP: case '?': usage(argv[0]); return 0;
      default: usage(argv[0]); return 1;
    }
  }
  if (optind + 4 != argc) {
    usage(argv[0]);
    return 1;
  }
  const char* pjName = argv[optind];
  const char* imName = argv[optind + 1];
  const int nx = atoi(argv[optind + 2]), ny = atoi(argv[optind + 3]);

  std::ostringstream label;
  label << "pjrec: " << nx << "x" << ny << ", " << filter;
  if (filterParam >= 0) label << ": alpha=" << filterParam;
  label << ", " << interp << ", preinterpolationFactor=" << preinterp << ", " << backproj;
  if (verbose) std::cout << "SRemark: " << label.str() << "\n";

  Projections proj;
  if (!proj.read(pjName)) {
    fprintf(stderr, "Unable to read projectfile file %s\n", pjName);
    return 1;
  }
  if (verbose) {
    std::ostringstream os;
    proj.printScanInfo(os);
    std::cout << os.str();
  }
  ImageFile im(nx, ny);
  {
    Reconstructor rec(proj, im, filter.c_str(), filterParam, method.c_str(), zeropad,
                      generation.c_str(), interp.c_str(), preinterp, backproj.c_str(), trace);
    if (rec.fail()) {
      std::cout << rec.failMessage();
      return 1;
    }
    rec.reconstructAllViews();
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  im.labelAdd(proj.getLabel());  // pjrec.cpp:355-356
  im.labelAdd(Array2dFileLabel::L_HISTORY, label.str().c_str(), secs);
  if (!im.fileWrite(imName)) {
    fprintf(stderr, "Unable to write image file %s\n", imName);
    return 1;
  }
  if (verbose) std::cout << "Run time: " << secs << " seconds\n";
  return 0;
}
