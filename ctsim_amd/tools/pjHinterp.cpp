// pjHinterp: helical 180LI interpolation and half-scan feathering of a .pj
// file on the GPU (tools/pjHinterp.cpp:61-149: Helical180LI(interpview),
// HalfScanFeather, write).
#include <getopt.h>

#include <cstdlib>
#include <cstring>
#include <iostream>

#include "ctsim/ctsim.h"

namespace {

enum { O_INTERPVIEW, O_TRACE, O_DEBUG, O_VERBOSE, O_HELP, O_VERSION, O_DEVICE };

const option kOptions[] = {{"interpview", 1, nullptr, O_INTERPVIEW}, {"trace", 1, nullptr, O_TRACE},
                           {"debug", 0, nullptr, O_DEBUG},           {"verbose", 0, nullptr, O_VERBOSE},
                           {"help", 0, nullptr, O_HELP},             {"version", 0, nullptr, O_VERSION},
                           {"device", 1, nullptr, O_DEVICE},         {nullptr, 0, nullptr, 0}};

void usage(const char* prog) {
  std::cout << "usage: " << prog << " proj-file interp-file [OPTIONS]\n"
            << "Helical 180LI interpolation of projection data on an MI355X GPU\n\n"
            << "  --interpview n   view to interpolate at (default: pi + fan angle)\n"
            << "  --device n  --trace none|console  --verbose --debug --version --help\n";
}

}  // namespace

int main(int argc, char* const argv[]) {
  int interpView = -1;
  while (true) {
    const int c = getopt_long(argc, argv, "", kOptions, nullptr);
    if (c == -1) break;
    switch (c) {
      case O_INTERPVIEW: {
        char* e = nullptr;
        interpView = static_cast<int>(strtol(optarg, &e, 10));
        if (e != optarg + strlen(optarg)) {
          std::cerr << "Error setting --interpview to " << optarg << std::endl;
          usage(argv[0]);
          return 1;
        }
        break;
      }
      case O_TRACE: case O_DEBUG: case O_VERBOSE: break;
      case O_DEVICE: ctsim_amd::setDevice(atoi(optarg)); break;
      case O_VERSION: std::cout << "Version " << ctsim_version() << "\n"; return 0;
      case O_HELP: case '?': usage(argv[0]); return 0;
      default: usage(argv[0]); return 1;
    }
  }
  if (optind + 2 != argc) {
    usage(argv[0]);
    return 1;
  }
  Projections projections;
  if (!projections.read(argv[optind])) {
    std::cerr << "Error reading input file " << argv[optind] << std::endl;
    return 1;
  }
  if (projections.Helical180LI(interpView) != 0) return 1;
  if (projections.HalfScanFeather() != 0) return 1;
  return projections.write(argv[optind + 1]) ? 0 : 1;
}
