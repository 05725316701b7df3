// phm2if: rasterise a phantom into a .if image on the GPU.  Command line
// after the reference's tools/phm2if.cpp (outfile nx ny, --phantom or
// --phmfile, --nsample, --view-ratio).
#include <getopt.h>

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "ctsim/ctsim.h"

namespace {

enum { O_PHANTOM, O_PHMFILE, O_DESC, O_NSAMPLE, O_VIEW_RATIO, O_TRACE, O_VERBOSE, O_HELP,
       O_DEBUG, O_VERSION, O_DEVICE };

const option kOptions[] = {
    {"phantom", 1, nullptr, O_PHANTOM},     {"phmfile", 1, nullptr, O_PHMFILE},
    {"desc", 1, nullptr, O_DESC},           {"nsample", 1, nullptr, O_NSAMPLE},
    {"view-ratio", 1, nullptr, O_VIEW_RATIO}, {"trace", 1, nullptr, O_TRACE},
    {"verbose", 0, nullptr, O_VERBOSE},     {"help", 0, nullptr, O_HELP},
    {"debug", 0, nullptr, O_DEBUG},         {"version", 0, nullptr, O_VERSION},
    {"device", 1, nullptr, O_DEVICE},       {nullptr, 0, nullptr, 0}};

void usage(const char* prog) {
  std::cout << "usage: " << prog << " outfile nx ny [OPTIONS]\n"
            << "Convert phantom object to an image file on an MI355X GPU\n\n"
            << "  --phantom shepp-logan|herman|unit-pulse  builtin phantom\n"
            << "  --phmfile file     phantom definition file\n"
            << "  --nsample n        samples per pixel along each axis (default 1)\n"
            << "  --view-ratio r     view diameter ratio (default 1)\n"
            << "  --desc text        description label\n"
            << "  --device n  --trace none|console  --verbose --debug --version --help\n";
}

}  // namespace

int main(int argc, char* const argv[]) {
  std::string phmName, phmFile, desc;
  int nsample = 1;
  double viewRatio = 1.0;
  bool verbose = false;
  while (true) {
    const int c = getopt_long(argc, argv, "", kOptions, nullptr);
    if (c == -1) break;
    switch (c) {
      case O_PHANTOM: phmName = optarg; break;
      case O_PHMFILE: phmFile = optarg; break;
      case O_DESC: desc = optarg; break;
      case O_NSAMPLE: nsample = atoi(optarg); break;
      case O_VIEW_RATIO: viewRatio = atof(optarg); break;
      case O_TRACE: break;
      case O_VERBOSE: verbose = true; break;
      case O_DEBUG: break;
      case O_DEVICE: ctsim_amd::setDevice(atoi(optarg)); break;
      case O_VERSION: std::cout << "Version " << ctsim_version() << "\n"; return 0;
      case O_HELP: case '?': usage(argv[0]); return 0;
      default: usage(argv[0]); return 1;
    }
  }
  if (optind + 3 != argc || (phmName.empty() && phmFile.empty())) {
    usage(argv[0]);
    return 1;
  }
  const char* outFile = argv[optind];
  const int nx = atoi(argv[optind + 1]), ny = atoi(argv[optind + 2]);
  Phantom phm;
  if (!phmName.empty() && !phm.createFromPhantom(phmName.c_str())) {
    std::cout << phm.failMessage() << "\n";
    return 1;
  }
  if (!phmFile.empty() && !phm.createFromFile(phmFile.c_str())) {
    std::cerr << "Unable to read phantom file " << phmFile << "\n";
    return 1;
  }
  ImageFile im(nx, ny);
  phm.convertToImagefile(im, viewRatio, nsample);
  std::ostringstream label;
  label << "phm2if: nx=" << nx << ", ny=" << ny << ", nsample=" << nsample << ", ";
  if (!phmFile.empty()) label << "PhantomFile=" << phmFile;
  else label << "Phantom=" << phmName;
  if (!desc.empty()) label << ": " << desc;
  im.labelAdd(label.str().c_str());
  if (!im.fileWrite(outFile)) {
    std::cerr << "Unable to write " << outFile << "\n";
    return 1;
  }
  if (verbose) std::cout << label.str() << "\n";
  return 0;
}
