// phm2pj: simulate a scan of a phantom into a .pj projection file on the GPU.
// Command line and defaults follow the reference's tools/phm2pj.cpp:34-51,
// 103-121, 272-277 (parallel geometry, nray 1, focal and centre-detector
// ratios 2, rotation 0.5 turn for parallel and 1 turn otherwise).
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "ctsim/ctsim.h"

namespace {

enum { O_PHANTOM, O_PHMFILE, O_DESC, O_NRAY, O_ROTANGLE, O_GEOMETRY, O_FOCAL_LENGTH,
       O_CENTER_DETECTOR_LENGTH, O_OFFSETVIEW, O_VIEW_RATIO, O_SCAN_RATIO, O_TRACE, O_VERBOSE,
       O_HELP, O_DEBUG, O_VERSION, O_PRECISION, O_DEVICE };

const option kOptions[] = {
    {"phantom", 1, nullptr, O_PHANTOM},
    {"phmfile", 1, nullptr, O_PHMFILE},
    {"desc", 1, nullptr, O_DESC},
    {"nray", 1, nullptr, O_NRAY},
    {"rotangle", 1, nullptr, O_ROTANGLE},
    {"geometry", 1, nullptr, O_GEOMETRY},
    {"focal-length", 1, nullptr, O_FOCAL_LENGTH},
    {"center-detector-length", 1, nullptr, O_CENTER_DETECTOR_LENGTH},
    {"offsetview", 1, nullptr, O_OFFSETVIEW},
    {"view-ratio", 1, nullptr, O_VIEW_RATIO},
    {"scan-ratio", 1, nullptr, O_SCAN_RATIO},
    {"trace", 1, nullptr, O_TRACE},
    {"verbose", 0, nullptr, O_VERBOSE},
    {"help", 0, nullptr, O_HELP},
    {"debug", 0, nullptr, O_DEBUG},
    {"version", 0, nullptr, O_VERSION},
    {"precision", 1, nullptr, O_PRECISION},
    {"device", 1, nullptr, O_DEVICE},
    {nullptr, 0, nullptr, 0}};

void usage(const char* prog) {
  std::cout << "usage: " << prog << " outfile ndet nview [OPTIONS]\n"
            << "Calculate (projections) through phantom object on an MI355X GPU\n\n"
            << "  --phantom shepp-logan|herman|unit-pulse   builtin phantom\n"
            << "  --phmfile file        phantom definition file\n"
            << "  --desc text           description stored in the file\n"
            << "  --nray n              samples per detector (default 1)\n"
            << "  --rotangle a          rotation in turns (default 0.5 parallel, 1 fan beam)\n"
            << "  --geometry parallel|equiangular|equilinear\n"
            << "  --focal-length r      source to centre, in view diameters (default 2)\n"
            << "  --center-detector-length r  centre to detector (default 2)\n"
            << "  --view-ratio r  --scan-ratio r  --offsetview n\n"
            << "  --precision f64|mixed (default f64)  --device n\n"
            << "  --trace none|console  --verbose --debug --version --help\n";
}

bool to_int(const char* s, int& v) {
  char* e = nullptr;
  v = static_cast<int>(strtol(s, &e, 10));
  return *s && e == s + strlen(s);
}

bool to_double(const char* s, double& v) {
  char* e = nullptr;
  v = strtod(s, &e);
  return *s && e == s + strlen(s);
}

}  // namespace

int main(int argc, char* const argv[]) {
  const auto t0 = std::chrono::steady_clock::now();
  std::string geometry = "parallel", desc, phmName, phmFile;
  int nray = 1, offsetView = 0, trace = Trace::TRACE_NONE;
  double focal = 2., centerDet = 2., viewRatio = 1., scanRatio = 1., rotAngle = -1;
  bool verbose = false;
  while (true) {
    const int c = getopt_long(argc, argv, "", kOptions, nullptr);
    if (c == -1) break;
    bool ok = true;
    switch (c) {
      case O_PHANTOM: phmName = optarg; break;
      case O_PHMFILE: phmFile = optarg; break;
      case O_DESC: desc = optarg; break;
      case O_NRAY: ok = to_int(optarg, nray); break;
      case O_ROTANGLE: ok = to_double(optarg, rotAngle); break;
      case O_GEOMETRY: geometry = optarg; break;
      case O_FOCAL_LENGTH: ok = to_double(optarg, focal); break;
      case O_CENTER_DETECTOR_LENGTH: ok = to_double(optarg, centerDet); break;
      case O_OFFSETVIEW: ok = to_int(optarg, offsetView); break;
      case O_VIEW_RATIO: ok = to_double(optarg, viewRatio); break;
      case O_SCAN_RATIO: ok = to_double(optarg, scanRatio); break;
      case O_VERBOSE: verbose = true; break;
      case O_DEBUG: break;
      case O_TRACE: ok = (trace = Trace::convertTraceNameToID(optarg)) != Trace::TRACE_INVALID; break;
      case O_PRECISION:
        if (!strcmp(optarg, "mixed")) ctsim_amd::setPrecision(CTSIM_PREC_MIXED);
        else if (!strcmp(optarg, "f64")) ctsim_amd::setPrecision(CTSIM_PREC_F64);
        else ok = false;
        break;
      case O_DEVICE: {
        int d = 0;
        ok = to_int(optarg, d);
        ctsim_amd::setDevice(d);
        break;
      }
      case O_VERSION: std::cout << "Version " << ctsim_version() << "\n"; return 0;
      case O_HELP: case '?': usage(argv[0]); return 0;
      default: ok = false;
    }
    if (!ok) {
      usage(argv[0]);
      return 1;
    }
  }
  if (optind + 3 != argc || (phmName.empty() && phmFile.empty())) {
    usage(argv[0]);
    return 1;
  }
  const char* outFile = argv[optind];
  int nDet = 0, nView = 0;
  if (!to_int(argv[optind + 1], nDet) || !to_int(argv[optind + 2], nView)) {
    usage(argv[0]);
    return 1;
  }
  if (rotAngle < 0) rotAngle = geometry == "parallel" ? 0.5 : 1.0;

  std::ostringstream label;
  label << "phm2pj: NDet=" << nDet << ", Nview=" << nView << ", NRay=" << nray
        << ", RotAngle=" << rotAngle << "OffsetView =" << offsetView << ", Geometry=" << geometry
        << ", ";
  if (!phmFile.empty()) label << "PhantomFile=" << phmFile;
  else label << "Phantom=" << phmName;
  if (!desc.empty()) label << ": " << desc;

  Phantom phm;
  if (!phmName.empty() && !phm.createFromPhantom(phmName.c_str())) {
    std::cout << phm.failMessage() << "\n\n";
    usage(argv[0]);
    return 1;
  }
  // The reference ignores createFromFile's result (phm2pj.cpp:305); a
  // file that cannot be opened leaves the phantom empty and the Scanner fails.
  if (!phmFile.empty()) phm.createFromFile(phmFile.c_str());
  // The reference multiplies the turn count by 2*pi in double (phm2pj.cpp:332).
  const double twoPi = 6.28318530717958647692;
  Scanner scanner(phm, geometry.c_str(), nDet, nView, offsetView, nray, rotAngle * twoPi, focal,
                  centerDet, viewRatio, scanRatio);
  if (scanner.fail()) {
    std::cout << "Scanner Creation Error: " << scanner.failMessage() << "\n";
    return 1;
  }
  Projections pj(scanner);
  scanner.collectProjections(pj, phm, 0, nView, offsetView, true, trace);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  pj.setCalcTime(secs);
  pj.setRemark(label.str());
  if (!pj.write(outFile)) {
    std::cerr << "Unable to write " << outFile << "\n";
    return 1;
  }
  if (verbose) {  // phm2pj.cpp:381-389
    phm.print(std::cout);
    std::cout << std::endl;
    std::ostringstream os;
    pj.printScanInfo(os);
    std::cout << os.str() << std::endl;
    std::cout << "  Remark: " << pj.remark() << std::endl;
    std::cout << "Run time: " << pj.calcTime() << " seconds\n";
  }
  return 0;
}
