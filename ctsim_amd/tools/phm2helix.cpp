// phm2helix: projections through a time-varying phantom on the GPU.
// Command line, defaults and per-view protocol follow the reference's
// tools/phm2helix.cpp:28-334: before view 0 and before every view v the
// external phantom program is run as `phmprog v nview phmfile`, the phantom
// it writes is read (Phantom::createFromFile) and view v alone is collected
// into the projections (collectProjections(pj, phm, v, 1, offsetView, true)).
// The Scanner's geometry comes from the view-0 phantom.  --precision mixed
// selects the closed-form ray sums; the default f64 path reproduces the
// reference's ray sums.
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "ctsim/ctsim.h"

namespace {

enum { O_PHANTOMPROG, O_PHMFILE, O_DESC, O_NRAY, O_ROTANGLE, O_GEOMETRY, O_FOCAL_LENGTH,
       O_CENTER_DETECTOR_LENGTH, O_VIEW_RATIO, O_SCAN_RATIO, O_OFFSETVIEW, O_TRACE, O_VERBOSE,
       O_HELP, O_DEBUG, O_VERSION, O_PRECISION, O_DEVICE };

const option kOptions[] = {
    {"phantom", 1, nullptr, O_PHANTOMPROG},
    {"desc", 1, nullptr, O_DESC},
    {"nray", 1, nullptr, O_NRAY},
    {"rotangle", 1, nullptr, O_ROTANGLE},
    {"geometry", 1, nullptr, O_GEOMETRY},
    {"focal-length", 1, nullptr, O_FOCAL_LENGTH},
    {"center-detector-length", 1, nullptr, O_CENTER_DETECTOR_LENGTH},
    {"view-ratio", 1, nullptr, O_VIEW_RATIO},
    {"scan-ratio", 1, nullptr, O_SCAN_RATIO},
    {"offsetview", 1, nullptr, O_OFFSETVIEW},
    {"trace", 1, nullptr, O_TRACE},
    {"verbose", 0, nullptr, O_VERBOSE},
    {"help", 0, nullptr, O_HELP},
    {"debug", 0, nullptr, O_DEBUG},
    {"version", 0, nullptr, O_VERSION},
    {"phmfile", 1, nullptr, O_PHMFILE},
    {"precision", 1, nullptr, O_PRECISION},
    {"device", 1, nullptr, O_DEVICE},
    {nullptr, 0, nullptr, 0}};

void usage(const char* prog) {
  std::cout << "usage: " << fileBasename(prog) << " outfile ndet nview phmprog [OPTIONS]\n"
            << "Calculate (projections) through time varying phantom object on an MI355X GPU\n\n"
            << "     outfile          Name of output file for projections\n"
            << "     ndet             Number of detectors\n"
            << "     nview            Number of rotated views\n"
            << "     phmprog          phantom program, run as: phmprog iview nview phmfile\n"
            << "     --phmfile        Temp phantom file name (default tmpphmfile)\n"
            << "     --desc  --nray  --rotangle (turns)  --geometry parallel|equilinear|equiangular\n"
            << "     --focal-length  --center-detector-length  --view-ratio  --scan-ratio\n"
            << "     --offsetview  --trace none|console  --precision f64|mixed  --device n\n"
            << "     --verbose  --debug  --version  --help\n";
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

// Runs the phantom program for one view (phm2helix.cpp:273-282, 299-309).
void run_program(const std::string& prog, int view, int nview, const std::string& file, bool debug) {
  std::ostringstream cmd;
  cmd << prog << " " << view << " " << nview << " " << file;
  if (debug) std::cout << cmd.str() << std::endl;
  if (system(cmd.str().c_str()) != 0)
    std::cerr << "Error executing external phantom program " << prog << " with command "
              << cmd.str() << std::endl;
}

}  // namespace

int main(int argc, char* const argv[]) {
  Timer timerProgram;
  std::string geometry = Scanner::convertGeometryIDToName(Scanner::GEOMETRY_PARALLEL);
  std::string desc, phmProg, phmFile = "tmpphmfile";
  int nray = 1, offsetView = 0, trace = Trace::TRACE_NONE;
  double focal = 2., centerDet = 2., viewRatio = 1., scanRatio = 1., rotAngle = -1;
  bool verbose = false, debug = false;
  while (true) {
    const int c = getopt_long(argc, argv, "", kOptions, nullptr);
    if (c == -1) break;
    bool ok = true;
    switch (c) {
      case O_PHANTOMPROG: break;  // declared by the reference, not read
      case O_PHMFILE: phmFile = optarg; break;
      case O_DESC: desc = optarg; break;
      case O_NRAY: ok = to_int(optarg, nray); break;
      case O_ROTANGLE: ok = to_double(optarg, rotAngle); break;
      case O_GEOMETRY: geometry = optarg; break;
      case O_FOCAL_LENGTH: ok = to_double(optarg, focal); break;
      case O_CENTER_DETECTOR_LENGTH: ok = to_double(optarg, centerDet); break;
      case O_VIEW_RATIO: ok = to_double(optarg, viewRatio); break;
      case O_SCAN_RATIO: ok = to_double(optarg, scanRatio); break;
      case O_OFFSETVIEW: ok = to_int(optarg, offsetView); break;
      case O_VERBOSE: verbose = true; break;
      case O_DEBUG: debug = true; break;
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
      case O_VERSION: std::cout << "Version: " << ctsim_version() << std::endl; return 0;
      case O_HELP: case '?': usage(argv[0]); return 0;
      default: ok = false;
    }
    if (!ok) {
      usage(argv[0]);
      return 1;
    }
  }
  if (optind + 4 != argc) {
    usage(argv[0]);
    return 1;
  }
  const char* outFile = argv[optind];
  int nDet = 0, nView = 0;
  if (!to_int(argv[optind + 1], nDet) || !to_int(argv[optind + 2], nView)) {
    usage(argv[0]);
    return 1;
  }
  phmProg = argv[optind + 3];
  if (rotAngle < 0) rotAngle = geometry == "parallel" ? 0.5 : 1.0;

  std::ostringstream label;
  label << "phm2helix: NDet=" << nDet << ", Nview=" << nView << ", NRay=" << nray
        << ", RotAngle=" << rotAngle << ", OffsetView =" << offsetView << ", Geometry=" << geometry
        << ", PhantomProg=" << phmProg << ", PhmFileName=" << phmFile;
  if (!desc.empty()) label << ": " << desc;

  // the reference multiplies the turn count by 2*pi in double (phm2helix.cpp:270)
  rotAngle *= 6.28318530717958647692;
  Phantom phm;
  run_program(phmProg, 0, nView, phmFile, debug);
  phm.createFromFile(phmFile.c_str());
  remove(phmFile.c_str());
  Scanner scanner(phm, geometry.c_str(), nDet, nView, offsetView, nray, rotAngle, focal, centerDet,
                  viewRatio, scanRatio);
  if (scanner.fail()) {
    std::cout << "Scanner Creation Error: " << scanner.failMessage() << std::endl;
    return 1;
  }
  Projections pj(scanner);
  for (int v = 0; v < nView; v++) {
    run_program(phmProg, v, nView, phmFile, debug);
    Phantom phmtmp;
    phmtmp.createFromFile(phmFile.c_str());
    scanner.collectProjections(pj, phmtmp, v, 1, scanner.offsetView(), true, trace);
    remove(phmFile.c_str());
  }
  pj.setCalcTime(timerProgram.timerEnd());
  pj.setRemark(label.str());
  if (!pj.write(outFile)) {
    std::cerr << "Unable to write " << outFile << "\n";
    return 1;
  }
  if (verbose) {
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
