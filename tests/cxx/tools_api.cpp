// The calls the reference's command-line tools make on the library, in one
// translation unit that compiles against either implementation:
//   default          include/ctsim/ctsim.h + libctsim_amd.so (the drop-in)
//   -DCTSIM_REFERENCE the reference's include/ct.h + the reference library
//                     built from its sources (oracle/_ref/libctsimref.so)
// Both builds print the same text when the drop-in keeps the reference's
// interface and behaviour.  No call here needs a GPU.
//
//   tools_api <projections.pj> <phantom-file> <malformed-phantom-file>
//
// Calls, by the tool lines they come from:
//   tools/pjrec.cpp:128-137   default option names from the static ID API
//   tools/pjrec.cpp:204       Trace::convertTraceNameToID
//   tools/pjrec.cpp:256-260   Projections::read, printScanInfo
//   tools/pjrec.cpp:355-356   ImageFile::labelAdd(getLabel()), typed labelAdd
//   tools/phm2pj.cpp:104      Scanner::convertGeometryIDToName
//   tools/phm2pj.cpp:154      Trace::convertTraceNameToID
//   tools/phm2pj.cpp:305      Phantom::createFromFile
//   tools/phm2pj.cpp:349,385  Projections::printScanInfo
//   tools/phm2pj.cpp:354      Projections::setNView
//   tools/phm2pj.cpp:382      Phantom::print
//   tools/phm2pj.cpp:126,374  Timer
#ifdef CTSIM_REFERENCE
#include "ct.h"
#include "timer.h"
#else
#include "ctsim/ctsim.h"
#endif

#include <iostream>
#include <sstream>
#include <string>

static void names(const char* what, int count, const char* (*name)(int), const char* (*title)(int),
                  int (*id)(const char*)) {
  for (int i = -1; i < count; i++) {
    const char* n = name(i);
    std::cout << what << " " << i << " [" << n << "] [" << (title ? title(i) : "") << "] -> "
              << (i >= 0 ? id(n) : id("no-such-name")) << "\n";
  }
  std::cout << what << " 99 [" << name(99) << "] [" << (title ? title(99) : "") << "]\n";
}

static const char* filterName(int i) { return SignalFilter::convertFilterIDToName(i); }
static const char* filterTitle(int i) { return SignalFilter::convertFilterIDToTitle(i); }
static int filterID(const char* n) { return SignalFilter::convertFilterNameToID(n); }
static const char* domainName(int i) { return SignalFilter::convertDomainIDToName(i); }
static const char* domainTitle(int i) { return SignalFilter::convertDomainIDToTitle(i); }
static int domainID(const char* n) { return SignalFilter::convertDomainNameToID(n); }
static const char* methodName(int i) { return ProcessSignal::convertFilterMethodIDToName(i); }
static const char* methodTitle(int i) { return ProcessSignal::convertFilterMethodIDToTitle(i); }
static int methodID(const char* n) { return ProcessSignal::convertFilterMethodNameToID(n); }
static const char* genName(int i) { return ProcessSignal::convertFilterGenerationIDToName(i); }
static const char* genTitle(int i) { return ProcessSignal::convertFilterGenerationIDToTitle(i); }
static int genID(const char* n) { return ProcessSignal::convertFilterGenerationNameToID(n); }
static const char* bpName(int i) { return Backprojector::convertBackprojectIDToName(i); }
static const char* bpTitle(int i) { return Backprojector::convertBackprojectIDToTitle(i); }
static int bpID(const char* n) { return Backprojector::convertBackprojectNameToID(n); }
static const char* interpName(int i) { return Backprojector::convertInterpIDToName(i); }
static const char* interpTitle(int i) { return Backprojector::convertInterpIDToTitle(i); }
static int interpID(const char* n) { return Backprojector::convertInterpNameToID(n); }
static const char* geomName(int i) { return Scanner::convertGeometryIDToName(i); }
static const char* geomTitle(int i) { return Scanner::convertGeometryIDToTitle(i); }
static int geomID(const char* n) { return Scanner::convertGeometryNameToID(n); }
static const char* phmName(int i) { return Phantom::convertPhantomIDToName(i); }
static const char* phmTitle(int i) { return Phantom::convertPhantomIDToTitle(i); }
static int phmID(const char* n) { return Phantom::convertNameToPhantomID(n); }

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  Timer timerProgram;

  // pjrec.cpp:132-136: the default names pjrec builds from IDs
  std::string sOptFilterName(SignalFilter::convertFilterIDToName(SignalFilter::FILTER_ABS_BANDLIMIT));
  std::string sOptFilterMethodName(
      ProcessSignal::convertFilterMethodIDToName(ProcessSignal::FILTER_METHOD_CONVOLUTION));
  std::string sOptFilterGenerationName(
      ProcessSignal::convertFilterGenerationIDToName(ProcessSignal::FILTER_GENERATION_DIRECT));
  std::string sOptInterpName(Backprojector::convertInterpIDToName(Backprojector::INTERP_LINEAR));
  std::string sOptBackprojectName(Backprojector::convertBackprojectIDToName(Backprojector::BPROJ_IDIFF));
  std::cout << "pjrec defaults: " << sOptFilterName << " " << sOptFilterMethodName << " "
            << sOptFilterGenerationName << " " << sOptInterpName << " " << sOptBackprojectName << "\n";
  // phm2pj.cpp:104
  std::string optGeometryName = Scanner::convertGeometryIDToName(Scanner::GEOMETRY_PARALLEL);
  std::cout << "phm2pj default geometry: " << optGeometryName << "\n";
  // pjrec.cpp:204, phm2pj.cpp:154
  for (const char* t : {"none", "console", "CONSOLE", "phantom", "proj", "plot", "clipping", "bogus"})
    std::cout << "trace " << t << " -> " << Trace::convertTraceNameToID(t)
              << (Trace::convertTraceNameToID(t) == Trace::TRACE_INVALID ? " (invalid)" : "") << "\n";

  // every ID and name of the static tables the tools and dialogs use; the
  // method table stops at FFT because the reference here is built without
  // FFTW (procsignal.cpp:31-42), which has no fftw/rfftw entries
  names("filter", SignalFilter::FILTER_TRIANGLE + 1, filterName, filterTitle, filterID);
  names("domain", SignalFilter::DOMAIN_SPATIAL + 1, domainName, domainTitle, domainID);
  names("method", ProcessSignal::FILTER_METHOD_FFT + 1, methodName, methodTitle, methodID);
  names("generation", ProcessSignal::FILTER_GENERATION_INVERSE_FOURIER + 1, genName, genTitle, genID);
  names("backproj", Backprojector::BPROJ_IDIFF + 1, bpName, bpTitle, bpID);
  names("interp", Backprojector::INTERP_CUBIC + 1, interpName, interpTitle, interpID);
  names("geometry", Scanner::GEOMETRY_LINOGRAM + 1, geomName, geomTitle, geomID);
  names("phantom", Phantom::PHM_UNITPULSE + 1, phmName, phmTitle, phmID);
  std::cout << "shepp -> " << SignalFilter::convertFilterNameToID("shepp") << " (FILTER_ABS_SINC "
            << SignalFilter::FILTER_ABS_SINC << ")\n";
  std::cout << "counts: " << Backprojector::getBackprojectCount() << " "
            << Backprojector::getInterpCount() << " " << Scanner::getGeometryCount() << " "
            << ProcessSignal::getFilterGenerationCount() << "\n";

  // phm2pj.cpp:305, 382: a phantom file and its printout
  {
    Phantom phm;
    const bool ok = phm.createFromFile(argv[2]);
    std::cout << "createFromFile " << ok << " nPElem " << phm.nPElem() << " name " << phm.name() << "\n";
    phm.print(std::cout);
    std::ostringstream defs;
    phm.printDefinitions(defs);
    std::cout << defs.str();
  }
  // a file with a short record: warned about (on standard output), reported
  // false, still read on
  {
    Phantom phm;
    const bool ok = phm.createFromFile(argv[3]);
    std::cout << "createFromFile(malformed) " << ok << " nPElem " << phm.nPElem() << "\n";
    phm.print(std::cout);
  }
  // built-in phantoms by name
  for (const char* n : {"herman", "shepp-logan", "unit-pulse"}) {
    Phantom phm(n);
    std::cout << "builtin " << n << " fail " << phm.fail() << " id " << phm.id() << "\n";
    phm.print(std::cout);
  }
  {
    Phantom bad("no-such-phantom");
    std::cout << "builtin bad fail " << bad.fail() << " [" << bad.failMessage() << "]\n";
  }

  // pjrec.cpp:256-260, phm2pj.cpp:349: a projection file's scan header
  Projections projGlobal;
  if (!projGlobal.read(argv[1])) {
    std::cout << "read failed\n";
    return 1;
  }
  {
    std::ostringstream os;
    projGlobal.printScanInfo(os);
    std::cout << os.str();
  }
  const Array2dFileLabel& pl = projGlobal.getLabel();
  std::cout << "projection label type " << pl.getLabelType() << " time " << pl.getCalcTime()
            << " [" << pl.getLabelString() << "] " << pl.getDateString() << "\n";
  pl.print(std::cout);
  pl.printBrief(std::cout);

  // pjrec.cpp:355-356: the image's history labels
  ImageFile im(8, 6);
  const std::string sRemark = "pjrec: 8x6, " + sOptFilterName + ", " + sOptInterpName;
  im.labelAdd(projGlobal.getLabel());
  im.labelAdd(Array2dFileLabel::L_HISTORY, sRemark.c_str(), 1.25);
  im.labelAdd(Array2dFileLabel::L_USER, "a note");
  im.labelAdd("plain", 2.5);
  std::cout << "labels " << im.getNumLabels() << "\n";
  for (unsigned i = 0; i < im.getNumLabels(); i++) {
    const Array2dFileLabel& l = im.labelGet(i);
    std::cout << "label " << i << " type " << l.getLabelType() << " time " << l.getCalcTime() << " ["
              << l.getLabelString() << "]\n";
  }
  im.labelGet(0).print(std::cout);

  // phm2pj.cpp:354: restrict the views
  Projections pjLocal(projGlobal.nView(), projGlobal.nDet());
  pjLocal.setNView(3);
  std::cout << "setNView " << pjLocal.nView() << " " << pjLocal.nDet() << "\n";

  const double t = timerProgram.timerEnd();
  std::cout << "timer " << (t >= 0 && t < 60 ? "ok" : "bad") << "\n";
  return 0;
}
