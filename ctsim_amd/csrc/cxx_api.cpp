// The C++ drop-in classes of include/ctsim/ctsim.h over the C ABI.
#include "ctsim/ctsim.h"

#include <hip/hip_runtime_api.h>
#include <strings.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <algorithm>
#include <fstream>
#include <iostream>

namespace {

int g_device = 0;
int g_precision = -1;
ctsim_ctx* g_ctx = nullptr;

const double kSqrt2 = 1.414213562373095049;

// Little-endian field I/O (the reference's frnetorderstream on little-endian
// hosts, libctsupport/fnetorderstream.cpp:135-198).
template <class T>
void put(std::ostream& os, T v) {
  os.write(reinterpret_cast<const char*>(&v), sizeof v);
}
template <class T>
bool get(std::istream& is, T& v) {
  is.read(reinterpret_cast<char*>(&v), sizeof v);
  return static_cast<bool>(is);
}

void now_fields(int* f) {
  time_t t = time(nullptr);
  tm* lt = localtime(&t);
  f[0] = lt->tm_year; f[1] = lt->tm_mon; f[2] = lt->tm_mday;
  f[3] = lt->tm_hour; f[4] = lt->tm_min; f[5] = lt->tm_sec;
}

// sys_error(ERR_WARNING, ...) prints to standard output and goes on
// (libctsupport/syserror.cpp:45-77).
void warn(const std::string& msg) { std::cout << "WARNING ERROR: " << msg << "\n"; }

}  // namespace

namespace ctsim_amd {

void setDevice(int device) { g_device = device; }

void setPrecision(int p) { g_precision = p == CTSIM_PREC_MIXED ? CTSIM_PREC_MIXED : CTSIM_PREC_F64; }

int precision() {
  if (g_precision < 0) {
    const char* e = getenv("CTSIM_PRECISION");
    g_precision = (e && !strcasecmp(e, "mixed")) ? CTSIM_PREC_MIXED : CTSIM_PREC_F64;
  }
  return g_precision;
}

ctsim_ctx* context() {
  if (!g_ctx && ctsim_ctx_create(&g_ctx, g_device, nullptr) != CTSIM_OK) {
    std::cerr << "ctsim_amd: cannot create a device context: " << ctsim_last_error() << "\n";
    std::abort();
  }
  return g_ctx;
}

}  // namespace ctsim_amd

// ----------------------------------------------------------------- Trace
int Trace::convertTraceNameToID(const char* name) {  // libctsim/trace.cpp:41-61
  static const char* const names[] = {"none", "console", "phantom", "proj", "plot", "clipping"};
  for (int i = 0; i < 6; i++)
    if (name && !strcasecmp(name, names[i])) return i;
  return TRACE_INVALID;
}

// ------------------------------------------------------------------ Timer
void Timer::timerReport(const char* msg) const {
  std::cout << msg << ": " << m_timeElapsed << " seconds" << std::endl;
}

// ------------------------------------------------------- Array2dFileLabel
// array2dfile.cpp:44-158: a new label carries the local date and time.
Array2dFileLabel::Array2dFileLabel() {
  int f[6];
  now_fields(f);
  setDateTime(f[0], f[1], f[2], f[3], f[4], f[5]);
}

Array2dFileLabel::Array2dFileLabel(const char* str, double ctime) : Array2dFileLabel() {
  m_strLabel = str ? str : "";
  m_labelType = L_USER;
  m_calcTime = ctime;
}

Array2dFileLabel::Array2dFileLabel(int type, const char* str, double ctime) : Array2dFileLabel() {
  m_strLabel = str ? str : "";
  m_labelType = type;
  m_calcTime = ctime;
}

void Array2dFileLabel::setDateTime(int year, int month, int day, int hour, int minute, int second) {
  m_year = year; m_month = month; m_day = day;
  m_hour = hour; m_minute = minute; m_second = second;
}

void Array2dFileLabel::getDateTime(int& year, int& month, int& day, int& hour, int& minute,
                                   int& second) const {
  year = m_year; month = m_month; day = m_day;
  hour = m_hour; minute = m_minute; second = m_second;
}

// "M/DD/YYYY HH:MM:SS" with tm-style year and month fields.
const std::string& Array2dFileLabel::getDateString() const {
  char buf[128];
  snprintf(buf, sizeof buf, "%2d/%02d/%4d %02d:%02d:%02d", m_month + 1, m_day, m_year + 1900,
           m_hour, m_minute, m_second);
  m_strDate = buf;
  return m_strDate;
}

// The reference prints a literal "%s" before the user and unknown labels'
// timestamps (array2dfile.cpp:125-131); kept so the text matches.
void Array2dFileLabel::print(std::ostream& os) const {
  if (m_labelType == L_HISTORY) {
    os << "History: " << std::endl << "  " << m_strLabel << std::endl;
    os << "  calc time = " << m_calcTime << " secs" << std::endl;
    os << "  Timestamp = " << getDateString() << std::endl;
  } else {
    if (m_labelType == L_USER) os << "Note: " << m_strLabel << std::endl;
    else os << "Unknown (" << m_labelType << "): " << m_strLabel << std::endl;
    os << "  Timestamp = %s" << getDateString() << std::endl;
  }
}

void Array2dFileLabel::printBrief(std::ostream& os) const {
  if (m_labelType == L_HISTORY) {
    os << "History (";
    if (m_calcTime > 0) os << m_calcTime << " secs, ";
    os << getDateString() << "): " << m_strLabel << std::endl;
  } else if (m_labelType == L_USER) {
    os << "Note (" << getDateString() << "): " << m_strLabel << std::endl;
  } else {
    os << "Unknown (" << getDateString() << "): " << m_strLabel << std::endl;
  }
}

// ------------------------------------------------------ static name tables
namespace {

// Index of name in the table (case-insensitive), or -1.
int lookup(const char* const* table, int n, const char* name) {
  for (int i = 0; i < n; i++)
    if (name && !strcasecmp(name, table[i])) return i;
  return -1;
}

const char* entry(const char* const* table, int n, int id) {
  return id >= 0 && id < n ? table[id] : "";
}

// libctsim/filter.cpp:46-75 (names 4 and 5 swapped against their IDs)
const char* const kFilterNames[] = {"abs_bandlimit", "abs_hamming", "abs_hanning", "abs_cosine",
                                    "shepp", "abs_sinc", "bandlimit", "sinc", "hamming",
                                    "hanning", "cosine", "triangle"};
const char* const kFilterTitles[] = {"Abs(w) * Bandlimit", "Abs(w) * Hamming", "Abs(w) * Hanning",
                                     "Abs(w) * Cosine", "Shepp", "Abs(w) * Sinc", "Bandlimit",
                                     "Sinc", "Hamming", "Hanning", "Cosine", "Triangle"};
const char* const kDomainNames[] = {"frequency", "spatial"};
const char* const kDomainTitles[] = {"Frequency", "Spatial"};
// libctsim/procsignal.cpp:41-73 (with the FFTW entries)
const char* const kMethodNames[] = {"convolution", "fourier", "fouier-table", "fft", "fftw",
                                    "rfftw"};
const char* const kMethodTitles[] = {"Convolution", "Fourier", "Fouier Trigometric Table", "FFT",
                                     "FFTW", "Real/Half-Complex FFTW"};
const char* const kGenerationNames[] = {"direct", "inverse-fourier"};
const char* const kGenerationTitles[] = {"Direct", "Inverse Fourier"};
// libctsim/backprojectors.cpp:34-97
const char* const kBackprojNames[] = {"trig", "table", "diff", "idiff"};
const char* const kBackprojTitles[] = {"Direct Trigometric", "Trigometric Table",
                                       "Difference Iteration", "Integer Difference Iteration"};
const char* const kInterpNames[] = {"nearest", "linear", "cubic"};
const char* const kInterpTitles[] = {"Nearest", "Linear", "Cubic"};
// libctsim/scanner.cpp:35-49
const char* const kGeometryNames[] = {"parallel", "equiangular", "equilinear", "linogram"};
const char* const kGeometryTitles[] = {"Parallel", "Equiangular", "Equilinear", "Linogram"};
// libctsim/phantom.cpp:37-49
const char* const kPhantomNames[] = {"herman", "shepp-logan", "unit-pulse"};
const char* const kPhantomTitles[] = {"Herman Head", "Shepp-Logan", "Unit Pulse"};

template <size_t N>
constexpr int count(const char* const (&)[N]) { return static_cast<int>(N); }

}  // namespace

int SignalFilter::getFilterCount() { return count(kFilterNames); }
int SignalFilter::getReconstructFilterCount() { return 4; }  // filter.cpp:44
const char* const* SignalFilter::getFilterNameArray() { return kFilterNames; }
const char* const* SignalFilter::getFilterTitleArray() { return kFilterTitles; }
int SignalFilter::convertFilterNameToID(const char* n) { return lookup(kFilterNames, count(kFilterNames), n); }
const char* SignalFilter::convertFilterIDToName(int id) { return entry(kFilterNames, count(kFilterNames), id); }
const char* SignalFilter::convertFilterIDToTitle(int id) { return entry(kFilterTitles, count(kFilterTitles), id); }
int SignalFilter::getDomainCount() { return count(kDomainNames); }
const char* const* SignalFilter::getDomainNameArray() { return kDomainNames; }
const char* const* SignalFilter::getDomainTitleArray() { return kDomainTitles; }
int SignalFilter::convertDomainNameToID(const char* n) { return lookup(kDomainNames, count(kDomainNames), n); }
const char* SignalFilter::convertDomainIDToName(int id) { return entry(kDomainNames, count(kDomainNames), id); }
const char* SignalFilter::convertDomainIDToTitle(int id) { return entry(kDomainTitles, count(kDomainTitles), id); }

int ProcessSignal::getFilterMethodCount() { return count(kMethodNames); }
const char* const* ProcessSignal::getFilterMethodNameArray() { return kMethodNames; }
const char* const* ProcessSignal::getFilterMethodTitleArray() { return kMethodTitles; }
int ProcessSignal::convertFilterMethodNameToID(const char* n) { return lookup(kMethodNames, count(kMethodNames), n); }
const char* ProcessSignal::convertFilterMethodIDToName(int id) { return entry(kMethodNames, count(kMethodNames), id); }
const char* ProcessSignal::convertFilterMethodIDToTitle(int id) { return entry(kMethodTitles, count(kMethodTitles), id); }
int ProcessSignal::getFilterGenerationCount() { return count(kGenerationNames); }
const char* const* ProcessSignal::getFilterGenerationNameArray() { return kGenerationNames; }
const char* const* ProcessSignal::getFilterGenerationTitleArray() { return kGenerationTitles; }
int ProcessSignal::convertFilterGenerationNameToID(const char* n) { return lookup(kGenerationNames, count(kGenerationNames), n); }
const char* ProcessSignal::convertFilterGenerationIDToName(int id) { return entry(kGenerationNames, count(kGenerationNames), id); }
const char* ProcessSignal::convertFilterGenerationIDToTitle(int id) { return entry(kGenerationTitles, count(kGenerationTitles), id); }

int Backprojector::getBackprojectCount() { return count(kBackprojNames); }
const char* const* Backprojector::getBackprojectNameArray() { return kBackprojNames; }
const char* const* Backprojector::getBackprojectTitleArray() { return kBackprojTitles; }
int Backprojector::convertBackprojectNameToID(const char* n) { return lookup(kBackprojNames, count(kBackprojNames), n); }
const char* Backprojector::convertBackprojectIDToName(int id) { return entry(kBackprojNames, count(kBackprojNames), id); }
const char* Backprojector::convertBackprojectIDToTitle(int id) { return entry(kBackprojTitles, count(kBackprojTitles), id); }
int Backprojector::getInterpCount() { return count(kInterpNames); }
const char* const* Backprojector::getInterpNameArray() { return kInterpNames; }
const char* const* Backprojector::getInterpTitleArray() { return kInterpTitles; }
int Backprojector::convertInterpNameToID(const char* n) { return lookup(kInterpNames, count(kInterpNames), n); }
const char* Backprojector::convertInterpIDToName(int id) { return entry(kInterpNames, count(kInterpNames), id); }
const char* Backprojector::convertInterpIDToTitle(int id) { return entry(kInterpTitles, count(kInterpTitles), id); }

int Scanner::getGeometryCount() { return count(kGeometryNames); }
const char* const* Scanner::getGeometryNameArray() { return kGeometryNames; }
const char* const* Scanner::getGeometryTitleArray() { return kGeometryTitles; }
int Scanner::convertGeometryNameToID(const char* n) { return lookup(kGeometryNames, count(kGeometryNames), n); }
const char* Scanner::convertGeometryIDToName(int id) { return entry(kGeometryNames, count(kGeometryNames), id); }
// The reference returns the name here, not the title (scanner.cpp:242-251).
const char* Scanner::convertGeometryIDToTitle(int id) { return entry(kGeometryNames, count(kGeometryNames), id); }

int Phantom::getPhantomCount() { return count(kPhantomNames); }
const char* const* Phantom::getPhantomNameArray() { return kPhantomNames; }
const char* const* Phantom::getPhantomTitleArray() { return kPhantomTitles; }
int Phantom::convertNameToPhantomID(const char* n) { return lookup(kPhantomNames, count(kPhantomNames), n); }
const char* Phantom::convertPhantomIDToName(int id) { return entry(kPhantomNames, count(kPhantomNames), id); }
// The reference returns the name here, not the title (phantom.cpp:103-110).
const char* Phantom::convertPhantomIDToTitle(int id) { return entry(kPhantomNames, count(kPhantomNames), id); }

// ---------------------------------------------------------------- Phantom
Phantom::Phantom() { ctsim_phantom_create(&m_hip, nullptr); }

Phantom::Phantom(const char* phmName) {
  ctsim_phantom_create(&m_hip, nullptr);
  createFromPhantom(phmName);
}

Phantom::~Phantom() { ctsim_phantom_destroy(m_hip); }

bool Phantom::createFromPhantom(const char* phmName) {
  const int id = convertNameToPhantomID(phmName);
  if (id == PHM_INVALID) {
    m_fail = true;
    m_failMessage = std::string("Invalid phantom name ") + (phmName ? phmName : "");
    return false;
  }
  m_name = phmName;
  return createFromPhantom(id);
}

// phantom.cpp:136-170: the built-in's elements are added to this phantom
// (an empty phantom takes the built-in's plan as it is, unit-pulse flag
// included).
bool Phantom::createFromPhantom(int phmid) {
  if (phmid < 0 || phmid >= getPhantomCount()) {
    m_fail = true;
    m_failMessage = "Illegal phantom id " + std::to_string(phmid);
    return false;
  }
  ctsim_phantom* p = nullptr;
  if (ctsim_phantom_create(&p, kPhantomNames[phmid]) != CTSIM_OK) {
    m_fail = true;
    m_failMessage = ctsim_last_error();
    return false;
  }
  double e[13];
  const int n = ctsim_phantom_extent(p, e);
  if (m_nPElem == 0) {
    ctsim_phantom_destroy(m_hip);
    m_hip = p;
    m_nPElem = n;
  } else {
    static const char* const types[] = {"", "rectangle", "triangle", "ellipse", "sector", "segment"};
    for (int i = n - 1; i >= 0; i--) {  // definition order (the list is newest-first)
      ctsim_phantom_element(p, i, e);
      addPElem(types[static_cast<int>(e[0])], e[1], e[2], e[3], e[4], e[12], e[6]);
    }
    ctsim_phantom_destroy(p);
  }
  m_id = phmid;
  return true;
}

// Phantom::createFromFile, phantom.cpp:184-213.  A record with fewer than
// seven fields warns, makes the result false and is still added; reading
// then resumes at the next token.  The reference reads indeterminate locals
// for the fields such a record lacks; here they keep the previous record's
// values (zero before the first record).
bool Phantom::createFromFile(const char* fname) {
  FILE* fp = fopen(fname, "r");
  if (!fp) return false;
  m_name = fname;
  bool good = true;
  double cx = 0, cy = 0, u = 0, v = 0, rot = 0, dens = 0;
  while (true) {
    char type[80] = "";
    const int status = fscanf(fp, "%79s %lf %lf %lf %lf %lf %lf", type, &cx, &cy, &u, &v, &rot, &dens);
    if (status == EOF) break;
    if (status != 7) {
      warn(std::string("Insufficient fields reading phantom file ") + fname +
           " [Phantom::createFromFile]");
      good = false;
    }
    addPElem(type, cx, cy, u, v, rot, dens);
  }
  fclose(fp);
  return good;
}

// Phantom::fileWrite, phantom.cpp:215-223: the definitions as text.
bool Phantom::fileWrite(const char* fname) {
  std::ofstream fs(fname);
  if (fs) printDefinitions(fs);
  return static_cast<bool>(fs);
}

namespace {
const char* elem_type_name(int t) {
  static const char* const names[] = {"Unknown", "rectangle", "triangle", "ellipse", "sector", "segment"};
  return t >= 1 && t <= 5 ? names[t] : names[0];
}
}  // namespace

void Phantom::print(std::ostream& os) const {
  double e[13];
  os << "Number of PElements: " << m_nPElem << "\n";
  os << "Limits: xmin=" << xmin() << ", ymin=" << ymin() << ", xmax=" << xmax() << ", ymax=" << ymax() << "\n";
  for (int i = 0; i < m_nPElem; i++) {
    ctsim_phantom_element(m_hip, i, e);
    os << "PhantomElement: nPoints=" << static_cast<int>(e[11]);
    os << ", atten=" << e[6] << " rot=" << e[5] * (180. / 3.14159265358979323846) << "\n";
    os << "xmin=" << e[7] << ", ymin=" << e[8] << ", xmax=" << e[9] << ", ymax=" << e[10] << "\n";
  }
}

void Phantom::printDefinitions(std::ostream& os) const {
  double e[13];
  for (int i = 0; i < m_nPElem; i++) {
    ctsim_phantom_element(m_hip, i, e);
    os << elem_type_name(static_cast<int>(e[0])) << " " << e[1] << " " << e[2] << " " << e[3]
       << " " << e[4] << " " << e[5] * (180. / 3.14159265358979323846) << " " << e[6] << "\n";
  }
}

void Phantom::addPElem(const char* type, double cx, double cy, double u, double v, double rot,
                       double atten) {
  if (ctsim_phantom_add_element(m_hip, type, cx, cy, u, v, rot, atten) != CTSIM_OK) {
    warn(ctsim_last_error());
    return;
  }
  m_nPElem++;
}

double Phantom::xmin() const { double e[4]; ctsim_phantom_extent(m_hip, e); return e[0]; }
double Phantom::xmax() const { double e[4]; ctsim_phantom_extent(m_hip, e); return e[1]; }
double Phantom::ymin() const { double e[4]; ctsim_phantom_extent(m_hip, e); return e[2]; }
double Phantom::ymax() const { double e[4]; ctsim_phantom_extent(m_hip, e); return e[3]; }
double Phantom::maxAxisLength() const {
  double e[4];
  ctsim_phantom_extent(m_hip, e);
  return (e[1] - e[0]) > (e[3] - e[2]) ? (e[1] - e[0]) : (e[3] - e[2]);
}
double Phantom::getDiameterBoundaryCircle() const { return kSqrt2 * maxAxisLength(); }

void Phantom::convertToImagefile(ImageFile& im, double dViewRatio, int nsample, int) const {
  const int nx = im.nx(), ny = im.ny();
  if (nx < 2 || ny < 2) return;  // phantom.cpp:469-470
  ctsim_ctx* ctx = ctsim_amd::context();
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * nx * ny) != hipSuccess) {
    warn("convertToImagefile: device allocation failed");
    return;
  }
  double ext[4];
  if (ctsim_phantom_rasterize(ctx, m_hip, nx, ny, dViewRatio, nsample, d, ext) == CTSIM_OK) {
    (void)hipMemcpy(im.data(), d, sizeof(float) * nx * ny, hipMemcpyDeviceToHost);
    im.setAxisExtent(ext[0], ext[1], ext[2], ext[3]);
    im.setAxisIncrement((ext[1] - ext[0]) / nx, (ext[3] - ext[2]) / ny);
  } else {
    warn(ctsim_last_error());
  }
  (void)hipFree(d);
}

// -------------------------------------------------------------- Scanner
Scanner::Scanner(const Phantom& phm, const char* geometryName, int nDet, int nView, int iOffsetView,
                 int nSample, double rot_anglen, double dFocalLengthRatio,
                 double dCenterDetectorRatio, double dViewRatio, double dScanRatio)
    : m_nDet(nDet), m_nView(nView), m_offsetView(iOffsetView) {
  if (ctsim_scanner_create(&m_hip, phm.handle(), geometryName, nDet, nView, iOffsetView, nSample,
                           rot_anglen, dFocalLengthRatio, dCenterDetectorRatio, dViewRatio,
                           dScanRatio) != CTSIM_OK) {
    m_fail = true;
    m_failMessage = ctsim_last_error();
    return;
  }
  ctsim_scanner_info(m_hip, m_info);
  m_geometry = static_cast<int>(m_info[11]);
}

Scanner::~Scanner() {
  if (m_dSino) (void)hipFree(m_dSino);
  ctsim_scanner_destroy(m_hip);
}

void Scanner::setNView(int nView) { m_nView = nView; }
void Scanner::setOffsetView(int iOffsetView) { m_offsetView = iOffsetView; }

void Scanner::collectProjections(Projections& proj, const Phantom& phm, int trace, SGP* pSGP) {
  collectProjections(proj, phm, 0, proj.nView(), m_offsetView, true, trace, pSGP);
}

void Scanner::collectProjections(Projections& proj, const Phantom& phm, int iStartView,
                                 int iNumViews, int iOffsetView, bool bStoreAtViewPosition,
                                 int trace, SGP* pSGP) {
  collectProjections(proj, phm, iStartView, iNumViews, iOffsetView,
                     bStoreAtViewPosition ? iStartView : 0, trace, pSGP);
}

void Scanner::collectProjections(Projections& proj, const Phantom& phm, int iStartView,
                                 int iNumViews, int iOffsetView, int iStorageOffset, int trace,
                                 SGP*) {
  if (m_fail || iNumViews <= 0) return;
  const int rows = proj.nView();
  if (iStorageOffset + iNumViews > rows || proj.nDet() != m_nDet) {
    warn("collectProjections: projections do not fit the scan");
    return;
  }
  ctsim_ctx* ctx = ctsim_amd::context();
  // one device buffer of the projections' rows, kept for later calls (a
  // time-varying scan collects one view per call)
  if (m_dRows < static_cast<size_t>(rows)) {
    if (m_dSino) (void)hipFree(m_dSino);
    m_dSino = nullptr;
    m_dRows = 0;
    if (hipMalloc(&m_dSino, sizeof(float) * static_cast<size_t>(rows) * m_nDet) != hipSuccess) {
      m_dSino = nullptr;
      warn("collectProjections: device allocation failed");
      return;
    }
    m_dRows = rows;
  }
  float* d = m_dSino;
  std::vector<double> ang(rows);
  const int rc = ctsim_collect_projections(ctx, m_hip, phm.handle(), ctsim_amd::precision(),
                                           iStartView, iNumViews, iOffsetView, iStorageOffset, d,
                                           ang.data());
  if (rc == CTSIM_OK) {
    ctsim_ctx_synchronize(ctx);
    float* h = proj.data() + static_cast<size_t>(iStorageOffset) * m_nDet;
    (void)hipMemcpy(h, d + static_cast<size_t>(iStorageOffset) * m_nDet,
                    sizeof(float) * static_cast<size_t>(iNumViews) * m_nDet, hipMemcpyDeviceToHost);
    for (int v = iStorageOffset; v < iStorageOffset + iNumViews; v++) {
      proj.getDetectorArray(v).setViewAngle(ang[v]);
      if (trace == Trace::TRACE_CONSOLE) std::cout << "Current View: " << v << std::endl;
    }
  } else {
    warn(ctsim_last_error());
  }
}

// ---------------------------------------------------------- Projections
Projections::Projections() { init(0, 0); }
Projections::Projections(int nView, int nDet) { init(nView, nDet); }
Projections::Projections(const Scanner& s) { initFromScanner(s); }

void Projections::init(int nView, int nDet) {
  m_label.setLabelType(Array2dFileLabel::L_HISTORY);  // projections.cpp:128-143
  m_nView = nView;
  m_nDet = nDet;
  m_data.assign(static_cast<size_t>(nView) * nDet, 0.f);
  m_views.assign(nView, DetectorArray());
  for (int v = 0; v < nView; v++) {
    m_views[v].m_detValues = m_data.data() + static_cast<size_t>(v) * nDet;
    m_views[v].m_nDet = nDet;
  }
  int f[6];
  now_fields(f);
  m_year = f[0]; m_month = f[1]; m_day = f[2]; m_hour = f[3]; m_minute = f[4]; m_second = f[5];
}

// Projections::initFromScanner, projections.cpp:145-161
void Projections::initFromScanner(const Scanner& s) {
  init(s.nView(), s.nDet());
  m_rotInc = s.rotInc();
  m_detInc = s.detInc();
  m_detStart = s.detStart();
  m_geometry = s.geometry();
  m_dFocalLength = s.focalLength();
  m_dSourceDetectorLength = s.sourceDetectorLength();
  m_dViewDiameter = s.viewDiameter();
  m_rotStart = s.offsetView() * s.rotInc();
  m_dFanBeamAngle = s.fanBeamAngle();
}

double Projections::phmLen() const { return m_dViewDiameter / kSqrt2; }
double Projections::viewLen() const { return m_dViewDiameter / kSqrt2; }

void Projections::setNView(int nView) { init(nView, m_nDet); }

void Projections::printScanInfo(std::ostringstream& os) const {
  os << "Number of detectors: " << m_nDet << "\n";
  os << "Number of views: " << m_nView << "\n";
  os << "Description: " << m_remark.c_str() << "\n";
  os << "Geometry: " << Scanner::convertGeometryIDToName(m_geometry) << "\n";
  os << "Focal Length: " << m_dFocalLength << "\n";
  os << "Source Detector Length: " << m_dSourceDetectorLength << "\n";
  os << "View Diameter: " << m_dViewDiameter << "\n";
  os << "Fan Beam Angle: " << m_dFanBeamAngle * (180. / 3.14159265358979323846) << "\n";
  os << "detStart: " << m_detStart << "\n";
  os << "detInc: " << m_detInc << "\n";
  os << "rotStart: " << m_rotStart << "\n";
  os << "rotInc: " << m_rotInc << "\n";
}
void Projections::setPhmLen(double p) { m_dViewDiameter = p * kSqrt2; }

// .pj layout, projections.cpp:485-546 (header) and 816-889 (views).
bool Projections::write(const char* fname) {
  std::ofstream fs(fname, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!fs) {
    warn(std::string("Error opening file ") + fname + " for output [projections_create]");
    return false;
  }
  const uint16_t sig = 'P' * 256 + 'J';
  const uint16_t rlen = static_cast<uint16_t>(m_remark.size());
  const uint16_t hsize = 2 + 2 + 4 + 4 + 4 + 9 * 8 + 7 * 2 + rlen;
  put<uint16_t>(fs, hsize);
  put<uint16_t>(fs, sig);
  put<uint32_t>(fs, m_nView);
  put<uint32_t>(fs, m_nDet);
  put<uint32_t>(fs, m_geometry);
  for (double d : {m_calcTime, m_rotStart, m_rotInc, m_detStart, m_detInc, m_dViewDiameter,
                   m_dFocalLength, m_dSourceDetectorLength, m_dFanBeamAngle})
    put<double>(fs, d);
  for (int f : {m_year, m_month, m_day, m_hour, m_minute, m_second}) put<uint16_t>(fs, f);
  put<uint16_t>(fs, rlen);
  fs.write(m_remark.data(), rlen);
  for (int v = 0; v < m_nView; v++) {
    put<double>(fs, m_views[v].viewAngle());
    put<uint32_t>(fs, m_nDet);
    fs.write(reinterpret_cast<const char*>(m_views[v].detValues()), sizeof(float) * m_nDet);
  }
  return static_cast<bool>(fs);
}

bool Projections::read(const char* fname) {
  std::ifstream fs(fname, std::ios::in | std::ios::binary);
  if (!fs) return false;
  uint16_t hsize, sig, f16[7];
  uint32_t nView, nDet, geom;
  double d[9];
  if (!get(fs, hsize) || !get(fs, sig) || !get(fs, nView) || !get(fs, nDet) || !get(fs, geom))
    return false;
  for (double& x : d)
    if (!get(fs, x)) return false;
  for (uint16_t& x : f16)
    if (!get(fs, x)) return false;
  if (sig != 'P' * 256 + 'J') {
    warn(std::string("File ") + fname + " does not have a valid projection file signature");
    return false;
  }
  std::string remark(f16[6], '\0');
  fs.read(&remark[0], f16[6]);
  if (!fs || static_cast<long>(fs.tellg()) != hsize) return false;
  init(nView, nDet);
  m_geometry = geom;
  m_calcTime = d[0]; m_rotStart = d[1]; m_rotInc = d[2]; m_detStart = d[3]; m_detInc = d[4];
  m_dViewDiameter = d[5]; m_dFocalLength = d[6]; m_dSourceDetectorLength = d[7]; m_dFanBeamAngle = d[8];
  m_year = f16[0]; m_month = f16[1]; m_day = f16[2]; m_hour = f16[3]; m_minute = f16[4]; m_second = f16[5];
  m_remark = remark;
  m_filename = fname;
  // projections.cpp:628-634
  m_label.setLabelType(Array2dFileLabel::L_HISTORY);
  m_label.setLabelString(m_remark);
  m_label.setCalcTime(m_calcTime);
  m_label.setDateTime(m_year, m_month, m_day, m_hour, m_minute, m_second);
  for (int v = 0; v < m_nView; v++) {
    double a;
    uint32_t nd;
    if (!get(fs, a) || !get(fs, nd)) break;
    m_views[v].setViewAngle(a);
    fs.read(reinterpret_cast<char*>(m_views[v].detValues()), sizeof(float) * nd);
    if (!fs) break;
  }
  return true;
}

namespace {

// Runs fn(ctx, d_in, d_out) on a device copy of the projections' data.
template <class F>
int on_device(const std::vector<float>& in, std::vector<float>& out, size_t outElems, F fn) {
  ctsim_ctx* ctx = ctsim_amd::context();
  float *d_in = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_in, sizeof(float) * std::max<size_t>(in.size(), 1)) != hipSuccess ||
      hipMalloc(&d_out, sizeof(float) * std::max<size_t>(outElems, 1)) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    warn("device allocation failed");
    return CTSIM_ERR_HIP;
  }
  (void)hipMemcpy(d_in, in.data(), sizeof(float) * in.size(), hipMemcpyHostToDevice);
  int rc = fn(ctx, d_in, d_out);
  if (rc == CTSIM_OK) {
    ctsim_ctx_synchronize(ctx);
    out.resize(outElems);
    (void)hipMemcpy(out.data(), d_out, sizeof(float) * outElems, hipMemcpyDeviceToHost);
  } else {
    warn(ctsim_last_error());
  }
  (void)hipFree(d_in);
  (void)hipFree(d_out);
  return rc;
}

}  // namespace

// Projections::Helical180LI, projections.cpp:200-358
int Projections::Helical180LI(int interpView) {
  if (m_geometry == Scanner::GEOMETRY_PARALLEL) {
    std::cerr << "Helical 180LI not yet implemented for PARALLEL geometry" << std::endl;
    return 2;
  }
  const int cap = static_cast<int>((3.14159265358979323846 + m_dFanBeamAngle) / m_rotInc) + 2;
  std::vector<float> out;
  std::vector<double> ang(cap);
  int n = 0;
  const int rc = on_device(m_data, out, static_cast<size_t>(cap) * m_nDet,
                           [&](ctsim_ctx* ctx, const float* din, float* dout) {
                             return ctsim_helical180li(ctx, m_geometry, m_nDet, m_nView, m_rotInc,
                                                       m_detInc, m_dFanBeamAngle,
                                                       m_dSourceDetectorLength, interpView, din,
                                                       dout, cap, &n, ang.data());
                           });
  if (rc != CTSIM_OK) return 1;
  const int nDet = m_nDet;
  init(n, nDet);
  std::copy(out.begin(), out.begin() + static_cast<size_t>(n) * nDet, m_data.begin());
  for (int v = 0; v < n; v++) m_views[v].setViewAngle(ang[v]);
  return 0;
}

// Projections::HalfScanFeather, projections.cpp:369-439
int Projections::HalfScanFeather() {
  if (m_geometry == Scanner::GEOMETRY_PARALLEL) {
    std::cerr << "HalfScanFeather not yet implemented for PARALLEL geometry" << std::endl;
    return 2;
  }
  std::vector<float> out;
  const int rc = on_device(m_data, out, m_data.size(),
                           [&](ctsim_ctx* ctx, const float* din, float* dout) {
                             return ctsim_halfscan_feather(ctx, m_geometry, m_nDet, m_nView,
                                                           m_rotInc, m_detInc, m_dFanBeamAngle,
                                                           m_dSourceDetectorLength, din, dout);
                           });
  if (rc != CTSIM_OK) return 1;
  std::copy(out.begin(), out.end(), m_data.begin());
  return 0;
}

// Projections::interpolateToParallel, projections.cpp:1250-1328
Projections* Projections::interpolateToParallel() const {
  if (m_geometry == Scanner::GEOMETRY_PARALLEL) return const_cast<Projections*>(this);
  std::vector<double> angles(m_nView);
  for (int v = 0; v < m_nView; v++) angles[v] = m_views[v].viewAngle();
  ctsim_rebin* rb = nullptr;
  if (ctsim_rebin_create(ctsim_amd::context(), &rb, m_geometry, m_nDet, m_nView, m_detStart,
                         m_detInc, m_dFocalLength, m_dSourceDetectorLength, m_dViewDiameter,
                         m_rotStart, angles.data()) != CTSIM_OK) {
    warn(ctsim_last_error());
    return nullptr;
  }
  double info[4];
  std::vector<double> newAngles(m_nView);
  ctsim_rebin_info(rb, info, newAngles.data());
  std::vector<float> out;
  const int rc = on_device(m_data, out, m_data.size(),
                           [&](ctsim_ctx* ctx, const float* din, float* dout) {
                             return ctsim_rebin_run(ctx, rb, din, dout);
                           });
  ctsim_rebin_destroy(rb);
  if (rc != CTSIM_OK) return nullptr;
  auto* p = new Projections(m_nView, m_nDet);
  p->m_geometry = Scanner::GEOMETRY_PARALLEL;
  p->m_dFocalLength = m_dFocalLength;
  p->m_dSourceDetectorLength = m_dSourceDetectorLength;
  p->m_dViewDiameter = m_dViewDiameter;
  p->m_dFanBeamAngle = m_dFanBeamAngle;
  p->m_calcTime = 0;
  p->m_remark = m_remark + "; Interpolate to Parallel";
  p->m_label.setLabelString(p->m_remark);  // projections.cpp:1264-1270
  p->m_label.setCalcTime(0);
  p->m_label.setDateTime(p->m_year, p->m_month, p->m_day, p->m_hour, p->m_minute, p->m_second);
  p->m_rotStart = info[0];
  p->m_rotInc = info[1];
  p->m_detStart = info[2];
  p->m_detInc = info[3];
  std::copy(out.begin(), out.end(), p->m_data.begin());
  for (int v = 0; v < m_nView; v++) p->m_views[v].setViewAngle(newAngles[v]);
  return p;
}

// ------------------------------------------------------------ ImageFile
ImageFile::ImageFile() {}
ImageFile::ImageFile(int nx, int ny) { setArraySize(nx, ny); }

bool ImageFile::setArraySize(int nx, int ny) {
  m_nx = nx;
  m_ny = ny;
  m_data.assign(static_cast<size_t>(nx) * ny, 0.f);
  m_cols.resize(nx);
  for (int ix = 0; ix < nx; ix++) m_cols[ix] = m_data.data() + static_cast<size_t>(ix) * ny;
  return true;
}

void ImageFile::arrayDataClear() { std::fill(m_data.begin(), m_data.end(), 0.f); }

void ImageFile::setAxisIncrement(double incX, double incY) {
  m_axisIncKnown = true;
  m_incX = incX;
  m_incY = incY;
}

// Array2dFile::setAxisExtent (array2dfile.cpp:390-397) stores only minX and
// maxY (maxX and minY keep their initial 0); files match the reference's.
void ImageFile::setAxisExtent(double minX, double maxX, double minY, double maxY) {
  (void)maxX;
  (void)minY;
  m_axisExtentKnown = true;
  m_minX = minX;
  m_maxY = maxY;
}

void ImageFile::labelAdd(const char* label, double calcTime) {
  labelAdd(Array2dFileLabel::L_HISTORY, label, calcTime);
}

void ImageFile::labelAdd(int type, const char* label, double calcTime) {
  labelAdd(Array2dFileLabel(type, label, calcTime));
}

void ImageFile::labelAdd(const Array2dFileLabel& label) { m_labels.push_back(label); }

// .if layout, array2dfile.cpp:443-520 (header, column-major data) and
// 601-633 (labels); pixel format 7 = float32, data type 1 = real.
bool ImageFile::fileWrite(const char* fname) {
  std::ofstream fs(fname, std::ios::out | std::ios::binary | std::ios::trunc);
  if (!fs) return false;
  const uint16_t hsize = 5 * 2 + 2 * 4 + 2 + 2 + 2 * 8 + 2 + 6 * 8;
  put<uint16_t>(fs, hsize);
  put<uint16_t>(fs, 'I' * 256 + 'F');
  put<uint16_t>(fs, 7);
  put<uint16_t>(fs, 4);
  put<uint16_t>(fs, static_cast<uint16_t>(m_labels.size()));
  put<uint32_t>(fs, m_nx);
  put<uint32_t>(fs, m_ny);
  put<uint16_t>(fs, 1);
  put<uint16_t>(fs, m_axisIncKnown ? 1 : 0);
  put<double>(fs, m_incX);
  put<double>(fs, m_incY);
  put<uint16_t>(fs, m_axisExtentKnown ? 1 : 0);
  for (double d : {m_minX, m_maxX, m_minY, m_maxY, 0.0, 1.0}) put<double>(fs, d);
  fs.write(reinterpret_cast<const char*>(m_data.data()), sizeof(float) * m_data.size());
  for (const Array2dFileLabel& l : m_labels) {
    int f[6];
    l.getDateTime(f[0], f[1], f[2], f[3], f[4], f[5]);
    put<uint16_t>(fs, l.getLabelType());
    for (int k = 0; k < 6; k++) put<uint16_t>(fs, f[k]);
    put<double>(fs, l.getCalcTime());
    const std::string& text = l.getLabelString();
    put<uint16_t>(fs, static_cast<uint16_t>(text.size()));
    fs.write(text.data(), text.size());
  }
  return static_cast<bool>(fs);
}

bool ImageFile::fileRead(const char* fname) {
  std::ifstream fs(fname, std::ios::in | std::ios::binary);
  if (!fs) return false;
  uint16_t hsize, sig, fmt, psize, nlab, dtype, incKnown, extKnown;
  uint32_t nx, ny;
  double inc[2], ext[6];
  if (!get(fs, hsize) || !get(fs, sig) || !get(fs, fmt) || !get(fs, psize) || !get(fs, nlab) ||
      !get(fs, nx) || !get(fs, ny) || !get(fs, dtype) || !get(fs, incKnown) || !get(fs, inc[0]) ||
      !get(fs, inc[1]) || !get(fs, extKnown))
    return false;
  for (double& d : ext)
    if (!get(fs, d)) return false;
  if (sig != 'I' * 256 + 'F' || fmt != 7 || psize != 4 || dtype != 1) {
    warn(std::string(fname) + ": only real float32 images are supported");
    return false;
  }
  setArraySize(nx, ny);
  fs.read(reinterpret_cast<char*>(m_data.data()), sizeof(float) * m_data.size());
  m_axisIncKnown = incKnown;
  m_incX = inc[0]; m_incY = inc[1];
  m_axisExtentKnown = extKnown;
  m_minX = ext[0]; m_maxX = ext[1]; m_minY = ext[2]; m_maxY = ext[3];
  m_labels.clear();
  for (int i = 0; i < nlab && fs; i++) {
    uint16_t t, dt[6], len;
    double calcTime = 0;
    get(fs, t);
    for (uint16_t& x : dt) get(fs, x);
    get(fs, calcTime);
    get(fs, len);
    std::string text(len, '\0');
    fs.read(&text[0], len);
    Array2dFileLabel l(t, text.c_str(), calcTime);
    l.setDateTime(dt[0], dt[1], dt[2], dt[3], dt[4], dt[5]);
    m_labels.push_back(l);
  }
  return static_cast<bool>(fs) || fs.eof();
}

// ImageFile::comparativeStatistics, imagefile.cpp:185-240 (same sums, same order).
bool ImageFile::comparativeStatistics(const ImageFile& c, double& d, double& r, double& e) const {
  if (c.nx() != m_nx && c.ny() != m_ny) {
    warn("Image sizes differ [ImageFile::comparativeStatistics]");
    return false;
  }
  if (m_nx == 0 || m_ny == 0) return false;
  const float* const* v = getArray();
  const float* const* w = c.getArray();
  double mean = 0.;
  for (unsigned ix = 0; ix < m_nx; ix++)
    for (unsigned iy = 0; iy < m_ny; iy++) mean += v[ix][iy];
  mean /= (m_nx * m_ny);
  double sq = 0., ab = 0., sqm = 0., av = 0.;
  for (unsigned ix = 0; ix < m_nx; ix++)
    for (unsigned iy = 0; iy < m_ny; iy++) {
      const double diff = v[ix][iy] - w[ix][iy];
      sq += diff * diff;
      ab += fabs(diff);
      const double dm = v[ix][iy] - mean;
      sqm += dm * dm;
      av += fabs(v[ix][iy]);
    }
  d = ::sqrt(sq / sqm);
  r = ab / av;
  const int hx = m_nx / 2, hy = m_ny / 2;
  double emax = -1;
  for (int ix = 0; ix < hx; ix++)
    for (int iy = 0; iy < hy; iy++) {
      const double a = 0.25 * (v[2 * ix][2 * iy] + v[2 * ix + 1][2 * iy] + v[2 * ix][2 * iy + 1] +
                               v[2 * ix + 1][2 * iy + 1]);
      const double b = 0.25 * (w[2 * ix][2 * iy] + w[2 * ix + 1][2 * iy] + w[2 * ix][2 * iy + 1] +
                               w[2 * ix + 1][2 * iy + 1]);
      const double err = fabs(a - b);
      if (err > emax) emax = err;
    }
  e = emax;
  return true;
}

// -------------------------------------------------------- ProcessSignal
ProcessSignal::ProcessSignal(const char* szFilterName, const char* szFilterMethodName,
                             double dBandwidth, double dSignalIncrement, int nSignalPoints,
                             double dFilterParam, const char* szDomainName,
                             const char* szFilterGenerationName, int iZeropad,
                             int iPreinterpolationFactor, int, int iGeometry, double dFocalLength,
                             double dSourceDetectorLength, SGP*)
    : m_n(nSignalPoints) {
  if (ctsim_procsignal_create(ctsim_amd::context(), &m_hip, szFilterName, szFilterMethodName,
                              dBandwidth, dSignalIncrement, nSignalPoints, dFilterParam,
                              szDomainName, szFilterGenerationName, iZeropad,
                              iPreinterpolationFactor, iGeometry, dFocalLength,
                              dSourceDetectorLength) != CTSIM_OK) {
    m_fail = true;
    m_failMessage = ctsim_last_error();
    return;
  }
  ctsim_procsignal_info(m_hip, &m_nFilterPoints, m_minMaxInc);
  m_filter.resize(m_nFilterPoints);
  ctsim_procsignal_get_filter(m_hip, m_filter.data(), m_nFilterPoints);
}

ProcessSignal::~ProcessSignal() {
  if (m_dIn) (void)hipFree(m_dIn);
  if (m_dOut) (void)hipFree(m_dOut);
  ctsim_procsignal_destroy(m_hip);
}

void ProcessSignal::filterSignal(const float input[], double output[]) const {
  if (m_fail) return;
  ctsim_ctx* ctx = ctsim_amd::context();
  if (!m_dIn) {
    (void)hipMalloc(&m_dIn, sizeof(float) * m_n);
    (void)hipMalloc(&m_dOut, sizeof(double) * m_n);
  }
  (void)hipMemcpy(m_dIn, input, sizeof(float) * m_n, hipMemcpyHostToDevice);
  if (ctsim_filter_views(ctx, m_hip, CTSIM_PREC_F64, static_cast<const float*>(m_dIn), 1, m_dOut) != CTSIM_OK) {
    warn(ctsim_last_error());
    return;
  }
  ctsim_ctx_synchronize(ctx);
  (void)hipMemcpy(output, m_dOut, sizeof(double) * m_n, hipMemcpyDeviceToHost);
}

// -------------------------------------------------------- Backprojector
Backprojector::Backprojector(const Projections& proj, ImageFile& im, const char* backprojName,
                             const char* interpName, int interpFactor,
                             const ReconstructionROI* pROI)
    : m_im(im), m_nDet(proj.nDet() * interpFactor) {
  double roi[4];
  if (pROI) {
    roi[0] = pROI->m_dXMin; roi[1] = pROI->m_dYMin; roi[2] = pROI->m_dXMax; roi[3] = pROI->m_dYMax;
  }
  if (ctsim_backprojector_create(ctsim_amd::context(), &m_hip, proj.geometry(), proj.nDet(),
                                 proj.detInc(), proj.nView(), proj.viewDiameter(),
                                 proj.focalLength(), proj.sourceDetectorLength(), backprojName,
                                 interpName, interpFactor, im.nx(), im.ny(),
                                 pROI ? roi : nullptr) != CTSIM_OK) {
    m_fail = true;
    m_failMessage = ctsim_last_error();
    return;
  }
  double g[8];
  ctsim_backprojector_info(m_hip, g);
  im.arrayDataClear();  // backprojectors.cpp:282
  im.setAxisIncrement(g[4], g[5]);
  im.setAxisExtent(g[0], g[1], g[2], g[3]);
  const size_t bytes = sizeof(float) * im.nx() * im.ny();
  if (hipMalloc(&m_dImage, bytes) != hipSuccess || hipMemset(m_dImage, 0, bytes) != hipSuccess) {
    m_fail = true;
    m_failMessage = "Backprojector: device image allocation failed";
  }
}

Backprojector::~Backprojector() {
  if (!m_fail && !m_postDone) {
    flush();
    download();
  }
  if (m_dViews) (void)hipFree(m_dViews);
  if (m_dImage) (void)hipFree(m_dImage);
  ctsim_backprojector_destroy(m_hip);
}

void Backprojector::BackprojectView(const double* viewData, double viewAngle) {
  if (m_fail) return;
  m_queue.insert(m_queue.end(), viewData, viewData + m_nDet);
  m_angles.push_back(viewAngle);
  if (m_angles.size() >= 256) flush();
}

int Backprojector::flush() {
  const int n = static_cast<int>(m_angles.size());
  if (n == 0) return CTSIM_OK;
  const int prec = ctsim_amd::precision();
  const size_t elem = prec == CTSIM_PREC_F64 ? sizeof(double) : sizeof(float);
  if (n > m_capacity) {
    if (m_dViews) (void)hipFree(m_dViews);
    m_dViews = nullptr;
    if (hipMalloc(&m_dViews, elem * static_cast<size_t>(n) * m_nDet) != hipSuccess) return CTSIM_ERR_HIP;
    m_capacity = n;
  }
  if (prec == CTSIM_PREC_F64) {
    (void)hipMemcpy(m_dViews, m_queue.data(), sizeof(double) * m_queue.size(), hipMemcpyHostToDevice);
  } else {
    std::vector<float> f(m_queue.begin(), m_queue.end());
    (void)hipMemcpy(m_dViews, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice);
  }
  const int rc = backprojectDevice(m_dViews, m_angles.data(), n, prec);
  ctsim_ctx_synchronize(ctsim_amd::context());
  m_queue.clear();
  m_angles.clear();
  return rc;
}

int Backprojector::backprojectDevice(const void* d_filtered, const double* angles, int nView,
                                     int precision) {
  if (m_fail) return CTSIM_ERR_INVALID;
  const int rc = ctsim_backproject_views(ctsim_amd::context(), m_hip, precision, d_filtered, angles,
                                         nView, m_dImage);
  if (rc != CTSIM_OK) warn(ctsim_last_error());
  return rc;
}

int Backprojector::download() {
  ctsim_ctx_synchronize(ctsim_amd::context());
  return hipMemcpy(m_im.data(), m_dImage, sizeof(float) * m_im.nx() * m_im.ny(),
                   hipMemcpyDeviceToHost) == hipSuccess ? CTSIM_OK : CTSIM_ERR_HIP;
}

void Backprojector::PostProcessing() {
  if (m_fail || m_postDone) return;
  flush();
  if (ctsim_backproject_postprocess(ctsim_amd::context(), m_hip, m_dImage) != CTSIM_OK)
    warn(ctsim_last_error());
  download();
  m_postDone = true;
}

// -------------------------------------------------------- Reconstructor
// Reconstructor ctor, reconstruct.cpp:50-103
Reconstructor::Reconstructor(const Projections& rProj, ImageFile& rIF, const char* filterName,
                             double filt_param, const char* filterMethodName, int zeropad,
                             const char* filterGenerationName, const char* interpName,
                             int interpFactor, const char* backprojectName, int iTrace,
                             ReconstructionROI* pROI, bool bRebinToParallel, SGP*)
    : m_rOriginalProj(rProj), m_bRebinToParallel(bRebinToParallel), m_iTrace(iTrace) {
  m_pProj = bRebinToParallel ? rProj.interpolateToParallel() : &rProj;
  if (!m_pProj) {
    m_bFail = true;
    m_strFailMessage = "Error rebinning to parallel geometry";
    return;
  }
  const Projections& proj = *m_pProj;
  const double bw = 1. / proj.detInc();
  m_pProcessSignal = new ProcessSignal(filterName, filterMethodName, bw, proj.detInc(), proj.nDet(),
                                       filt_param, "spatial", filterGenerationName, zeropad,
                                       interpFactor, iTrace, proj.geometry(), proj.focalLength(),
                                       proj.sourceDetectorLength());
  if (m_pProcessSignal->fail()) {
    m_bFail = true;
    m_strFailMessage = "Error creating ProcessSignal: " + m_pProcessSignal->failMessage();
    delete m_pProcessSignal;
    m_pProcessSignal = nullptr;
    return;
  }
  m_pBackprojector = new Backprojector(proj, rIF, backprojectName, interpName, interpFactor, pROI);
  if (m_pBackprojector->fail()) {
    m_bFail = true;
    m_strFailMessage = "Error creating backprojector: " + m_pBackprojector->failMessage();
    delete m_pBackprojector;
    m_pBackprojector = nullptr;
    delete m_pProcessSignal;
    m_pProcessSignal = nullptr;
  }
}

Reconstructor::~Reconstructor() {
  delete m_pBackprojector;
  delete m_pProcessSignal;
  if (m_bRebinToParallel && m_pProj != &m_rOriginalProj) delete m_pProj;
  if (m_dSino) (void)hipFree(m_dSino);
  if (m_dFiltered) (void)hipFree(m_dFiltered);
}

void Reconstructor::reconstructAllViews() {
  reconstructView(0, m_pProj->nView());
  postProcessing();
}

void Reconstructor::postProcessing() {
  if (m_pBackprojector) m_pBackprojector->PostProcessing();
}

// reconstructView, reconstruct.cpp:156-239: filter then backproject every
// view of the range, in order, on the device.
void Reconstructor::reconstructView(int iStartView, int iViewCount, SGP*, bool bBackprojectView,
                                    double) {
  if (m_bFail) return;
  if (iViewCount <= 0) iViewCount = m_pProj->nView() - iStartView;
  if (iViewCount <= 0) return;
  const int nDet = m_pProj->nDet(), nView = m_pProj->nView();
  const int prec = ctsim_amd::precision();
  if (!m_dSino) {
    (void)hipMalloc(reinterpret_cast<void**>(&m_dSino), sizeof(float) * static_cast<size_t>(nView) * nDet);
    (void)hipMalloc(&m_dFiltered, sizeof(double) * static_cast<size_t>(nView) * nDet);
    (void)hipMemcpy(m_dSino, m_pProj->data(), sizeof(float) * static_cast<size_t>(nView) * nDet,
                    hipMemcpyHostToDevice);
  }
  std::vector<double> angles(iViewCount);
  for (int v = 0; v < iViewCount; v++) {
    angles[v] = m_pProj->getDetectorArray(iStartView + v).viewAngle();
    if (m_iTrace == Trace::TRACE_CONSOLE)
      std::cout << "Reconstructing view " << iStartView + v << " (last = " << nView - 1 << ")\n";
  }
  const size_t elem = prec == CTSIM_PREC_F64 ? sizeof(double) : sizeof(float);
  void* out = static_cast<char*>(m_dFiltered) + elem * static_cast<size_t>(iStartView) * nDet;
  if (ctsim_filter_views(ctsim_amd::context(), m_pProcessSignal->handle(), prec,
                         m_dSino + static_cast<size_t>(iStartView) * nDet, iViewCount, out) != CTSIM_OK) {
    warn(ctsim_last_error());
    return;
  }
  if (bBackprojectView) m_pBackprojector->backprojectDevice(out, angles.data(), iViewCount, prec);
}
