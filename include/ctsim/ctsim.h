// ctsim/ctsim.h -- C++ drop-in for the CTSim hot-path classes, backed by the
// MI355X kernels of libctsim_amd.so through the C ABI of ctsim_hip.h.
//
// The classes keep the reference's names, constructor arguments, accessors,
// static name/ID tables and fail()/failMessage() convention, so the calls
// tools written against CTSim make (tools/pjrec.cpp, tools/phm2pj.cpp, MPI
// and GUI code aside) compile against this header:
//   Timer          include/timer.h:40-99
//   Array2dFileLabel include/array2dfile.h:45-100
//   SignalFilter   include/filter.h:47-165 (the static name/ID API; the
//                  filter itself is generated inside ProcessSignal)
//   Phantom        include/phantom.h:133-228
//   DetectorArray  include/scanner.h:41-62
//   Scanner        include/scanner.h:64-156
//   Projections    include/projections.h:103-219 (.pj files, Appendix B)
//   ImageFile      include/imagefile.h:170-241 (float32 images, .if files)
//   ProcessSignal  include/procsignal.h:37-160
//   Backprojector  include/backprojectors.h:40-106
//   Reconstructor  include/reconstruct.h:43-78
// GUI-only arguments (SGP*) are accepted and ignored.  Compute runs on the
// process's device (ctsim_amd::setDevice, default 0) in the precision chosen
// by ctsim_amd::setPrecision or the CTSIM_PRECISION environment variable
// ("f64", the default, reproduces the reference; "mixed" is the fast path).
#ifndef CTSIM_CTSIM_H
#define CTSIM_CTSIM_H

#include <chrono>
#include <cstring>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

#include "../ctsim_hip.h"

class SGP;  // the reference's graphics port; unused

namespace ctsim_amd {
void setDevice(int device);
void setPrecision(int precision);  // CTSIM_PREC_F64 or CTSIM_PREC_MIXED
int precision();
ctsim_ctx* context();              // the process-wide context (created on first use)
}  // namespace ctsim_amd

// Trace levels (include/trace.h); only TRACE_NONE / TRACE_CONSOLE have effect.
struct Trace {
  enum { TRACE_INVALID = -1, TRACE_NONE = 0, TRACE_CONSOLE = 1, TRACE_PHANTOM = 2,
         TRACE_PROJECTIONS = 3, TRACE_PLOT = 4, TRACE_CLIPPING = 5 };
  static int convertTraceNameToID(const char* name);
};

// File name without its directories (include/ctsupport.h:141-146).
inline const char* fileBasename(const char* filename) {
  const char* p = std::strrchr(filename, '/');
  return p ? p + 1 : filename;
}

// Wall-clock seconds since construction (include/timer.h:40-99).
class Timer {
 public:
  Timer() : m_timeStart(now()) {}
  virtual ~Timer() {}
  virtual double timerEnd() {
    m_timeEnd = now();
    m_timeElapsed = m_timeEnd - m_timeStart;
    return m_timeElapsed;
  }
  virtual void timerReport(const char* msg) const;
  virtual double timerEndAndReport(const char* msg) {
    const double t = timerEnd();
    timerReport(msg);
    return t;
  }
  double getTimeElapsed() const { return m_timeElapsed; }

 protected:
  static double now() {
    using namespace std::chrono;
    return duration<double>(system_clock::now().time_since_epoch()).count();
  }
  double m_timeStart;
  double m_timeEnd = 0;
  double m_timeElapsed = 0;
};

// An image or projection label (include/array2dfile.h:45-100): a type, a
// string, a calculation time and the local date and time of its creation.
class Array2dFileLabel {
 public:
  enum { L_EMPTY = 0, L_HISTORY = 1, L_USER = 2 };
  Array2dFileLabel();
  Array2dFileLabel(const char* str, double ctime = 0.);
  Array2dFileLabel(int type, const char* str, double ctime = 0.);
  const std::string& getLabelString() const { return m_strLabel; }
  double getCalcTime() const { return m_calcTime; }
  void setCalcTime(double calcTime) { m_calcTime = calcTime; }
  void setLabelType(int labelType) { m_labelType = labelType; }
  int getLabelType() const { return m_labelType; }
  std::string& setLabelString(const char* str) { return m_strLabel = str; }
  std::string& setLabelString(const std::string& str) { return m_strLabel = str; }
  void setDateTime(int year, int month, int day, int hour, int minute, int second);
  void getDateTime(int& year, int& month, int& day, int& hour, int& minute, int& second) const;
  const std::string& getDateString() const;
  void print(std::ostream& os) const;
  void printBrief(std::ostream& os) const;

 private:
  int m_labelType = L_EMPTY;
  std::string m_strLabel;
  double m_calcTime = 0;
  int m_year = 0, m_month = 0, m_day = 0, m_hour = 0, m_minute = 0, m_second = 0;
  mutable std::string m_strDate;
};

// The static filter name/ID API of SignalFilter (include/filter.h:49-70,
// 112-135; libctsim/filter.cpp:30-75, 231-300), including the reference's
// swapped names at IDs 4 and 5 ("shepp" names FILTER_ABS_SINC).
class SignalFilter {
 public:
  static constexpr int FILTER_INVALID = -1, FILTER_ABS_BANDLIMIT = 0, FILTER_ABS_G_HAMMING = 1,
                       FILTER_ABS_HANNING = 2, FILTER_ABS_COSINE = 3, FILTER_ABS_SINC = 4,
                       FILTER_SHEPP = 5, FILTER_BANDLIMIT = 6, FILTER_SINC = 7,
                       FILTER_G_HAMMING = 8, FILTER_HANNING = 9, FILTER_COSINE = 10,
                       FILTER_TRIANGLE = 11;
  static constexpr int DOMAIN_INVALID = -1, DOMAIN_FREQUENCY = 0, DOMAIN_SPATIAL = 1;
  static int getFilterCount();
  static int getReconstructFilterCount();
  static const char* const* getFilterNameArray();
  static const char* const* getFilterTitleArray();
  static int convertFilterNameToID(const char* filterName);
  static const char* convertFilterIDToName(int filterID);
  static const char* convertFilterIDToTitle(int filterID);
  static int getDomainCount();
  static const char* const* getDomainNameArray();
  static const char* const* getDomainTitleArray();
  static int convertDomainNameToID(const char* domainName);
  static const char* convertDomainIDToName(int domainID);
  static const char* convertDomainIDToTitle(int domainID);
};

struct ReconstructionROI {  // include/reconstruct.h:36-41
  double m_dXMin;
  double m_dYMin;
  double m_dXMax;
  double m_dYMax;
};

class Phantom {
 public:
  static constexpr int PHM_INVALID = -1, PHM_HERMAN = 0, PHM_SHEPP_LOGAN = 1, PHM_UNITPULSE = 2;
  Phantom();
  explicit Phantom(const char* phmName);
  ~Phantom();
  Phantom(const Phantom&) = delete;
  Phantom& operator=(const Phantom&) = delete;
  bool createFromPhantom(const char* phmName);
  bool createFromPhantom(int phmid);
  // phantom.cpp:184-213: one "type cx cy u v rot atten" element per record;
  // a record with missing fields warns, marks the file bad and is still
  // added, and reading goes on (fields the record lacks keep the previous
  // record's values, zero before the first).
  bool createFromFile(const char* fname);
  bool fileWrite(const char* fname);
  void addPElem(const char* type, double cx, double cy, double u, double v, double rot,
                double atten);
  bool fail() const { return m_fail; }
  const std::string& failMessage() const { return m_failMessage; }
  const std::string& name() const { return m_name; }
  int id() const { return m_id; }
  int nPElem() const { return m_nPElem; }
  // phantom.cpp:272-322: element count, extent, and per element its outline
  // point count, attenuation, rotation (degrees) and bounding box; the
  // definitions in the phantom file format.
  void print(std::ostream& os) const;
  void printDefinitions(std::ostream& os) const;
  static int getPhantomCount();
  static const char* const* getPhantomNameArray();
  static const char* const* getPhantomTitleArray();
  static int convertNameToPhantomID(const char* phmName);
  static const char* convertPhantomIDToName(int phmID);
  static const char* convertPhantomIDToTitle(int phmID);
  double xmin() const;
  double xmax() const;
  double ymin() const;
  double ymax() const;
  double maxAxisLength() const;
  double getDiameterBoundaryCircle() const;
  ctsim_phantom* handle() const { return m_hip; }
  // Rasterise on the device (phantom.cpp:448-538, phm2if).
  void convertToImagefile(class ImageFile& im, double dViewRatio, int nsample,
                          int trace = Trace::TRACE_NONE) const;

 private:
  ctsim_phantom* m_hip = nullptr;
  bool m_fail = false;
  std::string m_failMessage, m_name;
  int m_nPElem = 0;
  int m_id = PHM_INVALID;
};

class DetectorArray {
 public:
  int nDet() const { return m_nDet; }
  double viewAngle() const { return m_viewAngle; }
  float* detValues() { return m_detValues; }
  const float* detValues() const { return m_detValues; }
  void setViewAngle(double a) { m_viewAngle = a; }

 private:
  friend class Projections;
  float* m_detValues = nullptr;
  int m_nDet = 0;
  double m_viewAngle = 0;
};

class Scanner;

class Projections {
 public:
  Projections();
  Projections(int nView, int nDet);
  explicit Projections(const Scanner& scanner);
  void initFromScanner(const Scanner& scanner);
  // projections.cpp:935-949: the scan header as text.
  void printScanInfo(std::ostringstream& os) const;
  // projections.cpp:164-168: reallocates nView zeroed views (the MPI tools
  // restrict a rank's views this way).
  void setNView(int nView);
  bool read(const char* fname);
  bool read(const std::string& fname) { return read(fname.c_str()); }
  bool write(const char* fname);
  bool write(const std::string& fname) { return write(fname.c_str()); }
  DetectorArray& getDetectorArray(int iview) { return m_views[iview]; }
  const DetectorArray& getDetectorArray(int iview) const { return m_views[iview]; }
  int nDet() const { return m_nDet; }
  int nView() const { return m_nView; }
  int geometry() const { return m_geometry; }
  double detInc() const { return m_detInc; }
  double rotInc() const { return m_rotInc; }
  double detStart() const { return m_detStart; }
  double rotStart() const { return m_rotStart; }
  double focalLength() const { return m_dFocalLength; }
  double sourceDetectorLength() const { return m_dSourceDetectorLength; }
  double viewDiameter() const { return m_dViewDiameter; }
  double phmLen() const;
  double fanBeamAngle() const { return m_dFanBeamAngle; }
  double calcTime() const { return m_calcTime; }
  const char* remark() const { return m_remark.c_str(); }
  double viewLen() const;
  const std::string& getFilename() const { return m_filename; }
  // The history label a read file or setRemark gives (projections.cpp:
  // 128-161, 628-634); pjrec copies it into the image's labels.
  Array2dFileLabel& getLabel() { return m_label; }
  const Array2dFileLabel& getLabel() const { return m_label; }
  void setCalcTime(double t) { m_calcTime = t; }
  void setRemark(const char* r) { m_remark = r; m_label.setLabelString(r); }
  void setRemark(const std::string& r) { setRemark(r.c_str()); }
  void setRotInc(double r) { m_rotInc = r; }
  void setDetInc(double d) { m_detInc = d; }
  void setPhmLen(double p);
  // Contiguous [nView][nDet] float data (the DetectorArrays point into it).
  const float* data() const { return m_data.data(); }
  float* data() { return m_data.data(); }
  // On the device (projections.cpp:200-439, 1250-1328).  Helical180LI and
  // HalfScanFeather return 0 on success like the reference; the equilinear
  // geometry is accepted (the reference supports equiangular only).
  int Helical180LI(int interpView);
  int HalfScanFeather();
  // A new parallel-beam Projections (the caller deletes it), or this object
  // itself for parallel data, as the reference.
  Projections* interpolateToParallel() const;

 private:
  void init(int nView, int nDet);
  int m_nView = 0, m_nDet = 0, m_geometry = 0;
  double m_detInc = 0, m_rotInc = 0, m_detStart = 0, m_rotStart = 0, m_dFocalLength = 0;
  double m_dSourceDetectorLength = 0, m_dViewDiameter = 0, m_dFanBeamAngle = 0, m_calcTime = 0;
  int m_year = 0, m_month = 0, m_day = 0, m_hour = 0, m_minute = 0, m_second = 0;
  std::string m_remark, m_filename;
  Array2dFileLabel m_label;
  std::vector<float> m_data;
  std::vector<DetectorArray> m_views;
};

class Scanner {
 public:
  static const int GEOMETRY_INVALID = -1, GEOMETRY_PARALLEL = 0, GEOMETRY_EQUIANGULAR = 1,
                   GEOMETRY_EQUILINEAR = 2, GEOMETRY_LINOGRAM = 3;
  Scanner(const Phantom& phm, const char* geometryName, int nDet, int nView, int iOffsetView,
          int nSample, double rot_anglen, double dFocalLengthRatio, double dCenterDetectorRatio,
          double dViewRatio, double dScanRatio);
  ~Scanner();
  Scanner(const Scanner&) = delete;
  Scanner& operator=(const Scanner&) = delete;
  void collectProjections(Projections& proj, const Phantom& phm, int trace = Trace::TRACE_NONE,
                          SGP* pSGP = nullptr);
  void collectProjections(Projections& proj, const Phantom& phm, int iStartView, int iNumViews,
                          int iOffsetView, bool bStoreAtViewPosition, int trace = Trace::TRACE_NONE,
                          SGP* pSGP = nullptr);
  void collectProjections(Projections& proj, const Phantom& phm, int iStartView, int iNumViews,
                          int iOffsetView, int iStorageOffset, int trace = Trace::TRACE_NONE,
                          SGP* pSGP = nullptr);
  // Declared by the reference (scanner.h:87-88) and never defined there, so
  // it has no behaviour to follow: here they set the view count and view
  // offset later collectProjections calls use (rotInc is kept).
  void setNView(int nView);
  void setOffsetView(int iOffsetView);
  bool fail() const { return m_fail; }
  const std::string& failMessage() const { return m_failMessage; }
  unsigned int nDet() const { return m_nDet; }
  unsigned int nView() const { return m_nView; }
  unsigned int offsetView() const { return m_offsetView; }
  unsigned int startView() const { return 0; }
  double rotInc() const { return m_info[2]; }
  double detInc() const { return m_info[0]; }
  double detLen() const { return m_info[8]; }
  double detStart() const { return m_info[1]; }
  double focalLength() const { return m_info[4]; }
  double sourceDetectorLength() const { return m_info[5]; }
  double centerDetectorLength() const { return m_info[9]; }
  double viewDiameter() const { return m_info[3]; }
  double scanDiameter() const { return m_info[10]; }
  double fanBeamAngle() const { return m_info[6]; }
  int geometry() const { return m_geometry; }
  static int getGeometryCount();
  static const char* const* getGeometryNameArray();
  static const char* const* getGeometryTitleArray();
  static int convertGeometryNameToID(const char* name);
  static const char* convertGeometryIDToName(int id);
  static const char* convertGeometryIDToTitle(int id);

 private:
  ctsim_scanner* m_hip = nullptr;
  bool m_fail = false;
  std::string m_failMessage;
  int m_nDet = 0, m_nView = 0, m_offsetView = 0, m_geometry = -1;
  double m_info[12] = {};
  float* m_dSino = nullptr;  // device rows, kept across collectProjections calls
  size_t m_dRows = 0;
};

class ImageFile {
 public:
  ImageFile();
  ImageFile(int nx, int ny);
  bool setArraySize(int nx, int ny);
  unsigned int nx() const { return m_nx; }
  unsigned int ny() const { return m_ny; }
  float** getArray() { return m_cols.empty() ? nullptr : m_cols.data(); }
  const float* const* getArray() const { return m_cols.empty() ? nullptr : m_cols.data(); }
  float* data() { return m_data.data(); }          // column-major [nx][ny]
  const float* data() const { return m_data.data(); }
  void arrayDataClear();
  void setAxisIncrement(double incX, double incY);
  void setAxisExtent(double minX, double maxX, double minY, double maxY);
  // array2dfile.cpp:636-657: a history label, a label of the given type, or
  // a copy of a label.
  void labelAdd(const char* label, double calcTime = 0.);
  void labelAdd(int type, const char* label, double calcTime = 0.);
  void labelAdd(const Array2dFileLabel& label);
  unsigned int getNumLabels() const { return static_cast<unsigned int>(m_labels.size()); }
  const Array2dFileLabel& labelGet(int i) const { return m_labels[i]; }
  bool fileWrite(const char* fname);
  bool fileWrite(const std::string& fname) { return fileWrite(fname.c_str()); }
  bool fileRead(const char* fname);
  bool fileRead(const std::string& fname) { return fileRead(fname.c_str()); }
  // Herman's d, r, e with this image as the reference (imagefile.cpp:185-240).
  bool comparativeStatistics(const ImageFile& imComp, double& d, double& r, double& e) const;

 private:
  unsigned int m_nx = 0, m_ny = 0;
  std::vector<float> m_data;
  std::vector<float*> m_cols;
  bool m_axisIncKnown = false, m_axisExtentKnown = false;
  double m_incX = 0, m_incY = 0, m_minX = 0, m_maxX = 0, m_minY = 0, m_maxY = 0;
  std::vector<Array2dFileLabel> m_labels;
};

class ProcessSignal {
 public:
  // procsignal.cpp:31-73, with the FFTW methods this build implements.
  static constexpr int FILTER_METHOD_INVALID = -1, FILTER_METHOD_CONVOLUTION = 0,
                       FILTER_METHOD_FOURIER = 1, FILTER_METHOD_FOURIER_TABLE = 2,
                       FILTER_METHOD_FFT = 3, FILTER_METHOD_FFTW = 4, FILTER_METHOD_RFFTW = 5;
  static constexpr int FILTER_GENERATION_INVALID = -1, FILTER_GENERATION_DIRECT = 0,
                       FILTER_GENERATION_INVERSE_FOURIER = 1;
  enum { FORWARD = -1, BACKWARD = 1 };
  static int getFilterMethodCount();
  static const char* const* getFilterMethodNameArray();
  static const char* const* getFilterMethodTitleArray();
  static int convertFilterMethodNameToID(const char* filterMethodName);
  static const char* convertFilterMethodIDToName(int idFilterMethod);
  static const char* convertFilterMethodIDToTitle(int idFilterMethod);
  static int getFilterGenerationCount();
  static const char* const* getFilterGenerationNameArray();
  static const char* const* getFilterGenerationTitleArray();
  static int convertFilterGenerationNameToID(const char* fgName);
  static const char* convertFilterGenerationIDToName(int idFG);
  static const char* convertFilterGenerationIDToTitle(int idFG);
  ProcessSignal(const char* szFilterName, const char* szFilterMethodName, double dBandwidth,
                double dSignalIncrement, int nSignalPoints, double dFilterParam,
                const char* szDomainName, const char* szFilterGenerationName, int iZeropad,
                int iPreinterpolationFactor, int iTraceLevel, int iGeometry, double dFocalLength,
                double dSourceDetectorLength, SGP* pSGP = nullptr);
  ~ProcessSignal();
  ProcessSignal(const ProcessSignal&) = delete;
  ProcessSignal& operator=(const ProcessSignal&) = delete;
  // One view, FP64 (procsignal.cpp:544-629).
  void filterSignal(const float input[], double output[]) const;
  bool fail() const { return m_fail; }
  const std::string& failMessage() const { return m_failMessage; }
  int getNFilterPoints() const { return m_nFilterPoints; }
  double getFilterMin() const { return m_minMaxInc[0]; }
  double getFilterMax() const { return m_minMaxInc[1]; }
  double getFilterIncrement() const { return m_minMaxInc[2]; }
  const double* getFilter() const { return m_filter.data(); }
  ctsim_procsignal* handle() const { return m_hip; }
  int nSignalPoints() const { return m_n; }
  void setTraceLevel(int traceLevel) { m_traceLevel = traceLevel; }
  int idFilterGeneration() const { return m_idFilterGeneration; }

 private:
  ctsim_procsignal* m_hip = nullptr;
  int m_traceLevel = 0, m_idFilterGeneration = FILTER_GENERATION_INVALID;
  bool m_fail = false;
  std::string m_failMessage;
  int m_n = 0, m_nFilterPoints = 0;
  double m_minMaxInc[3] = {};
  std::vector<double> m_filter;
  mutable void* m_dIn = nullptr;
  mutable void* m_dOut = nullptr;
};

class Backprojector {
 public:
  // backprojectors.cpp:29-97 (frequency preinterpolation and B-splines are
  // compiled out of the reference, so the name table has three entries).
  static constexpr int BPROJ_INVALID = -1, BPROJ_TRIG = 0, BPROJ_TABLE = 1, BPROJ_DIFF = 2,
                       BPROJ_IDIFF = 3;
  static constexpr int INTERP_INVALID = -1, INTERP_NEAREST = 0, INTERP_LINEAR = 1,
                       INTERP_CUBIC = 2, INTERP_FREQ_PREINTERPOLATION = 3;
  static int getBackprojectCount();
  static const char* const* getBackprojectNameArray();
  static const char* const* getBackprojectTitleArray();
  static const char* convertBackprojectIDToName(int bprojID);
  static const char* convertBackprojectIDToTitle(int bprojID);
  static int getInterpCount();
  static const char* const* getInterpNameArray();
  static const char* const* getInterpTitleArray();
  static const char* convertInterpIDToName(int interpID);
  static const char* convertInterpIDToTitle(int interpID);
  Backprojector(const Projections& proj, ImageFile& im, const char* backprojName,
                const char* interpName, int interpFactor, const ReconstructionROI* pROI = nullptr);
  ~Backprojector();
  Backprojector(const Backprojector&) = delete;
  Backprojector& operator=(const Backprojector&) = delete;
  // Views are queued and backprojected in order on the device; the host image
  // is updated by PostProcessing() (or the destructor, unscaled).
  void BackprojectView(const double* viewData, double viewAngle);
  void PostProcessing();
  bool fail() const { return m_fail; }
  const std::string& failMessage() const { return m_failMessage; }
  // Device-side batch: nView filtered rows already on the device.
  int backprojectDevice(const void* d_filtered, const double* angles, int nView, int precision);
  static int convertBackprojectNameToID(const char* name);
  static int convertInterpNameToID(const char* name);

 private:
  int flush();
  int download();
  ctsim_backprojector* m_hip = nullptr;
  ImageFile& m_im;
  bool m_fail = false, m_postDone = false;
  std::string m_failMessage;
  int m_nDet = 0;
  float* m_dImage = nullptr;
  void* m_dViews = nullptr;
  int m_capacity = 0;
  std::vector<double> m_queue, m_angles;
};

class Reconstructor {
 public:
  Reconstructor(const Projections& rProj, ImageFile& rIF, const char* filterName,
                double filt_param, const char* filterMethodName, int zeropad,
                const char* filterGenerationName, const char* interpName, int interpFactor,
                const char* backprojectName, int iTrace, ReconstructionROI* pROI = nullptr,
                bool bRebinToParallel = false, SGP* pSGP = nullptr);
  ~Reconstructor();
  Reconstructor(const Reconstructor&) = delete;
  Reconstructor& operator=(const Reconstructor&) = delete;
  bool fail() const { return m_bFail; }
  const std::string& failMessage() const { return m_strFailMessage; }
  void reconstructAllViews();
  void reconstructView(int iStartView = 0, int iViewCount = -1, SGP* pSGP = nullptr,
                       bool bBackprojectView = true, double dGraphWidth = 1.);
  void postProcessing();

 private:
  const Projections& m_rOriginalProj;
  const Projections* m_pProj = nullptr;  // the original, or its parallel rebinning
  bool m_bRebinToParallel = false;
  ProcessSignal* m_pProcessSignal = nullptr;
  Backprojector* m_pBackprojector = nullptr;
  int m_iTrace = 0;
  bool m_bFail = false;
  std::string m_strFailMessage;
  float* m_dSino = nullptr;
  void* m_dFiltered = nullptr;
};

#endif  // CTSIM_CTSIM_H
