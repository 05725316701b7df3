// Thin extern "C" harness over the real CTSim library, compiled from the
// reference sources by oracle/Makefile into oracle/_ref/libctsimref.so.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (ctsim_amd/, include/)
// may load this library; it exists so tests and the CPU-baseline leg of
// bench.py can run the reference's own code on the same inputs.
//
// Every entry point mirrors a driver of the reference:
//   ref_pipeline      ~ tools/phm2pj.cpp:333-371 + tools/pjrec.cpp:342-347
//                       (filtering and backprojection are driven view by view
//                        exactly like Reconstructor::reconstructView,
//                        libctsim/reconstruct.cpp:156-239, so the filtered
//                        views can be captured)
//   ref_filter        ~ ProcessSignal ctor + filterSignal (procsignal.cpp:82-629)
//   ref_backproject   ~ Backprojector ctor + BackprojectView + PostProcessing
//                       (backprojectors.cpp:100-336) on caller-supplied views
//   ref_stats         ~ ImageFile::comparativeStatistics (imagefile.cpp:185-240)
#include "ct.h"
#include <chrono>
#include <thread>
#include <vector>

static std::string g_err;

static const char* elemTypeName(int t) {
  switch (t) {
    case 1: return "rectangle";
    case 2: return "triangle";
    case 3: return "ellipse";
    case 4: return "sector";
    case 5: return "segment";
  }
  return "invalid";
}

static void buildPhantom(Phantom& phm, const int* types, const double* params, int nElem,
                         const char* builtin) {
  if (builtin && builtin[0]) phm.createFromPhantom(builtin);
  for (int i = 0; i < nElem; i++) {
    const double* p = params + 6 * i;
    phm.addPElem(elemTypeName(types[i]), p[0], p[1], p[2], p[3], p[4], p[5]);
  }
}

static void copyImage(const ImageFile& im, float* out) {
  ImageFileArrayConst v = im.getArray();
  for (unsigned ix = 0; ix < im.nx(); ix++)
    for (unsigned iy = 0; iy < im.ny(); iy++) out[(size_t)ix * im.ny() + iy] = v[ix][iy];
}

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

// info[0..15]: detInc, detStart, rotInc, viewDiameter, focalLength,
// sourceDetectorLength, fanBeamAngle, phmLen, phm xmin, xmax, ymin, ymax,
// detLen, centerDetectorLength, scanDiameter, nFilterPoints
int ref_pipeline(const int* types, const double* params, int nElem, const char* builtin,
                 const char* geometry, int nDet, int nView, int offsetView, int nSample,
                 double rotLen, double focalRatio, double centerRatio, double viewRatio,
                 double scanRatio, const char* filterName, double filterParam,
                 const char* methodName, int zeropad, const char* generationName,
                 const char* interpName, int interpFactor, const char* backprojName, int nx,
                 int ny, float* sino, double* angles, double* filtered, float* image,
                 float* imageRecon, double* info, double* timings) {
  Phantom phm;
  buildPhantom(phm, types, params, nElem, builtin);
  if (phm.fail()) { g_err = phm.failMessage(); return 1; }
  Scanner scanner(phm, geometry, nDet, nView, offsetView, nSample, rotLen, focalRatio,
                  centerRatio, viewRatio, scanRatio);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  Projections pj(scanner);
  auto t0 = std::chrono::steady_clock::now();
  scanner.collectProjections(pj, phm, 0, nView, offsetView, true, Trace::TRACE_NONE);
  auto t1 = std::chrono::steady_clock::now();
  for (int v = 0; v < nView; v++) {
    const DetectorArray& da = pj.getDetectorArray(v);
    for (int d = 0; d < nDet; d++) sino[(size_t)v * nDet + d] = da.detValues()[d];
    angles[v] = da.viewAngle();
  }
  if (info) {
    info[0] = pj.detInc(); info[1] = pj.detStart(); info[2] = pj.rotInc();
    info[3] = pj.viewDiameter(); info[4] = pj.focalLength(); info[5] = pj.sourceDetectorLength();
    info[6] = scanner.fanBeamAngle(); info[7] = pj.phmLen(); info[8] = phm.xmin();
    info[9] = phm.xmax(); info[10] = phm.ymin(); info[11] = phm.ymax();
    info[12] = scanner.detLen(); info[13] = scanner.centerDetectorLength();
    info[14] = scanner.scanDiameter(); info[15] = 0;
  }
  if (!filterName || nx <= 0) return 0;

  // Same construction as Reconstructor::Reconstructor (reconstruct.cpp:70-85).
  ImageFile im(nx, ny);
  double bw = 1. / pj.detInc();
  ProcessSignal ps(filterName, methodName, bw, pj.detInc(), pj.nDet(), filterParam, "spatial",
                   generationName, zeropad, interpFactor, Trace::TRACE_NONE, pj.geometry(),
                   pj.focalLength(), pj.sourceDetectorLength());
  if (ps.fail()) { g_err = ps.failMessage(); return 3; }
  if (info) info[15] = ps.getNFilterPoints();
  Backprojector bp(pj, im, backprojName, interpName, interpFactor, NULL);
  if (bp.fail()) { g_err = bp.failMessage(); return 4; }
  int nFilt = pj.nDet() * interpFactor;
  double* buf = new double[nFilt];
  double tf = 0, tb = 0;
  for (int v = 0; v < nView; v++) {
    const DetectorArray& da = pj.getDetectorArray(v);
    auto a = std::chrono::steady_clock::now();
    ps.filterSignal(da.detValues(), buf);
    auto b = std::chrono::steady_clock::now();
    if (filtered) memcpy(filtered + (size_t)v * nFilt, buf, sizeof(double) * nFilt);
    bp.BackprojectView(buf, da.viewAngle());
    auto c = std::chrono::steady_clock::now();
    tf += std::chrono::duration<double>(b - a).count();
    tb += std::chrono::duration<double>(c - b).count();
  }
  bp.PostProcessing();
  delete[] buf;
  copyImage(im, image);
  if (timings) {
    timings[0] = std::chrono::duration<double>(t1 - t0).count();
    timings[1] = tf;
    timings[2] = tb;
  }
  if (imageRecon) {
    ImageFile im2(nx, ny);
    Reconstructor rec(pj, im2, filterName, filterParam, methodName, zeropad, generationName,
                      interpName, interpFactor, backprojName, Trace::TRACE_NONE);
    if (rec.fail()) { g_err = rec.failMessage(); return 5; }
    rec.reconstructAllViews();
    copyImage(im2, imageRecon);
  }
  return 0;
}

int ref_filter(const char* filterName, const char* methodName, double bw, double inc, int n,
               double param, const char* domainName, const char* generationName, int zeropad,
               int preinterp, int geometry, double focalLength, double sdl, const float* in,
               int nView, double* out, double* filterOut, int filterCap, int* nFilterPoints,
               double* filterMinMaxInc) {
  ProcessSignal ps(filterName, methodName, bw, inc, n, param, domainName, generationName, zeropad,
                   preinterp, Trace::TRACE_NONE, geometry, focalLength, sdl);
  if (ps.fail()) { g_err = ps.failMessage(); return 1; }
  int nf = ps.getNFilterPoints();
  if (nFilterPoints) *nFilterPoints = nf;
  if (filterMinMaxInc) {
    filterMinMaxInc[0] = ps.getFilterMin();
    filterMinMaxInc[1] = ps.getFilterMax();
    filterMinMaxInc[2] = ps.getFilterIncrement();
  }
  if (filterOut) {
    const double* f = ps.getFilter();
    for (int i = 0; i < nf && i < filterCap; i++) filterOut[i] = f[i];
  }
  for (int v = 0; v < nView; v++) ps.filterSignal(in + (size_t)v * n, out + (size_t)v * n);
  return 0;
}

int ref_backproject(const int* types, const double* params, int nElem, const char* builtin,
                    const char* geometry, int nDet, int nView, int offsetView, int nSample,
                    double rotLen, double focalRatio, double centerRatio, double viewRatio,
                    double scanRatio, const char* backprojName, const char* interpName,
                    int interpFactor, int nx, int ny, const double* roi, const double* filtered,
                    const double* angles, int postProcess, float* image) {
  Phantom phm;
  buildPhantom(phm, types, params, nElem, builtin);
  Scanner scanner(phm, geometry, nDet, nView, offsetView, nSample, rotLen, focalRatio,
                  centerRatio, viewRatio, scanRatio);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  Projections pj(scanner);
  ImageFile im(nx, ny);
  ReconstructionROI r;
  if (roi) { r.m_dXMin = roi[0]; r.m_dXMax = roi[1]; r.m_dYMin = roi[2]; r.m_dYMax = roi[3]; }
  Backprojector bp(pj, im, backprojName, interpName, interpFactor, roi ? &r : NULL);
  if (bp.fail()) { g_err = bp.failMessage(); return 4; }
  int nFilt = nDet * interpFactor;
  for (int v = 0; v < nView; v++) bp.BackprojectView(filtered + (size_t)v * nFilt, angles[v]);
  if (postProcess) bp.PostProcessing();
  copyImage(im, image);
  return 0;
}

int ref_stats(int nx, int ny, const float* refImage, const float* cmpImage, double* dre) {
  ImageFile a(nx, ny), b(nx, ny);
  ImageFileArray va = a.getArray(), vb = b.getArray();
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      va[ix][iy] = refImage[(size_t)ix * ny + iy];
      vb[ix][iy] = cmpImage[(size_t)ix * ny + iy];
    }
  return a.comparativeStatistics(b, dre[0], dre[1], dre[2]) ? 0 : 1;
}

}  // extern "C"

extern "C" {

// Filter + backproject the first nUse views of a given sinogram, as
// Reconstructor::reconstructView does (reconstruct.cpp:156-239), timing the
// two stages.  The geometry comes from a Scanner built on the same phantom.
int ref_fbp_sample(const int* types, const double* params, int nElem, const char* builtin,
                   const char* geometry, int nDet, int nView, int nSample, double rotLen,
                   double focalRatio, double centerRatio, const char* filterName,
                   const char* methodName, int zeropad, const char* interpName,
                   const char* backprojName, int nx, int ny, const float* sino,
                   const double* angles, int nUse, float* image, double* timings) {
  Phantom phm;
  buildPhantom(phm, types, params, nElem, builtin);
  Scanner scanner(phm, geometry, nDet, nView, 0, nSample, rotLen, focalRatio, centerRatio, 1.0, 1.0);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  Projections pj(scanner);
  ImageFile im(nx, ny);
  ProcessSignal ps(filterName, methodName, 1. / pj.detInc(), pj.detInc(), pj.nDet(), -1.0,
                   "spatial", "direct", zeropad, 1, Trace::TRACE_NONE, pj.geometry(),
                   pj.focalLength(), pj.sourceDetectorLength());
  if (ps.fail()) { g_err = ps.failMessage(); return 3; }
  Backprojector bp(pj, im, backprojName, interpName, 1, NULL);
  if (bp.fail()) { g_err = bp.failMessage(); return 4; }
  double* buf = new double[nDet];
  double tf = 0, tb = 0;
  for (int v = 0; v < nUse; v++) {
    auto a = std::chrono::steady_clock::now();
    ps.filterSignal(sino + (size_t)v * nDet, buf);
    auto b = std::chrono::steady_clock::now();
    bp.BackprojectView(buf, angles[v]);
    auto c = std::chrono::steady_clock::now();
    tf += std::chrono::duration<double>(b - a).count();
    tb += std::chrono::duration<double>(c - b).count();
  }
  bp.PostProcessing();
  delete[] buf;
  copyImage(im, image);
  timings[0] = tf;
  timings[1] = tb;
  return 0;
}

// All host cores as the reference's GUI reconstructs
// (src/backgroundsupr.cpp:111-152, src/threadrecon.cpp:175-260): nThreads
// workers over contiguous blocks of views [0, nUse) (nUse/nThreads each,
// the first nUse%nThreads one more), each with its own Reconstructor over
// the whole Projections writing its own image (reconstructView one view at a
// time, then postProcessing), the images then summed in float.  seconds[0]
// is the wall time of the workers and the sum.
int ref_fbp_threads(const int* types, const double* params, int nElem, const char* builtin,
                    const char* geometry, int nDet, int nView, int nSample, double rotLen,
                    double focalRatio, double centerRatio, const char* filterName,
                    const char* methodName, int zeropad, const char* interpName,
                    const char* backprojName, int nx, int ny, const float* sino,
                    const double* angles, int nUse, int nThreads, float* image,
                    double* seconds) {
  Phantom phm;
  buildPhantom(phm, types, params, nElem, builtin);
  Scanner scanner(phm, geometry, nDet, nView, 0, nSample, rotLen, focalRatio, centerRatio, 1.0, 1.0);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  if (nThreads < 1 || nUse < 1 || nUse > nView) { g_err = "bad thread or view count"; return 3; }
  Projections pj(scanner);
  for (int v = 0; v < nView; v++) {
    DetectorArray& d = pj.getDetectorArray(v);
    d.setViewAngle(angles[v]);
    for (int i = 0; i < nDet; i++) d.detValues()[i] = sino[(size_t)v * nDet + i];
  }
  std::vector<ImageFile*> child(nThreads);
  for (int t = 0; t < nThreads; t++) child[t] = new ImageFile(nx, ny);
  std::vector<int> fails(nThreads, 0);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> workers;
  const int base = nUse / nThreads, extra = nUse % nThreads;
  int start = 0;
  for (int t = 0; t < nThreads; t++) {
    const int count = base + (t < extra ? 1 : 0);
    workers.emplace_back([&, t, start, count]() {
      Reconstructor rec(pj, *child[t], filterName, -1.0, methodName, zeropad, "direct", interpName,
                        1, backprojName, Trace::TRACE_NONE, NULL, false);
      if (rec.fail()) { fails[t] = 1; return; }
      for (int u = 0; u < count; u++) rec.reconstructView(start + u, 1);
      rec.postProcessing();
    });
    start += count;
  }
  for (std::thread& w : workers) w.join();
  ImageFile sum(nx, ny);
  sum.arrayDataClear();
  ImageFileArray a = sum.getArray();
  for (int t = 0; t < nThreads; t++) {
    ImageFileArrayConst c = child[t]->getArray();
    for (int ix = 0; ix < nx; ix++)
      for (int iy = 0; iy < ny; iy++) a[ix][iy] += c[ix][iy];
  }
  seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int t = 0; t < nThreads; t++) delete child[t];
  for (int f : fails)
    if (f) { g_err = "Reconstructor failed"; return 4; }
  copyImage(sum, image);
  return 0;
}

}  // extern "C"

extern "C" {

// File formats through the reference's own writers and readers
// (Projections::write/read, projections.cpp:486-546 and 817-889;
// Array2dFile::fileWrite/fileRead, array2dfile.cpp:443-633), so tests can
// check the drop-in's .pj and .if files byte for byte against them.
int ref_write_pj(const char* path, const char* builtin, const char* geometry, int nDet, int nView,
                 int nSample, double rotLen, double focalRatio, double centerRatio,
                 const float* data, const double* angles, double calcTime, const char* remark) {
  Phantom phm;
  buildPhantom(phm, nullptr, nullptr, 0, builtin);
  Scanner scanner(phm, geometry, nDet, nView, 0, nSample, rotLen, focalRatio, centerRatio, 1.0, 1.0);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  Projections pj(scanner);
  pj.setCalcTime(calcTime);
  pj.setRemark(remark);
  for (int v = 0; v < nView; v++) {
    DetectorArray& d = pj.getDetectorArray(v);
    d.setViewAngle(angles[v]);
    for (int i = 0; i < nDet; i++) d.detValues()[i] = data[(size_t)v * nDet + i];
  }
  return pj.write(path) ? 0 : 1;
}

// Reads a .pj file: info = {nView, nDet, geometry, calcTime, rotStart,
// rotInc, detStart, detInc, viewDiameter, focalLength, SDL, 0} (the fan-beam
// angle has no public accessor);
// data/angles receive at most cap views.
int ref_read_pj(const char* path, double* info, float* data, double* angles, int cap) {
  Projections pj;
  if (!pj.read(path)) { g_err = "read failed"; return 1; }
  const double v[12] = {(double)pj.nView(), (double)pj.nDet(), (double)pj.geometry(),
                        pj.calcTime(), pj.rotStart(), pj.rotInc(), pj.detStart(), pj.detInc(),
                        pj.viewDiameter(), pj.focalLength(), pj.sourceDetectorLength(), 0.0};
  for (int i = 0; i < 12; i++) info[i] = v[i];
  for (int k = 0; k < pj.nView() && k < cap; k++) {
    const DetectorArray& d = pj.getDetectorArray(k);
    angles[k] = d.viewAngle();
    for (int i = 0; i < pj.nDet(); i++) data[(size_t)k * pj.nDet() + i] = d.detValues()[i];
  }
  return 0;
}

// image is column-major [nx][ny], as copyImage writes it.
int ref_write_if(const char* path, int nx, int ny, const float* image, double incX, double incY,
                 double minX, double maxX, double minY, double maxY, const char* label,
                 double calcTime) {
  ImageFile im(nx, ny);
  ImageFileArray a = im.getArray();
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) a[ix][iy] = image[(size_t)ix * ny + iy];
  im.setAxisIncrement(incX, incY);
  im.setAxisExtent(minX, maxX, minY, maxY);
  im.labelAdd(label, calcTime);
  return im.fileWrite(path) ? 0 : 1;
}

int ref_read_if(const char* path, int* nxny, float* image, int cap) {
  ImageFile im;
  if (!im.fileRead(path)) { g_err = "read failed"; return 1; }
  nxny[0] = im.nx();
  nxny[1] = im.ny();
  if ((long)im.nx() * im.ny() > cap) return 2;
  copyImage(im, image);
  return 0;
}

}  // extern "C"

extern "C" {

// Projections::Helical180LI then (if feather) HalfScanFeather on a sinogram
// whose geometry comes from a Scanner on a built-in phantom.  out receives
// at most cap views; returns the number of output views, or < 0.
int ref_helical(const char* builtin, const char* geometry, int nDet, int nView, int nSample,
                double rotLen, double focalRatio, double centerRatio, const float* data,
                int interpView, int feather, float* out, double* anglesOut, int cap,
                double* info4) {
  Phantom phm;
  buildPhantom(phm, nullptr, nullptr, 0, builtin);
  Scanner scanner(phm, geometry, nDet, nView, 0, nSample, rotLen, focalRatio, centerRatio, 1.0, 1.0);
  if (scanner.fail()) { g_err = scanner.failMessage(); return -10; }
  Projections pj(scanner);
  for (int v = 0; v < nView; v++)
    for (int i = 0; i < nDet; i++) pj.getDetectorArray(v).detValues()[i] = data[(size_t)v * nDet + i];
  info4[0] = pj.rotInc();
  info4[1] = pj.detInc();
  info4[2] = scanner.fanBeamAngle();
  info4[3] = pj.sourceDetectorLength();
  if (pj.Helical180LI(interpView) != 0) { g_err = "Helical180LI failed"; return -1; }
  if (feather && pj.HalfScanFeather() != 0) { g_err = "HalfScanFeather failed"; return -2; }
  const int n = pj.nView();
  for (int v = 0; v < n && v < cap; v++) {
    anglesOut[v] = pj.getDetectorArray(v).viewAngle();
    for (int i = 0; i < nDet; i++) out[(size_t)v * nDet + i] = pj.getDetectorArray(v).detValues()[i];
  }
  return n;
}

}  // extern "C"

extern "C" {

// Projections::interpolateToParallel of a sinogram whose geometry comes from
// a Scanner on a built-in phantom.  info8: detStart, detInc, focalLength,
// sdl, viewDiameter, rotStart of the input, new rotInc, new detInc.
int ref_rebin(const char* builtin, const char* geometry, int nDet, int nView, double rotLen,
              double focalRatio, double centerRatio, const float* data, float* out,
              double* anglesIn, double* anglesOut, double* info8) {
  Phantom phm;
  buildPhantom(phm, nullptr, nullptr, 0, builtin);
  Scanner scanner(phm, geometry, nDet, nView, 0, 1, rotLen, focalRatio, centerRatio, 1.0, 1.0);
  if (scanner.fail()) { g_err = scanner.failMessage(); return 2; }
  Projections pj(scanner);
  double a = 0;
  for (int v = 0; v < nView; v++, a += pj.rotInc()) {  // collectProjections' angles (scanner.cpp:330)
    pj.getDetectorArray(v).setViewAngle(a);
    anglesIn[v] = a;
    for (int i = 0; i < nDet; i++) pj.getDetectorArray(v).detValues()[i] = data[(size_t)v * nDet + i];
  }
  Projections* p = pj.interpolateToParallel();
  info8[0] = pj.detStart();
  info8[1] = pj.detInc();
  info8[2] = pj.focalLength();
  info8[3] = pj.sourceDetectorLength();
  info8[4] = pj.viewDiameter();
  info8[5] = pj.rotStart();
  info8[6] = p->rotInc();
  info8[7] = p->detInc();
  for (int v = 0; v < nView; v++) {
    anglesOut[v] = p->getDetectorArray(v).viewAngle();
    for (int i = 0; i < nDet; i++) out[(size_t)v * nDet + i] = p->getDetectorArray(v).detValues()[i];
  }
  if (p != &pj) delete p;
  return 0;
}

}  // extern "C"

extern "C" {

// Phantom::convertToImagefile (phm2if) into img [nx][ny] column-major.
int ref_rasterize(const int* types, const double* params, int nElem, const char* builtin, int nx,
                  int ny, double viewRatio, int nsample, float* img) {
  Phantom phm;
  buildPhantom(phm, types, params, nElem, builtin);
  ImageFile im(nx, ny);
  phm.convertToImagefile(im, viewRatio, nsample, Trace::TRACE_NONE);
  copyImage(im, img);
  return 0;
}

}  // extern "C"
