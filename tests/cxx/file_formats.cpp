// Host-only driver of the drop-in's Scanner geometry and .pj / .if file code
// (include/ctsim/ctsim.h), for the CPU format tests; it makes no GPU call.
//   file_formats pj-write out.pj builtin geometry nDet nView nSample rotLen focal center
//                data.f32 angles.f64 calcTime remark
//   file_formats pj-read  in.pj data.f32 angles.f64      (prints the header fields)
//   file_formats if-write out.if nx ny data.f32 incX incY minX maxX minY maxY label calcTime
//   file_formats if-read  in.if data.f32                 (prints nx ny)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ctsim/ctsim.h"

template <class T>
static std::vector<T> load(const char* path, size_t n) {
  std::vector<T> v(n);
  FILE* f = fopen(path, "rb");
  if (!f || fread(v.data(), sizeof(T), n, f) != n) {
    fprintf(stderr, "cannot read %s\n", path);
    exit(2);
  }
  fclose(f);
  return v;
}

template <class T>
static void save(const char* path, const T* p, size_t n) {
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(p, sizeof(T), n, f) != n) exit(2);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  const std::string mode = argv[1];
  if (mode == "pj-write" && argc == 15) {
    Phantom phm(argv[3]);
    const int nDet = atoi(argv[5]), nView = atoi(argv[6]);
    Scanner sc(phm, argv[4], nDet, nView, 0, atoi(argv[7]), atof(argv[8]), atof(argv[9]),
               atof(argv[10]), 1.0, 1.0);
    if (sc.fail()) return 3;
    Projections pj(sc);
    const auto data = load<float>(argv[11], (size_t)nView * nDet);
    const auto ang = load<double>(argv[12], nView);
    for (int v = 0; v < nView; v++) {
      memcpy(pj.getDetectorArray(v).detValues(), &data[(size_t)v * nDet], sizeof(float) * nDet);
      pj.getDetectorArray(v).setViewAngle(ang[v]);
    }
    pj.setCalcTime(atof(argv[13]));
    pj.setRemark(argv[14]);
    return pj.write(argv[2]) ? 0 : 4;
  }
  if (mode == "pj-read" && argc == 5) {
    Projections pj;
    if (!pj.read(argv[2])) return 4;
    std::vector<double> ang(pj.nView());
    for (int v = 0; v < pj.nView(); v++) ang[v] = pj.getDetectorArray(v).viewAngle();
    save(argv[3], pj.data(), (size_t)pj.nView() * pj.nDet());
    save(argv[4], ang.data(), ang.size());
    printf("%d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", pj.nView(),
           pj.nDet(), pj.geometry(), pj.calcTime(), pj.rotStart(), pj.rotInc(), pj.detStart(),
           pj.detInc(), pj.viewDiameter(), pj.focalLength(), pj.sourceDetectorLength(),
           pj.fanBeamAngle());
    printf("%s\n", pj.remark());
    return 0;
  }
  if (mode == "if-write" && argc == 14) {
    const int nx = atoi(argv[3]), ny = atoi(argv[4]);
    ImageFile im(nx, ny);
    const auto data = load<float>(argv[5], (size_t)nx * ny);
    memcpy(im.data(), data.data(), sizeof(float) * data.size());
    im.setAxisIncrement(atof(argv[6]), atof(argv[7]));
    im.setAxisExtent(atof(argv[8]), atof(argv[9]), atof(argv[10]), atof(argv[11]));
    im.labelAdd(argv[12], atof(argv[13]));
    return im.fileWrite(argv[2]) ? 0 : 4;
  }
  if (mode == "if-read" && argc == 4) {
    ImageFile im;
    if (!im.fileRead(argv[2])) return 4;
    save(argv[3], im.data(), (size_t)im.nx() * im.ny());
    printf("%u %u\n", im.nx(), im.ny());
    return 0;
  }
  fprintf(stderr, "bad arguments\n");
  return 1;
}
