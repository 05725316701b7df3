// Host side of the ray-sum stage: phantom elements (transforms and polygon
// bounding boxes) and scanner geometry, restated from libctsim/phantom.cpp and
// libctsim/scanner.cpp.  Compiled with g++ -O2 -ffp-contract=off so every
// table is bit-identical to the reference's.
#include <strings.h>

#include <cmath>
#include <cstring>

#include "scan_plan.h"

namespace ctsim {

namespace {

using Mtx = double[3][3];

// libctsupport/xform.cpp:99-160
void ident(Mtx m) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = i == j ? 1.0 : 0.0;
}
void xlat(Mtx m, double x, double y) { ident(m); m[2][0] = x; m[2][1] = y; }
void scale(Mtx m, double sx, double sy) { ident(m); m[0][0] = sx; m[1][1] = sy; }
void rot(Mtx m, double t) {
  const double c = cos(t), s = sin(t);
  ident(m);
  m[0][0] = c; m[0][1] = s;
  m[1][0] = -s; m[1][1] = c;
}
void mult(const Mtx a, const Mtx b, Mtx r) {
  Mtx t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t[i][j] = 0;
      for (int k = 0; k < 3; k++) t[i][j] += a[i][k] * b[k][j];
    }
  memcpy(r, t, sizeof(Mtx));
}
void xform(const Mtx m, double& x, double& y) {
  const double xt = x * m[0][0] + y * m[1][0] + m[2][0];
  const double yt = x * m[0][1] + y * m[1][1] + m[2][1];
  x = xt;
  y = yt;
}
void pack(const Mtx m, double* o) {
  o[0] = m[0][0]; o[1] = m[0][1]; o[2] = m[1][0]; o[3] = m[1][1]; o[4] = m[2][0]; o[5] = m[2][1];
}

// PhantomElement::numCirclePoints, phantom.cpp:821-827
int num_circle_points(double theta) {
  theta = theta >= kTwoPi ? kTwoPi : (theta <= 0. ? 0. : theta);
  return static_cast<int>(360 * theta / kTwoPi + 1.5);
}

// PhantomElement::calcArcPoints, phantom.cpp:771-793
void calc_arc(double* x, double* y, int pts, double xc, double yc, double r, double start, double stop) {
  const double th = (stop - start) / (pts - 1);
  const double c = cos(th), s = sin(th);
  x[0] = r * cos(start) + xc;
  y[0] = r * sin(start) + yc;
  double xp = x[0] - xc, yp = y[0] - yc;
  for (int i = 1; i < pts; i++) {
    const double xr = c * xp - s * yp;
    const double yr = s * xp + c * yp;
    x[i] = xr + xc;
    y[i] = yr + yc;
    xp = xr;
    yp = yr;
  }
}

int type_of(const char* name) {  // PhantomElement::convertNameToType, phantom.cpp:598-614
  if (!name) return PE_INVALID;
  if (!strcasecmp(name, "rectangle")) return PE_RECT;
  if (!strcasecmp(name, "triangle")) return PE_TRI;
  if (!strcasecmp(name, "ellipse")) return PE_ELLIPSE;
  if (!strcasecmp(name, "sector")) return PE_SECTOR;
  if (!strcasecmp(name, "segment")) return PE_SEGMENT;
  return PE_INVALID;
}

}  // namespace

// Phantom::addPElem + PhantomElement ctor, phantom.cpp:236-256, 554-566,
// 629-655 (transforms), 669-756 (outline -> bounding box).
int phantom_add(PhantomPlan& p, const char* typeName, double cx, double cy, double u, double v,
                double rotDeg, double atten) {
  const int type = type_of(typeName);
  if (type == PE_INVALID)
    return fail(CTSIM_ERR_INVALID, std::string("Unknown PhantomElement type ") + (typeName ? typeName : "(null)"));
  ElemGPU e{};
  e.type = type;
  e.u = u;
  e.v = v;
  e.atten = atten;
  const double rr = rotDeg * (kPi / 180.);
  Mtx o2p, p2o, t;
  scale(o2p, u, v);
  rot(t, rr);
  mult(o2p, t, o2p);
  xlat(t, cx, cy);
  mult(o2p, t, o2p);
  xlat(p2o, -cx, -cy);
  rot(t, -rr);
  mult(p2o, t, p2o);
  scale(t, 1 / u, 1 / v);
  mult(p2o, t, p2o);
  pack(o2p, e.o2p);
  pack(p2o, e.p2o);

  std::vector<double> xs, ys;
  int n = 0;
  if (type == PE_RECT) {
    xs = {-u, u, u, -u, -u};
    ys = {-v, -v, v, v, -v};
    n = 5;
  } else if (type == PE_TRI) {
    xs = {-u, u, 0.0, -u};
    ys = {0.0, 0.0, v, 0.0};
    n = 4;
  } else if (type == PE_ELLIPSE) {
    n = num_circle_points(kTwoPi);
    xs.resize(n);
    ys.resize(n);
    calc_arc(xs.data(), ys.data(), n, 0.0, 0.0, 1.0, 0.0, kTwoPi);
    for (int i = 0; i < n; i++) { xs[i] *= u; ys[i] *= v; }
  } else {
    const double radius = sqrt(u * u + v * v);
    const double theta = atan(u / v);
    const double start = 3.0 * kHalfPi - theta, stop = 3.0 * kHalfPi + theta;
    const int cpts = num_circle_points(stop - start);
    if (type == PE_SECTOR) {
      n = 3 + cpts;
      xs.resize(n);
      ys.resize(n);
      xs[0] = 0.0; ys[0] = v;
      xs[1] = -u; ys[1] = 0.0;
      calc_arc(xs.data() + 2, ys.data() + 2, cpts, 0.0, v, radius, start, stop);
      xs[cpts + 2] = 0.0; ys[cpts + 2] = v;
    } else {
      n = cpts + 1;
      xs.resize(n);
      ys.resize(n);
      calc_arc(xs.data(), ys.data(), cpts, 0.0, v, radius, start, stop);
      xs[cpts] = -u; ys[cpts] = 0.0;
    }
  }
  // rotate2d, xlat2d (xform.cpp:32-68), minmax_array (ctsupport.h:233-244)
  const double ct = cos(rr), st = sin(rr);
  for (int i = 0; i < n; i++) {
    const double xr = xs[i] * ct - ys[i] * st;
    const double yr = xs[i] * st + ys[i] * ct;
    xs[i] = xr;
    ys[i] = yr;
  }
  for (int i = 0; i < n; i++) { xs[i] += cx; ys[i] += cy; }
  double xmn = xs[0], xmx = xs[0], ymn = ys[0], ymx = ys[0];
  for (int i = 1; i < n; i++) {
    if (xs[i] < xmn) xmn = xs[i];
    else if (xs[i] > xmx) xmx = xs[i];
    if (ys[i] < ymn) ymn = ys[i];
    else if (ys[i] > ymx) ymx = ys[i];
  }
  // SCALE_PELEM_EXTENT = 0 (phantom.cpp:29): the extent is not widened
  const double xf = (xmx - xmn) * 0.0, yf = (ymx - ymn) * 0.0;
  xmn -= xf; ymn -= yf; xmx += xf; ymx += yf;
  e.lim[0] = xmn; e.lim[1] = ymn; e.lim[2] = xmx; e.lim[3] = ymx;
  p.elems.push_back(e);
  p.infos.push_back(ElemInfo{type, n, cx, cy, u, v, rr, atten, rotDeg});
  if (p.xmin > xmn) p.xmin = xmn;
  if (p.xmax < xmx) p.xmax = xmx;
  if (p.ymin > ymn) p.ymin = ymn;
  if (p.ymax < ymx) p.ymax = ymx;
  p.dirty = true;
  return CTSIM_OK;
}

// Phantom::createFromPhantom, addStdHerman, addStdSheppLogan,
// phantom.cpp:128-170, 390-433
int phantom_builtin(PhantomPlan& p, const char* name) {
  if (!name) return fail(CTSIM_ERR_INVALID, "Invalid phantom name (null)");
  struct E { const char* t; double cx, cy, u, v, r, a; };
  static const E herman[] = {
      {"ellipse", 0.000, 1.50, 0.375, 0.3000, 90.00, -0.003},
      {"ellipse", 0.675, -0.75, 0.225, 0.1500, 140.00, 0.010},
      {"ellipse", 0.750, 1.50, 0.375, 0.2250, 50.00, 0.003},
      {"segment", 1.375, -7.50, 1.100, 0.6250, 19.20, -0.204},
      {"segment", 1.375, -7.50, 1.100, 4.3200, 19.21, 0.204},
      {"segment", 0.000, -2.25, 1.125, 0.3750, 0.00, -0.003},
      {"segment", 0.000, -2.25, 1.125, 3.0000, 0.00, 0.003},
      {"segment", -1.000, 3.75, 1.000, 0.5000, 135.00, -0.003},
      {"segment", -1.000, 3.75, 1.000, 3.0000, 135.00, 0.003},
      {"segment", 1.000, 3.75, 1.000, 0.5000, 225.00, -0.003},
      {"segment", 1.000, 3.75, 1.000, 3.0000, 225.00, 0.003},
      {"triangle", 5.025, 3.75, 1.125, 0.5000, 110.75, 0.206},
      {"triangle", -5.025, 3.75, 1.125, 0.9000, -110.75, 0.206},
      {"ellipse", 0.000, 0.00, 8.625, 6.4687, 90.00, 0.416},
      {"ellipse", 0.000, 0.00, 7.875, 5.7187, 90.00, -0.206}};
  static const E shepp[] = {
      {"ellipse", 0.0000, 0.0000, 0.6900, 0.9200, 0.0, 1.00},
      {"ellipse", 0.0000, -0.0184, 0.6624, 0.8740, 0.0, -0.98},
      {"ellipse", 0.2200, 0.0000, 0.1100, 0.3100, -18.0, -0.02},
      {"ellipse", -0.2200, 0.0000, 0.1600, 0.4100, 18.0, -0.02},
      {"ellipse", 0.0000, 0.3500, 0.2100, 0.2500, 0.0, 0.01},
      {"ellipse", 0.0000, 0.1000, 0.0460, 0.0460, 0.0, 0.01},
      {"ellipse", 0.0000, -0.1000, 0.0460, 0.0460, 0.0, 0.01},
      {"ellipse", -0.0800, -0.6050, 0.0460, 0.0230, 0.0, 0.01},
      {"ellipse", 0.0000, -0.6050, 0.0230, 0.0230, 0.0, 0.01},
      {"ellipse", 0.0600, -0.6050, 0.0230, 0.0230, 0.0, 0.01},
      {"ellipse", 0.5538, -0.3858, 0.0330, 0.2060, -18.0, 0.03}};
  const E* list = nullptr;
  int n = 0;
  if (!strcasecmp(name, "herman")) { list = herman; n = 15; }
  else if (!strcasecmp(name, "shepp-logan")) { list = shepp; n = 11; }
  else if (!strcasecmp(name, "unit-pulse")) {
    p.unitPulse = true;
    CTSIM_CHECK(phantom_add(p, "rectangle", 0., 0., 100., 100., 0., 0.));
    return phantom_add(p, "ellipse", 0., 0., 1., 1., 0., 1.);
  } else {
    return fail(CTSIM_ERR_INVALID, std::string("Invalid phantom name ") + name);
  }
  for (int i = 0; i < n; i++)
    CTSIM_CHECK(phantom_add(p, list[i].t, list[i].cx, list[i].cy, list[i].u, list[i].v, list[i].r, list[i].a));
  return CTSIM_OK;
}

// Scanner::Scanner, scanner.cpp:86-224
int scanner_init(ScannerPlan& s, const PhantomPlan& p, const char* geometry, int nDet, int nView,
                 int offsetView, int nSample, double rotLen, double fRatio, double cRatio,
                 double viewRatio, double scanRatio) {
  static const char* const names[] = {"parallel", "equiangular", "equilinear", "linogram"};
  s.geom = -1;
  for (int i = 0; i < 4; i++)
    if (geometry && !strcasecmp(geometry, names[i])) s.geom = i;
  if (s.geom < 0)
    return fail(CTSIM_ERR_INVALID, std::string("Invalid geometry name ") + (geometry ? geometry : "(null)"));
  if (s.geom == GEOM_LINOGRAM)
    return fail(CTSIM_ERR_UNSUPPORTED, "linogram geometry: the reference defines no scanner for it");
  if (nView < 1 || nDet < 1) return fail(CTSIM_ERR_INVALID, "nView & nDet must be greater than 0");
  if (nSample < 1)
    return fail(CTSIM_ERR_INVALID, "nSample must be at least 1 (the reference would divide by zero)");
  if (p.elems.empty()) return fail(CTSIM_ERR_INVALID, "phantom has no elements");
  s.nDet = nDet;
  s.nView = nView;
  s.offsetView = offsetView;
  s.startView = 0;  // the reference leaves m_startView uninitialised (scanner.h:124)
  s.nSample = static_cast<unsigned>(nSample);
  const double maxAxis = (p.xmax - p.xmin) > (p.ymax - p.ymin) ? (p.xmax - p.xmin) : (p.ymax - p.ymin);
  s.viewDiameter = kSqrt2 * maxAxis * viewRatio;
  s.focalLength = (s.viewDiameter / 2) * fRatio;
  s.centerDetLength = (s.viewDiameter / 2) * cRatio;
  s.sdl = s.focalLength + s.centerDetLength;
  s.scanDiameter = s.viewDiameter * scanRatio;
  s.xc = p.xmin + (p.xmax - p.xmin) / 2;
  s.yc = p.ymin + (p.ymax - p.ymin) / 2;
  s.rotLen = rotLen;
  s.rotInc = s.rotLen / s.nView;
  const bool even = (nDet % 2) == 0;
  if (s.geom == GEOM_PARALLEL) {
    s.fanBeamAngle = 0;
    s.detLen = s.scanDiameter;
    s.detStart = -s.detLen / 2;
    s.detInc = s.detLen / s.nDet;
    double endOff = 0;
    if (even) { s.detInc = s.detLen / (s.nDet - 1); endOff = s.detInc; }
    const double h = s.detLen / 2;
    s.xs1 = s.xc - h; s.ys1 = s.yc + s.focalLength;
    s.xs2 = s.xc + h + endOff; s.ys2 = s.yc + s.focalLength;
    s.xd1 = s.xc - h; s.yd1 = s.yc - s.centerDetLength;
    s.xd2 = s.xc + h + endOff; s.yd2 = s.yc - s.centerDetLength;
    s.detLen += endOff;
  } else if (s.geom == GEOM_EQUILINEAR) {
    if (s.scanDiameter / 2 >= s.focalLength)
      return fail(CTSIM_ERR_INVALID, "Invalid geometry: Focal length must be larger than scan length");
    const double a = asin((s.scanDiameter / 2) / s.focalLength);
    const double h = s.sdl * tan(a);
    s.detLen = h * 2;
    s.detStart = -h;
    s.detInc = s.detLen / s.nDet;
    double endOff = 0;
    if (even) { s.detInc = s.detLen / (s.nDet - 1); endOff = s.detInc; s.detLen += endOff; }
    s.fanBeamAngle = a * 2;
    s.xs1 = s.xc; s.ys1 = s.yc + s.focalLength;
    s.xs2 = s.xc; s.ys2 = s.yc + s.focalLength;
    s.xd1 = s.xc - h; s.yd1 = s.yc - s.centerDetLength;
    s.xd2 = s.xc + h + endOff; s.yd2 = s.yc - s.centerDetLength;
  } else {
    if (s.scanDiameter / 2 > s.focalLength)
      return fail(CTSIM_ERR_INVALID, "Invalid geometry: Focal length must be larger than scan length");
    const double a = asin((s.scanDiameter / 2) / s.focalLength);
    s.detLen = 2 * a;
    s.detStart = -a;
    s.detInc = s.detLen / s.nDet;
    double endOff = 0;
    if (even) { s.detInc = s.detLen / (s.nDet - 1); endOff = s.detInc; }
    const double a1 = acos((s.scanDiameter / 2) / s.centerDetLength);
    const double angScale = 2 * (kHalfPi + a - a1) / s.detLen;
    s.angDetLen = angScale * (s.detLen + endOff);
    s.angDetInc = angScale * s.detInc;
    s.fanBeamAngle = a * 2;
    s.xs1 = s.xc; s.ys1 = s.yc + s.focalLength;
    s.xs2 = s.xc; s.ys2 = s.yc + s.focalLength;
    s.detLen += endOff;
  }
  Mtx t;
  xlat(s.rotIncM, -s.xc, -s.yc);
  rot(t, s.rotInc);
  mult(s.rotIncM, t, s.rotIncM);
  xlat(t, s.xc, s.yc);
  mult(s.rotIncM, t, s.rotIncM);
  return CTSIM_OK;
}

// Endpoints of every view, Scanner::collectProjections, scanner.cpp:302-419:
// the start angle's rotation matrix, then one incremental rotation per view.
void scanner_view_ends(const ScannerPlan& s, int iStartView, int iNumViews, int iOffsetView,
                       ViewEnds* out) {
  const double start = (iStartView + iOffsetView) * s.rotInc;
  Mtx m0, t;
  xlat(m0, -s.xc, -s.yc);
  rot(t, start);
  mult(m0, t, m0);
  xlat(t, s.xc, s.yc);
  mult(m0, t, m0);
  double xd1 = 0, yd1 = 0, xd2 = 0, yd2 = 0;
  if (s.geom != GEOM_EQUIANGULAR) {
    xd1 = s.xd1; yd1 = s.yd1; xd2 = s.xd2; yd2 = s.yd2;
    xform(m0, xd1, yd1);
    xform(m0, xd2, yd2);
  }
  double xs1 = s.xs1, ys1 = s.ys1, xs2 = s.xs2, ys2 = s.ys2;
  xform(m0, xs1, ys1);
  xform(m0, xs2, ys2);
  double va = start;
  for (int v = 0; v < iNumViews; v++, va += s.rotInc) {
    out[v] = {xd1, yd1, xd2, yd2, xs1, ys1, xs2, ys2, va + 3 * kHalfPi, va};
    xform(s.rotIncM, xs1, ys1);
    xform(s.rotIncM, xs2, ys2);
    if (s.geom != GEOM_EQUIANGULAR) {
      xform(s.rotIncM, xd1, yd1);
      xform(s.rotIncM, xd2, yd2);
    }
  }
}

}  // namespace ctsim
