// Ray-sum kernels for gfx950 (Scanner::collectProjections,
// libctsim/scanner.cpp:280-656).
//
// One lane owns one (view, detector): it walks the detector's sample rays and
// sums every phantom element's clipped chord times its attenuation in FP64,
// then stores sum / nSample as float (scanner.cpp:522).  All lanes of a wave
// visit the same elements in the same order, so element data is read with
// wave-uniform loads.  The detector and source positions are repeated sums
// along the detector array (scanner.cpp:490-532); a first pass with one lane
// per view performs that walk so every lane starts from the reference's
// values.
//
//   F64 path    the reference's clipping step order: bounding-box reject by
//               Cohen-Sutherland on the polygon outline's box, map to the
//               element frame, circle clip by translate / rotate (atan2, cos,
//               sin) / clamp / rotate back, map back, length.
//   mixed path  the same box reject, then closed-form FP64 chords for
//               ellipses (quadratic in the element frame) and rectangles
//               (slab clip); triangles, segments and sectors keep the
//               reference's clip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "crmath.h"
#include "scan_plan.h"

#pragma clang fp contract(off)

namespace ctsim {
namespace {

// Per-view constants of projectSingleView, scanner.cpp:449-473.
struct ViewSetup {
  double sdx, sdy, ddx, ddy, ddx2, ddy2, aInc, aSampInc;
};

__device__ __forceinline__ ViewSetup view_setup(const ViewEnds& e, int nDet, unsigned nSample,
                                                int geom, double angDetInc) {
  ViewSetup s;
  s.sdx = (e.xs2 - e.xs1) / nDet;
  s.sdy = (e.ys2 - e.ys1) / nDet;
  s.ddx = s.ddy = s.ddx2 = s.ddy2 = s.aInc = s.aSampInc = 0;
  if (geom == GEOM_EQUIANGULAR) {
    s.aInc = angDetInc;
    s.aSampInc = s.aInc / nSample;
  } else {
    s.ddx = (e.xd2 - e.xd1) / nDet;
    s.ddy = (e.yd2 - e.yd1) / nDet;
    s.ddx2 = s.ddx / nSample;
    s.ddy2 = s.ddy / nSample;
  }
  return s;
}

struct WalkArgs {
  const ViewEnds* ends;
  double4* walk;  // [nView][nDet]: xs, ys, xd (or the detector angle), yd
  int* counter;   // the ray-sum kernel's work counter, zeroed here
  int nView, nDet, geom;
  unsigned nSample;
  double angDetInc, angDetLen;
};

// The detector walk of projectSingleView (scanner.cpp:449-531): per view
// four chains of rounded sums (xs, ys, and xd, yd or the detector angle),
// one lane per chain, 16 views per block.  Each run of 64 detectors is
// gathered in LDS and stored as whole rows of the walk.
constexpr int kWalkViews = 16;
constexpr int kWalkRow = 64 * 4 + 4;  // doubles per view in LDS (padded against bank conflicts)
__global__ __launch_bounds__(64) void raysum_walk_kernel(WalkArgs a) {
  __shared__ double tile[kWalkViews * kWalkRow];
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.counter = 0;
  const int lane = threadIdx.x, vl = lane >> 2, comp = lane & 3, v0 = blockIdx.x * kWalkViews;
  const int v = v0 + vl;
  double x = 0.0, inc = 0.0;
  if (v < a.nView) {
    const ViewEnds e = a.ends[v];
    const ViewSetup s = view_setup(e, a.nDet, a.nSample, a.geom, a.angDetInc);
    const bool equi = a.geom == GEOM_EQUIANGULAR;
    if (comp == 0) { x = e.xs1 + (s.sdx / 2); inc = s.sdx; }
    else if (comp == 1) { x = e.ys1 + (s.sdy / 2); inc = s.sdy; }
    else if (comp == 2) {
      if (equi) { x = e.detAngle - (a.angDetLen / 2) + s.aSampInc / 2; inc = s.aInc; }
      else { x = e.xd1 + s.ddx2 / 2; inc = s.ddx; }
    } else if (!equi) {  // yd (the equiangular walk keeps 0 here)
      x = e.yd1 + s.ddy2 / 2;
      inc = s.ddy;
    }
  }
  double* mine = tile + vl * kWalkRow + comp;
  for (int d0 = 0; d0 < a.nDet; d0 += 64) {
    const int n = min(64, a.nDet - d0);
    if (n == 64) {  // a whole run: unrolled, the stores off the chain of adds
#pragma unroll
      for (int j = 0; j < 64; j++) {
        mine[4 * j] = x;
        x += inc;
      }
    } else {
      for (int j = 0; j < n; j++) {
        mine[4 * j] = x;
        x += inc;
      }
    }
    __syncthreads();
    // row tv's cells tj, 64 lanes a round (n == 64: one view row per round)
    for (int t = lane; t < kWalkViews * n; t += 64) {
      const int tv = n == 64 ? t >> 6 : t / n, tj = n == 64 ? t & 63 : t - tv * n;
      if (v0 + tv < a.nView) {
        const double* c = tile + tv * kWalkRow + 4 * tj;
        a.walk[static_cast<size_t>(v0 + tv) * a.nDet + d0 + tj] = make_double4(c[0], c[1], c[2], c[3]);
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------- reference clipping

__device__ __forceinline__ void xf(const double* m, double& x, double& y) {
  const double xt = x * m[0] + y * m[2] + m[4];
  const double yt = x * m[1] + y * m[3] + m[5];
  x = xt;
  y = yt;
}

// normalizeAngle, libctsupport/mathfuncs.cpp:72-81
__device__ __forceinline__ double normalize_angle(double t) {
  while (t < 0.) t += kTwoPi;
  while (t >= kTwoPi) t -= kTwoPi;
  return t;
}

// rectcode + clip_rect, libctsupport/clip.cpp:332-395
__device__ __forceinline__ int rectcode(double x, double y, const double* r) {
  int c = 0;
  if (x < r[0]) c = 1;
  else if (x > r[2]) c = 2;
  if (y < r[1]) c += 4;
  else if (y > r[3]) c += 8;
  return c;
}
__device__ bool clip_rect(double& x1, double& y1, double& x2, double& y2, const double* r) {
  double x = 0, y = 0;
  int c1 = rectcode(x1, y1, r), c2 = rectcode(x2, y2, r);
  while (c1 || c2) {
    if (c1 & c2) return false;
    const int c = c1 ? c1 : c2;
    if (c & 1) { y = y1 + (y2 - y1) * (r[0] - x1) / (x2 - x1); x = r[0]; }
    else if (c & 2) { y = y1 + (y2 - y1) * (r[2] - x1) / (x2 - x1); x = r[2]; }
    else if (c & 4) { x = x1 + (x2 - x1) * (r[1] - y1) / (y2 - y1); y = r[1]; }
    else if (c & 8) { x = x1 + (x2 - x1) * (r[3] - y1) / (y2 - y1); y = r[3]; }
    if (c == c1) { x1 = x; y1 = y; c1 = rectcode(x1, y1, r); }
    else { x2 = x; y2 = y; c2 = rectcode(x2, y2, r); }
  }
  return true;
}

// clip_circle, libctsupport/clip.cpp:126-198 (rot_mtx2 / xform_mtx2 inlined)
__device__ bool clip_circle(double& x1, double& y1, double& x2, double& y2, double cx, double cy,
                            double radius, double t1, double t2) {
  double xc1 = x1, yc1 = y1, xc2 = x2, yc2 = y2, ccx = cx, ccy = cy;
  const double xtrans = -xc1, ytrans = -yc1;
  xc1 += xtrans; yc1 += ytrans; xc2 += xtrans; yc2 += ytrans; ccx += xtrans; ccy += ytrans;
  const double theta = -crm_atan2(yc2, xc2);
  // correctly rounded cos and sin are exactly even and odd, so rotating back
  // by -theta below reuses them as the reference's cos(-theta), sin(-theta)
  double c, s;
  crm_sincos(theta, &s, &c);
  const double cb = c, sb = -s;
  {  // xform by rot(theta): x' = x c + y (-s) + 0, y' = x s + y c + 0
    double xt = xc2 * c + yc2 * -s + 0.0, yt = xc2 * s + yc2 * c + 0.0;
    xc2 = xt; yc2 = yt;
    xt = ccx * c + ccy * -s + 0.0; yt = ccx * s + ccy * c + 0.0;
    ccx = xt; ccy = yt;
  }
  t1 += theta;
  t2 += theta;
  t1 = normalize_angle(t1);
  t2 = normalize_angle(t2);
  if (xc2 < -kDEps || fabs(yc2) > kFEps) return false;  // reference: ERR_SEVERE, reject
  if (fabs(ccy) > radius) return false;
  const double temp = sqrt(radius * radius - ccy * ccy);
  const double xcmin = ccx - temp, xcmax = ccx + temp;
  if (fabs(t2 - t1) < kDEps) {
    if (xc1 < xcmin) xc1 = xcmin;
    if (xc2 > xcmax) xc2 = xcmax;
  } else if (t1 < t2) {
    if (t1 < kPi && t2 > kPi)
      if (xc1 < xcmin) xc1 = xcmin;
  } else if (t1 > t2) {
    if (t1 < kPi)
      if (xc1 < xcmin) xc1 = xcmin;
    if (xc2 > xcmax) xc2 = xcmax;
  }
  c = cb;
  s = sb;
  {
    double xt = xc1 * c + yc1 * -s + 0.0, yt = xc1 * s + yc1 * c + 0.0;
    xc1 = xt; yc1 = yt;
    xt = xc2 * c + yc2 * -s + 0.0; yt = xc2 * s + yc2 * c + 0.0;
    xc2 = xt; yc2 = yt;
  }
  xc1 += -xtrans; yc1 += -ytrans; xc2 += -xtrans; yc2 += -ytrans;
  x1 = xc1; y1 = yc1; x2 = xc2; y2 = yc2;
  return true;
}

// tcode + clip_triangle, libctsupport/clip.cpp:243-316
__device__ __forceinline__ int tcode(double x, double y, double m, double b, bool clipX) {
  int c = 0;
  if (clipX && y < 0.) c = 1;
  if (y > -m * x + b + kDEps) c += 2;
  if (y > m * x + b + kDEps) c += 4;
  return c;
}
__device__ bool clip_triangle(double& x1, double& y1, double& x2, double& y2, double u, double v,
                              bool clipX) {
  const double m = v / u, b = v;
  int c1 = tcode(x1, y1, m, b, clipX), c2 = tcode(x2, y2, m, b, clipX);
  while (c1 || c2) {
    if (c1 & c2) return false;
    const int c = c1 ? c1 : c2;
    double x = 0, y = 0;
    if (c & 1) {
      x = x1 + (x2 - x1) * (0.0 - y1) / (y2 - y1);
      y = 0.0;
    } else if (c & 2) {
      const double dx = x2 - x1, dy = y2 - y1;
      if (fabs(dx) > kDEps) x = (-y1 + b + x1 * dy / dx) / (m + dy / dx);
      else x = x1;
      y = -m * x + b;
    } else if (c & 4) {
      const double dx = x2 - x1, dy = y2 - y1;
      if (fabs(dx) > kDEps) {
        x = (y1 - b - x1 * dy / dx);
        x /= (m - dy / dx);
      } else {
        x = x1;
      }
      y = m * x + b;
    }
    if (c == c1) { x1 = x; y1 = y; c1 = tcode(x1, y1, m, b, clipX); }
    else { x2 = x; y2 = y; c2 = tcode(x2, y2, m, b, clipX); }
  }
  return true;
}

// clip_segment (libctsupport/clip.cpp:44-77) and clip_sector (clip.cpp:
// 90-111) share the ellipse's clip_circle call below, so the kernel holds
// one copy of the circle clip (and its atan2, sin and cos): each case does
// the same operations in the same order as the reference.

// False only when the segment certainly misses the box: separating axes
// (the box's two axes and the segment's normal) on the box widened by a
// margin far above the rounding of the reference's Cohen-Sutherland clip,
// so every segment that clip keeps passes.  No division.
__device__ __forceinline__ bool seg_may_hit_box(double x1, double y1, double x2, double y2, const double* r) {
  const double scale = fmax(fmax(fabs(r[0]), fabs(r[2])), fmax(fmax(fabs(r[1]), fabs(r[3])), 1.0));
  const double m = 1e-9 * scale;
  const double bx0 = r[0] - m, by0 = r[1] - m, bx1 = r[2] + m, by1 = r[3] + m;
  if (fmax(x1, x2) < bx0 || fmin(x1, x2) > bx1 || fmax(y1, y2) < by0 || fmin(y1, y2) > by1) return false;
  const double nx = y1 - y2, ny = x2 - x1;  // the segment's normal
  const double s0 = nx * (bx0 - x1) + ny * (by0 - y1), s1 = nx * (bx1 - x1) + ny * (by0 - y1);
  const double s2 = nx * (bx0 - x1) + ny * (by1 - y1), s3 = nx * (bx1 - x1) + ny * (by1 - y1);
  const double tol = m * (fabs(nx) + fabs(ny));
  const double lo = fmin(fmin(s0, s1), fmin(s2, s3)), hi = fmax(fmax(s0, s1), fmax(s2, s3));
  return lo <= tol && hi >= -tol;
}

// True only when the segment certainly crosses the box: the same test on the
// box narrowed by that margin, with strict inequalities.  Such a segment is
// kept by the reference's Cohen-Sutherland clip, whose clipped points the
// chord discards, so the clip can be skipped.
__device__ __forceinline__ bool seg_surely_hits_box(double x1, double y1, double x2, double y2, const double* r) {
  const double scale = fmax(fmax(fabs(r[0]), fabs(r[2])), fmax(fmax(fabs(r[1]), fabs(r[3])), 1.0));
  const double m = 1e-9 * scale;
  const double bx0 = r[0] + m, by0 = r[1] + m, bx1 = r[2] - m, by1 = r[3] - m;
  if (!(bx0 < bx1 && by0 < by1)) return false;
  if (!(fmax(x1, x2) > bx0 && fmin(x1, x2) < bx1 && fmax(y1, y2) > by0 && fmin(y1, y2) < by1)) return false;
  const double nx = y1 - y2, ny = x2 - x1;
  const double s0 = nx * (bx0 - x1) + ny * (by0 - y1), s1 = nx * (bx1 - x1) + ny * (by0 - y1);
  const double s2 = nx * (bx0 - x1) + ny * (by1 - y1), s3 = nx * (bx1 - x1) + ny * (by1 - y1);
  const double tol = m * (fabs(nx) + fabs(ny));
  const double lo = fmin(fmin(s0, s1), fmin(s2, s3)), hi = fmax(fmax(s0, s1), fmax(s2, s3));
  return lo < -tol && hi > tol;
}

// False only when the line through the segment certainly misses the unit
// circle of the element frame m (an ellipse): its distance from the centre,
// |p1 x p2| / |p2 - p1| in that frame, exceeds 1 by a margin far above the
// rounding of the reference's circle clip, which then rejects the segment
// (clip.cpp:152, |ccy| > radius).
__device__ __forceinline__ bool line_may_hit_unit_circle(const double* m, double x1, double y1, double x2,
                                                         double y2) {
  const double X1 = x1 * m[0] + y1 * m[2] + m[4], Y1 = x1 * m[1] + y1 * m[3] + m[5];
  const double X2 = x2 * m[0] + y2 * m[2] + m[4], Y2 = x2 * m[1] + y2 * m[3] + m[5];
  const double dx = X2 - X1, dy = Y2 - Y1, cr = X1 * Y2 - X2 * Y1;
  const double sc = fmax(fmax(fabs(X1), fabs(Y1)), fmax(fmax(fabs(X2), fabs(Y2)), 1.0));
  const double r = 1.0 + 1e-9 * sc;
  return cr * cr <= r * r * (dx * dx + dy * dy);
}

// PhantomElement::clipLineWorldCoords / clipLineNormalizedCoords and
// Scanner::projectLineAgainstPElem, phantom.cpp:830-898, scanner.cpp:633-656
__device__ __forceinline__ double chord_ref(const ElemGPU& e, double x1, double y1, double x2, double y2) {
  if (!seg_surely_hits_box(x1, y1, x2, y2, e.lim)) {  // else the box clip keeps the segment
    double cx1 = x1, cy1 = y1, cx2 = x2, cy2 = y2;
    if (!clip_rect(cx1, cy1, cx2, cy2, e.lim)) return 0.0;
  }
  xf(e.p2o, x1, y1);
  xf(e.p2o, x2, y2);
  bool ok = false;
  const int type = e.type;
  if (type == PE_RECT) {
    const double unit[4] = {-1.0, -1.0, 1.0, 1.0};
    ok = clip_rect(x1, y1, x2, y2, unit);
  } else if (type == PE_TRI) {
    ok = clip_triangle(x1, y1, x2, y2, 1.0, 1.0, true);
  } else if (type == PE_ELLIPSE || type == PE_SEGMENT || type == PE_SECTOR) {
    const bool scaled = type != PE_ELLIPSE;  // segment, sector: x * u, y * v
    const double u = e.u, v = e.v;
    double xc1 = x1, yc1 = y1, xc2 = x2, yc2 = y2;
    double ccy = 0.0, radius = 1.0;
    ok = true;
    if (scaled) {
      xc1 = x1 * u; yc1 = y1 * v; xc2 = x2 * u; yc2 = y2 * v;
      if (type == PE_SEGMENT && yc1 > 0. && yc2 > 0.) ok = false;
      radius = sqrt(u * u + v * v);
      ccy = v;
    }
    if (ok) ok = clip_circle(xc1, yc1, xc2, yc2, 0.0, ccy, radius, 0.0, 0.0);
    if (ok && type == PE_SEGMENT) {
      if (yc1 > 0. && yc2 > 0.) {
        ok = false;
      } else if (yc1 > 0.) {
        xc1 = xc1 + (xc2 - xc1) * (0.0 - yc1) / (yc2 - yc1);
        yc1 = 0.0;
      } else if (yc2 > 0.) {
        xc2 = xc1 + (xc2 - xc1) * (0.0 - yc1) / (yc2 - yc1);
        yc2 = 0.0;
      }
    }
    if (ok && type == PE_SECTOR) ok = clip_triangle(xc1, yc1, xc2, yc2, u, v, false);
    if (ok) {
      if (scaled) {
        x1 = xc1 / u; y1 = yc1 / v; x2 = xc2 / u; y2 = yc2 / v;
      } else {
        x1 = xc1; y1 = yc1; x2 = xc2; y2 = yc2;
      }
    }
  }
  if (!ok) return 0.0;
  xf(e.o2p, x1, y1);
  xf(e.o2p, x2, y2);
  const double len = sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
  return len * e.atten;
}

// ------------------------------------------------- closed-form chords
#pragma clang fp contract(fast)

// Segment-box test by slabs (the reference rejects with Cohen-Sutherland on
// the same box; the two agree except for rays grazing a box corner).
__device__ __forceinline__ bool seg_hits_box(double x1, double y1, double dx, double dy,
                                             const double* r) {
  double t0 = 0.0, t1 = 1.0;
  if (dx != 0.0) {
    const double inv = 1.0 / dx;
    double ta = (r[0] - x1) * inv, tb = (r[2] - x1) * inv;
    if (ta > tb) { const double t = ta; ta = tb; tb = t; }
    t0 = fmax(t0, ta);
    t1 = fmin(t1, tb);
  } else if (x1 < r[0] || x1 > r[2]) {
    return false;
  }
  if (dy != 0.0) {
    const double inv = 1.0 / dy;
    double ta = (r[1] - y1) * inv, tb = (r[3] - y1) * inv;
    if (ta > tb) { const double t = ta; ta = tb; tb = t; }
    t0 = fmax(t0, ta);
    t1 = fmin(t1, tb);
  } else if (y1 < r[1] || y1 > r[3]) {
    return false;
  }
  return t0 <= t1;
}

// The reference's clip chain as a call: the closed form takes it only for
// triangles, segments and sectors, whose registers then stay out of the
// ellipse and rectangle path.
__device__ __attribute__((noinline)) double chord_ref_call(const ElemGPU& e, double x1, double y1, double x2,
                                                           double y2) {
  return chord_ref(e, x1, y1, x2, y2);
}

// Parameter interval of the segment inside the element, in the element frame.
__device__ double chord_closed(const ElemGPU& e, double x1, double y1, double x2, double y2,
                               double segLen) {
  const double dx = x2 - x1, dy = y2 - y1;
  if (!seg_hits_box(x1, y1, dx, dy, e.lim)) return 0.0;
  if (e.type == PE_ELLIPSE || e.type == PE_RECT) {
    const double px = x1 * e.p2o[0] + y1 * e.p2o[2] + e.p2o[4];
    const double py = x1 * e.p2o[1] + y1 * e.p2o[3] + e.p2o[5];
    const double qx = dx * e.p2o[0] + dy * e.p2o[2];
    const double qy = dx * e.p2o[1] + dy * e.p2o[3];
    double t0, t1;
    if (e.type == PE_ELLIPSE) {
      // |p + t q|^2 = 1
      const double a = qx * qx + qy * qy, b = px * qx + py * qy, c = px * px + py * py - 1.0;
      const double disc = b * b - a * c;
      if (!(disc > 0.0)) return 0.0;
      const double sq = sqrt(disc);
      t0 = (-b - sq) / a;
      t1 = (-b + sq) / a;
    } else {
      t0 = -1e300;
      t1 = 1e300;
      if (qx != 0.0) {
        double ta = (-1.0 - px) / qx, tb = (1.0 - px) / qx;
        if (ta > tb) { const double t = ta; ta = tb; tb = t; }
        t0 = fmax(t0, ta);
        t1 = fmin(t1, tb);
      } else if (px < -1.0 || px > 1.0) {
        return 0.0;
      }
      if (qy != 0.0) {
        double ta = (-1.0 - py) / qy, tb = (1.0 - py) / qy;
        if (ta > tb) { const double t = ta; ta = tb; tb = t; }
        t0 = fmax(t0, ta);
        t1 = fmin(t1, tb);
      } else if (py < -1.0 || py > 1.0) {
        return 0.0;
      }
    }
    t0 = fmax(t0, 0.0);
    t1 = fmin(t1, 1.0);
    return t1 > t0 ? (t1 - t0) * segLen * e.atten : 0.0;
  }
  return chord_ref_call(e, x1, y1, x2, y2);
}

#pragma clang fp contract(off)

struct RaysumArgs {
  const ViewEnds* ends;
  const double4* walk;
  const ElemGPU* __restrict__ elems;
  float* sino;
  int* counter;  // items taken so far (zeroed by the walk kernel)
  int nView, nDet, nElem, geom, storageOffset;
  unsigned nSample;
  double angDetInc, centerDetLength;
  int unitPulse;
};

// Detector end (a) and source end (b) of sample i of detector d's ray, as
// raysum_ray walks them (to rounding: the hull test below has a margin).
__device__ __forceinline__ void ray_ends(const RaysumArgs& a, const ViewSetup& s, int v, int d, unsigned i,
                                         double& ax, double& ay, double& bx, double& by) {
  const double4 w = a.walk[static_cast<size_t>(v) * a.nDet + d];
  bx = w.x;
  by = w.y;
  if (a.geom == GEOM_EQUIANGULAR) {
    const double ang = w.z + i * s.aSampInc;
    ax = a.centerDetLength * cos(ang);
    ay = a.centerDetLength * sin(ang);
  } else {
    ax = w.z + i * s.ddx2;
    ay = w.w + i * s.ddy2;
  }
}

// Candidate elements of one wave's rays (detectors d0..d1 of view v, every
// sample): lane k tests element 64 c + k's box against the hull of the
// rays, the quadrilateral of the first and the last ray (widened by the
// sagitta of the detector arc in the equiangular geometry and a margin), by
// separating axes; bit k of mask c is set unless the box is certainly
// outside.  Every ray of the wave lies in the hull, so a clear bit marks an
// element whose chords are all exactly zero.
constexpr int kCandChunks = 16;  // up to 1024 elements; more take no culling
__device__ __forceinline__ void wave_candidates(const RaysumArgs& a, const ViewSetup& s, int v, int d0, int d1,
                                                unsigned long long* masks, int nChunks) {
  double px[4], py[4];
  ray_ends(a, s, v, d0, 0, px[0], py[0], px[1], py[1]);
  ray_ends(a, s, v, d1, a.nSample - 1, px[3], py[3], px[2], py[2]);
  double mg = 1e-9 * fmax(fmax(fmax(fabs(px[0]), fabs(px[1])), fmax(fabs(px[2]), fabs(px[3]))),
                          fmax(fmax(fmax(fabs(py[0]), fabs(py[1])), fmax(fabs(py[2]), fabs(py[3]))), 1.0));
  if (a.geom == GEOM_EQUIANGULAR) {  // the detector ends lie on an arc bulging past the chord
    const double cx = px[3] - px[0], cy = py[3] - py[0];
    mg += (cx * cx + cy * cy) / (2.0 * a.centerDetLength) + 1e-12;
  }
  const double qx0 = fmin(fmin(px[0], px[1]), fmin(px[2], px[3])), qx1 = fmax(fmax(px[0], px[1]), fmax(px[2], px[3]));
  const double qy0 = fmin(fmin(py[0], py[1]), fmin(py[2], py[3])), qy1 = fmax(fmax(py[0], py[1]), fmax(py[2], py[3]));
  const int lane = threadIdx.x;
  for (int c = 0; c < nChunks; c++) {
    const int k = 64 * c + lane;
    bool cand = false;
    if (k < a.nElem) {
      const double* r = a.elems[k].lim;
      cand = !(qx1 < r[0] - mg || qx0 > r[2] + mg || qy1 < r[1] - mg || qy0 > r[3] + mg);
#pragma unroll
      for (int ed = 0; ed < 4 && cand; ed++) {
        const int e2 = (ed + 1) & 3;
        const double nx = py[ed] - py[e2], ny = px[e2] - px[ed];
        double h0 = 1e300, h1 = -1e300;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const double p = nx * px[j] + ny * py[j];
          h0 = fmin(h0, p);
          h1 = fmax(h1, p);
        }
        const double c00 = nx * r[0] + ny * r[1], c10 = nx * r[2] + ny * r[1];
        const double c01 = nx * r[0] + ny * r[3], c11 = nx * r[2] + ny * r[3];
        const double b0 = fmin(fmin(c00, c10), fmin(c01, c11)), b1 = fmax(fmax(c00, c10), fmax(c01, c11));
        const double tol = mg * (fabs(nx) + fabs(ny)) + 1e-300;
        if (b1 < h0 - tol || b0 > h1 + tol) cand = false;
      }
    }
    const unsigned long long bits = __ballot(cand);
    if (lane == 0) masks[c] = bits;
  }
}

// A wave's deferred chords.  The (lane, element) pairs whose chord is
// needed are queued in the reference's element order and computed 64 at a
// time, one pair per lane, so lanes whose rays miss an element do not idle
// through its clip; each lane then adds its own chords in queue order, the
// order of the element list.  Rings of kQueue entries: at most 63 pairs
// wait before an element's (up to 64) are added, and a flush of 64 follows.
constexpr int kQueue = 128;
struct WaveQueue {
  double4 ray[64];                  // each lane's ray: detector end, source end
  double val[kQueue];               // computed chords, by pair slot
  unsigned long long vmask[kQueue];  // per queued element: the lanes that need it
  int vbase[kQueue];                // ... and the slot of its first pair
  int pair[kQueue];                 // lane | element << 6, by pair slot
};

__device__ __forceinline__ int lanes_below(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(m), 0));
}

__device__ __forceinline__ unsigned long long uniform_u64(unsigned long long x) {
  return (static_cast<unsigned long long>(static_cast<unsigned>(
              __builtin_amdgcn_readfirstlane(static_cast<unsigned>(x >> 32)))) << 32) |
         static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(x)));
}

// One lane's ray sum: detector d of view v (a lane past the last detector,
// active false, only helps its wave's chords).  masks: the nChunks
// candidate masks (wave_candidates), or nullptr for every element.
__device__ __forceinline__ void raysum_ray(const RaysumArgs& a, const ViewSetup& s, int v, int d, bool active,
                                           const unsigned long long* masks, int nChunks, WaveQueue* qp) {
  const int lane = threadIdx.x;
  const double4 w = a.walk[static_cast<size_t>(v) * a.nDet + (active ? d : a.nDet - 1)];
  const double xs = w.x, ys = w.y;
  double xd = w.z, yd = w.w, ang = w.z;
  double sum = 0.0;
  for (unsigned i = 0; i < a.nSample; i++) {
    if (a.geom == GEOM_EQUIANGULAR) {
      double sa, ca;
      crm_sincos(ang, &sa, &ca);
      xd = a.centerDetLength * ca;
      yd = a.centerDetLength * sa;
    }
    qp->ray[lane] = make_double4(xd, yd, xs, ys);
    // projectSingleLine: the element list is newest-first (phantom.cpp:247).
    // A segment that misses an element's box, or an ellipse's line that
    // misses it, adds the reference's exact zero, which changes no sum: such
    // pairs are never queued (elements outside the wave's hull by the masks,
    // then lane by lane).
    double rsum = 0.0;
    int enq = 0, done = 0, nv = 0, vc = 0;
    auto flush = [&](int n) {
      __syncthreads();  // rays, pairs and visits written above are visible
      if (lane < n) {
        const int slot = (done + lane) & (kQueue - 1);
        const int pr = qp->pair[slot];
        const double4 r = qp->ray[pr & 63];
        qp->val[slot] = chord_ref(a.elems[pr >> 6], r.x, r.y, r.z, r.w);
      }
      __syncthreads();
      const int prev = done;
      done += n;
      // each lane adds its chords among slots [prev, done), element by element
      while (vc < nv) {
        const int vs = vc & (kQueue - 1);
        const unsigned long long vm = uniform_u64(qp->vmask[vs]);
        const int vb = __builtin_amdgcn_readfirstlane(qp->vbase[vs]);
        if ((vm >> lane) & 1) {
          const int slot = vb + lanes_below(vm);
          if (slot >= prev && slot < done) rsum += qp->val[slot & (kQueue - 1)];
        }
        if (vb + __popcll(vm) > done) break;  // the rest of this element comes later
        vc++;
      }
    };
    for (int c = nChunks - 1; c >= 0; c--) {
      unsigned long long m;
      if (masks) {
        m = uniform_u64(masks[c]);
      } else {  // every element of the chunk
        const int rem = a.nElem - 64 * c;
        m = rem >= 64 ? ~0ull : (1ull << rem) - 1;
      }
      while (m) {
        const int b = 63 - __builtin_clzll(m);
        m &= ~(1ull << b);
        const int k = 64 * c + b;
        // an ellipse by its line alone (a line that misses it clips to zero;
        // the rays span the phantom, so the line test queues next to no pair
        // whose segment the clip then rejects)
        const bool may = active && (a.elems[k].type == PE_ELLIPSE
                                        ? line_may_hit_unit_circle(a.elems[k].p2o, xd, yd, xs, ys)
                                        : seg_may_hit_box(xd, yd, xs, ys, a.elems[k].lim));
        const unsigned long long hit = __ballot(may);
        if (!hit) continue;
        if (may) qp->pair[(enq + lanes_below(hit)) & (kQueue - 1)] = lane | (k << 6);
        if (lane == 0) {
          qp->vmask[nv & (kQueue - 1)] = hit;
          qp->vbase[nv & (kQueue - 1)] = enq;
        }
        nv++;
        enq += __popcll(hit);
        if (enq - done >= 64) flush(64);
      }
    }
    if (enq > done) flush(enq - done);
    sum += rsum;
    if (a.geom == GEOM_EQUIANGULAR) ang += s.aSampInc;
    else { xd += s.ddx2; yd += s.ddy2; }
  }
  if (active) {
    float* out = a.sino + static_cast<size_t>(v + a.storageOffset) * a.nDet;
    out[d] = static_cast<float>(sum / a.nSample);
  }
}

// A persistent wave per block takes items (view, 64 detectors) from a
// counter until none is left: the rays' work differs widely (a ray through
// the phantom's middle meets many elements), and taking items as waves free
// up keeps every CU busy to the end.
__device__ __forceinline__ void raysum_items(const RaysumArgs& a, WaveQueue* q) {
  const int blocks = (a.nDet + 63) / 64, nItems = blocks * a.nView;
  const int nChunks = (a.nElem + 63) / 64;
  const bool cull = nChunks <= kCandChunks;
  __shared__ unsigned long long masks[kCandChunks];  // the wave's candidate masks
  while (true) {
    int item = 0;
    if (threadIdx.x == 0) item = atomicAdd(a.counter, 1);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= nItems) break;
    const int v = item / blocks, d0 = (item - v * blocks) * 64;
    const int d = d0 + static_cast<int>(threadIdx.x), d1 = min(d0 + 63, a.nDet - 1);
    if (a.unitPulse) {  // scanner.cpp:475-478
      if (d < a.nDet) a.sino[static_cast<size_t>(v + a.storageOffset) * a.nDet + d] = d == a.nDet / 2 ? 1.f : 0.f;
      continue;
    }
    const ViewEnds e = a.ends[v];
    const ViewSetup s = view_setup(e, a.nDet, a.nSample, a.geom, a.angDetInc);
    if (cull) wave_candidates(a, s, v, d0, d1, masks, nChunks);
    raysum_ray(a, s, v, d, d < a.nDet, cull ? masks : nullptr, nChunks, q);
  }
}

// The FP64 path at 4 waves per SIMD: the clip chain's registers are capped
// at 128, the few it then keeps in scratch cost less than the latency a
// fourth wave hides.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void raysum_f64_kernel(RaysumArgs a) {
  __shared__ WaveQueue q;
  raysum_items(a, &q);
}

// The closed-form chords, a few dozen operations each, one lane per ray over
// every element: culling costs them more than it saves.
__global__ __launch_bounds__(256) void raysum_closed_kernel(RaysumArgs a) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (d >= a.nDet) return;
  float* out = a.sino + static_cast<size_t>(v + a.storageOffset) * a.nDet;
  if (a.unitPulse) {  // scanner.cpp:475-478
    out[d] = d == a.nDet / 2 ? 1.f : 0.f;
    return;
  }
  const ViewSetup s = view_setup(a.ends[v], a.nDet, a.nSample, a.geom, a.angDetInc);
  const double4 w = a.walk[static_cast<size_t>(v) * a.nDet + d];
  const double xs = w.x, ys = w.y;
  double xd = w.z, yd = w.w, ang = w.z;
  double sum = 0.0;
  for (unsigned i = 0; i < a.nSample; i++) {
    if (a.geom == GEOM_EQUIANGULAR) {
      double sa, ca;
      crm_sincos(ang, &sa, &ca);
      xd = a.centerDetLength * ca;
      yd = a.centerDetLength * sa;
    }
    const double segLen = sqrt((xs - xd) * (xs - xd) + (ys - yd) * (ys - yd));
    double rsum = 0.0;
    for (int k = a.nElem - 1; k >= 0; k--) rsum += chord_closed(a.elems[k], xd, yd, xs, ys, segLen);
    sum += rsum;
    if (a.geom == GEOM_EQUIANGULAR) ang += s.aSampInc;
    else { xd += s.ddx2; yd += s.ddy2; }
  }
  out[d] = static_cast<float>(sum / a.nSample);
}

}  // namespace

int raysum_launch(ctsim_ctx* ctx, const ScannerPlan& s, PhantomPlan& p, int precision,
                  int iStartView, int iNumViews, int iOffsetView, int iStorageOffset,
                  float* d_sino, double* h_angles) {
  if (iNumViews <= 0) return CTSIM_OK;
  const hipStream_t st = ctx->stream;
  if (p.dirty || !p.d_elems.ptr) {
    CTSIM_CHECK(p.d_elems.reserve(p.elems.size() * sizeof(ElemGPU)));
    CTSIM_HIP(hipMemcpy(p.d_elems.ptr, p.elems.data(), p.elems.size() * sizeof(ElemGPU),
                        hipMemcpyHostToDevice));
    p.dirty = false;
  }
  const size_t eb = static_cast<size_t>(iNumViews) * sizeof(ViewEnds);
  CTSIM_CHECK(ctx->stage.reserve(eb, st));
  ViewEnds* ends = static_cast<ViewEnds*>(ctx->stage.ptr);
  scanner_view_ends(s, iStartView, iNumViews, iOffsetView, ends);
  if (h_angles)
    for (int v = 0; v < iNumViews; v++) h_angles[v + iStorageOffset] = ends[v].viewAngle;
  CTSIM_CHECK(ctx->stage.upload(ctx->viewConsts, eb, st));
  // the walk, then the ray-sum kernel's work counter
  const size_t walkBytes = static_cast<size_t>(iNumViews) * s.nDet * sizeof(double4);
  CTSIM_CHECK(ctx->scratch.reserve(walkBytes + 256));
  int* counter = reinterpret_cast<int*>(static_cast<unsigned char*>(ctx->scratch.ptr) + walkBytes);
  WalkArgs wa{ctx->viewConsts.as<ViewEnds>(), ctx->scratch.as<double4>(), counter, iNumViews, s.nDet,
              s.geom, s.nSample, s.angDetInc, s.angDetLen};
  hipLaunchKernelGGL(raysum_walk_kernel, dim3((iNumViews + kWalkViews - 1) / kWalkViews), dim3(64), 0, st, wa);
  CTSIM_HIP(hipGetLastError());
  RaysumArgs ra{ctx->viewConsts.as<ViewEnds>(), ctx->scratch.as<double4>(), p.d_elems.as<ElemGPU>(),
                d_sino, counter, iNumViews, s.nDet, static_cast<int>(p.elems.size()), s.geom, iStorageOffset,
                s.nSample, s.angDetInc, s.centerDetLength, p.unitPulse ? 1 : 0};
  if (precision == CTSIM_PREC_F64) {
    // persistent single-wave blocks, as many as are resident at once, no
    // more than the items
    const int items = ((s.nDet + 63) / 64) * iNumViews;
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(raysum_f64_kernel), 64,
                                                     0) != hipSuccess ||
        perCU < 1)
      perCU = 8;
    hipLaunchKernelGGL(raysum_f64_kernel, dim3(std::min(items, ctx->numCUs * perCU)), dim3(64), 0, st, ra);
  } else {
    hipLaunchKernelGGL(raysum_closed_kernel, dim3((s.nDet + 255) / 256, iNumViews), dim3(256), 0, st, ra);
  }
  CTSIM_HIP(hipGetLastError());
  return CTSIM_OK;
}

}  // namespace ctsim

using namespace ctsim;

extern "C" {

int ctsim_phantom_create(ctsim_phantom** out, const char* builtinName) {
  if (!out) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_create: NULL argument");
  *out = nullptr;
  auto* p = new ctsim_phantom();
  if (builtinName && builtinName[0]) {
    const int rc = phantom_builtin(p->plan, builtinName);
    if (rc) { delete p; return rc; }
  }
  *out = p;
  return CTSIM_OK;
}

void ctsim_phantom_destroy(ctsim_phantom* p) { delete p; }

int ctsim_phantom_add_element(ctsim_phantom* p, const char* type, double cx, double cy, double u,
                              double v, double rotDeg, double atten) {
  if (!p) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_add_element: NULL phantom");
  return phantom_add(p->plan, type, cx, cy, u, v, rotDeg, atten);
}

int ctsim_phantom_extent(const ctsim_phantom* p, double* ext4) {
  if (!p || !ext4) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_extent: NULL argument");
  ext4[0] = p->plan.xmin; ext4[1] = p->plan.xmax; ext4[2] = p->plan.ymin; ext4[3] = p->plan.ymax;
  return static_cast<int>(p->plan.elems.size());
}

int ctsim_phantom_element(const ctsim_phantom* p, int index, double* out13) {
  if (!p || !out13) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_element: NULL argument");
  const int n = static_cast<int>(p->plan.elems.size());
  if (index < 0 || index >= n) return fail(CTSIM_ERR_INVALID, "ctsim_phantom_element: index out of range");
  const int k = n - 1 - index;  // the reference's list is newest-first
  const ElemInfo& in = p->plan.infos[k];
  const ElemGPU& e = p->plan.elems[k];
  const double v[13] = {static_cast<double>(in.type), in.cx, in.cy, in.u, in.v, in.rotRad, in.atten,
                        e.lim[0], e.lim[1], e.lim[2], e.lim[3], static_cast<double>(in.nOutline),
                        in.rotDeg};
  for (int i = 0; i < 13; i++) out13[i] = v[i];
  return CTSIM_OK;
}

int ctsim_scanner_create(ctsim_scanner** out, const ctsim_phantom* phm, const char* geometry,
                         int nDet, int nView, int offsetView, int nSample, double rotLen,
                         double focalLengthRatio, double centerDetectorRatio, double viewRatio,
                         double scanRatio) {
  if (!out || !phm) return fail(CTSIM_ERR_INVALID, "ctsim_scanner_create: NULL argument");
  *out = nullptr;
  auto* s = new ctsim_scanner();
  const int rc = scanner_init(s->plan, phm->plan, geometry, nDet, nView, offsetView, nSample, rotLen,
                              focalLengthRatio, centerDetectorRatio, viewRatio, scanRatio);
  if (rc) { delete s; return rc; }
  *out = s;
  return CTSIM_OK;
}

void ctsim_scanner_destroy(ctsim_scanner* s) { delete s; }

int ctsim_scanner_info(const ctsim_scanner* sc, double* o) {
  if (!sc || !o) return fail(CTSIM_ERR_INVALID, "ctsim_scanner_info: NULL argument");
  const ScannerPlan& s = sc->plan;
  o[0] = s.detInc; o[1] = s.detStart; o[2] = s.rotInc; o[3] = s.viewDiameter;
  o[4] = s.focalLength; o[5] = s.sdl; o[6] = s.fanBeamAngle; o[7] = s.viewDiameter / kSqrt2;
  o[8] = s.detLen; o[9] = s.centerDetLength; o[10] = s.scanDiameter; o[11] = s.geom;
  return CTSIM_OK;
}

int ctsim_collect_projections(ctsim_ctx* ctx, const ctsim_scanner* sc, ctsim_phantom* phm,
                              int precision, int iStartView, int iNumViews, int iOffsetView,
                              int iStorageOffset, float* d_sino, double* h_angles) {
  if (!ctx || !sc || !phm || (iNumViews > 0 && !d_sino))
    return fail(CTSIM_ERR_INVALID, "ctsim_collect_projections: NULL argument");
  if (precision != CTSIM_PREC_F64 && precision != CTSIM_PREC_MIXED)
    return fail(CTSIM_ERR_INVALID, "ctsim_collect_projections: unknown precision");
  if (iNumViews < 0 || iStartView < 0 || iStorageOffset < 0)
    return fail(CTSIM_ERR_INVALID, "ctsim_collect_projections: negative view range");
  CTSIM_CHECK(ctx->activate());
  return raysum_launch(ctx, sc->plan, phm->plan, precision, iStartView, iNumViews, iOffsetView,
                       iStorageOffset, d_sino, h_angles);
}

}  // extern "C"
