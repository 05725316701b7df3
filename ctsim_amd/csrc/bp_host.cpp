// Host side of the backprojector plan.  Compiled with g++ -O2 -ffp-contract=off
// so every FP64 table comes out bit-identical to the reference's own.
#include <strings.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>
#include <cstring>

#include "bp_plan.h"

namespace ctsim {

namespace {
const char* const kBackprojNames[] = {"trig", "table", "diff", "idiff"};   // backprojectors.cpp:34-40
const char* const kInterpNames[] = {"nearest", "linear", "cubic"};        // backprojectors.cpp:64-76

int lookup(const char* name, const char* const* table, int n) {
  if (!name) return -1;
  for (int i = 0; i < n; i++)
    if (strcasecmp(name, table[i]) == 0) return i;
  return -1;
}

// nearest<long>, include/ctsupport.h:192-194
long nearest_long(double x) { return x > 0 ? static_cast<long>(x + 0.5) : static_cast<long>(x - 0.5); }

// g(z) = atan(sqrt z) / sqrt z, so that atan(t) = t g(t^2).
double atan_ratio(double z) {
  if (z < 1e-12) return 1.0 - z / 3.0;
  const double s = sqrt(z);
  return atan(s) / s;
}

// Polynomial of degree D in z interpolating g at the D+1 Chebyshev nodes of
// [0, zmax] (near-minimax), by Gaussian elimination on the Vandermonde system.
void cheb_fit(double zmax, int D, double* c) {
  double A[9][10];
  for (int j = 0; j <= D; j++) {
    const double z = 0.5 * zmax * (1 - cos(kPi * (j + 0.5) / (D + 1)));
    double zp = 1;
    for (int k = 0; k <= D; k++, zp *= z) A[j][k] = zp;
    A[j][D + 1] = atan_ratio(z);
  }
  for (int col = 0; col <= D; col++) {
    int piv = col;
    for (int r = col + 1; r <= D; r++)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    for (int k = 0; k <= D + 1; k++) std::swap(A[col][k], A[piv][k]);
    for (int r = 0; r <= D; r++) {
      if (r == col) continue;
      const double f = A[r][col] / A[col][col];
      for (int k = col; k <= D + 1; k++) A[r][k] -= f * A[col][k];
    }
  }
  for (int k = 0; k <= D; k++) c[k] = A[k][D + 1] / A[k][k];
}

// Lowest degree whose relative error on [0, zmax] is below 3e-8 (a position
// error under one FP32 ulp of a detector index); coefficients carry the
// 1/detInc scale so the kernel forms the detector offset as t * P(t^2).
void fit_atan(BpPlan& p, double zmax) {
  for (int D = 4; D <= 8; D += 2) {
    double c[9];
    cheb_fit(zmax, D, c);
    double err = 0;
    for (int i = 0; i <= 4000; i++) {
      const double z = zmax * i / 4000.0;
      double v = c[D];
      for (int k = D - 1; k >= 0; k--) v = v * z + c[k];
      err = fmax(err, fabs(v / atan_ratio(z) - 1));
    }
    if (err <= 3e-8) {
      p.atanDeg = D;
      for (int k = 0; k <= D; k++) p.atanCoef[k] = static_cast<float>(c[k] / p.detInc);
      return;
    }
  }
  p.atanDeg = 0;
}
}  // namespace

// Backprojector::initBackprojector + Backproject ctor + BackprojectTable and
// BackprojectDiff ctors, backprojectors.cpp:135-180, 264-319, 413-429, 491-502.
int bp_plan_init(BpPlan& p, int geometry, int nDet, double detInc, int nView,
                 double viewDiameter, double F, double sdl, const char* backprojName,
                 const char* interpName, int interpFactor, int nx, int ny, const double* roi) {
  p.family = lookup(backprojName, kBackprojNames, 4);
  if (p.family < 0)
    return fail(CTSIM_ERR_INVALID,
                std::string("Invalid backprojection name ") + (backprojName ? backprojName : "(null)"));
  p.interp = lookup(interpName, kInterpNames, 3);
  if (p.interp < 0)
    return fail(CTSIM_ERR_INVALID,
                std::string("Invalid interpolation name ") + (interpName ? interpName : "(null)"));
  if (geometry == GEOM_EQUILINEAR) p.kind = KIND_EQUILINEAR;
  else if (geometry == GEOM_EQUIANGULAR) p.kind = KIND_EQUIANGULAR;
  else if (geometry == GEOM_PARALLEL) p.kind = p.family;
  else
    return fail(CTSIM_ERR_INVALID,
                "Unable to select a backprojection method [Backprojector::initBackprojector]");
  if (nDet < 2 || nView < 1 || nx < 1 || ny < 1)
    return fail(CTSIM_ERR_INVALID, "backprojector: nDet >= 2, nView >= 1 and nx, ny >= 1 required");
  if (interpFactor != 1)
    return fail(CTSIM_ERR_UNSUPPORTED,
                "backprojector: interpolation factor other than 1 needs frequency preinterpolation, "
                "which the reference compiles out (backprojectors.cpp:64-76)");
  p.geom = geometry;
  p.interpFactor = interpFactor;
  p.detInc = detInc;
  p.nDet = nDet;
  p.nView = nView;
  p.iDetCenter = (nDet - 1) / 2;
  p.rotScale = geometry == GEOM_PARALLEL ? kPi / nView : (2 * kPi) / nView;
  p.nx = nx;
  p.ny = ny;
  p.viewDiameter = viewDiameter;
  double phmLen = viewDiameter / kSqrt2;  // include/projections.h:165
  p.xMin = -phmLen / 2;
  p.xMax = p.xMin + phmLen;
  p.yMin = -phmLen / 2;
  p.yMax = p.yMin + phmLen;
  if (roi) {  // ReconstructionROI member order: xMin, yMin, xMax, yMax
    if (roi[0] > p.xMin) p.xMin = roi[0];
    if (roi[2] < p.xMax) p.xMax = roi[2];
    if (roi[1] > p.yMin) p.yMin = roi[1];
    if (roi[3] < p.yMax) p.yMax = roi[3];
    if (p.xMin > p.xMax) std::swap(p.xMin, p.xMax);
    if (p.yMin > p.yMax) std::swap(p.yMin, p.yMax);
  }
  p.xInc = (p.xMax - p.xMin) / nx;
  p.yInc = (p.yMax - p.yMin) / ny;
  p.F = F;
  p.sdl = sdl;

  // Pixel-centre tables by repeated addition (backprojectors.cpp:374-377, 424-428).
  p.xs.resize(nx);
  p.ys.resize(ny);
  {
    double x = p.xMin + p.xInc / 2;
    for (int ix = 0; ix < nx; x += p.xInc, ix++) p.xs[ix] = static_cast<float>(x);
    double y = p.yMin + p.yInc / 2;
    for (int iy = 0; iy < ny; y += p.yInc, iy++) p.ys[iy] = static_cast<float>(y);
  }
  if (p.kind == KIND_TRIG || p.kind == KIND_TABLE || p.kind == KIND_EQUIANGULAR ||
      p.kind == KIND_EQUILINEAR) {
    // BackprojectTable ctor; BackprojectTrig evaluates the same r and phi per
    // view from the same repeated-addition x, y, so it shares the table.
    p.r.resize(static_cast<size_t>(nx) * ny);
    p.phi.resize(static_cast<size_t>(nx) * ny);
    double x, y;
    int ix, iy;
    for (x = p.xMin + p.xInc / 2, ix = 0; ix < nx; x += p.xInc, ix++)
      for (y = p.yMin + p.yInc / 2, iy = 0; iy < ny; y += p.yInc, iy++) {
        p.r[static_cast<size_t>(ix) * ny + iy] = sqrt(x * x + y * y);
        p.phi[static_cast<size_t>(ix) * ny + iy] = atan2(y, x);
      }
  }
  if (p.kind == KIND_DIFF || p.kind == KIND_IDIFF) {
    double x = p.xMin + p.xInc / 2;
    double y = p.yMin + p.yInc / 2;
    p.start_r = sqrt(x * x + y * y);
    p.start_phi = atan2(y, x);
  }
  p.windowF64 = bp_window_bound(p, p.tileX, p.tileY);
  p.windowF32 = p.windowF64;
  {
    // tan(gamma) = r cos / (F + r sin) stays within [-1, 1] when every image
    // point is closer to the origin than F / sqrt(2) (with a margin)
    const double rx = fmax(fabs(p.xMin), fabs(p.xMax)), ry = fmax(fabs(p.yMin), fabs(p.yMax));
    const double R = sqrt(rx * rx + ry * ry);
    p.smallTan = p.F > 0 && R * 1.4143 * 1.001 < p.F;
    if (p.geom == GEOM_EQUIANGULAR && p.F > R) {
      // every pixel sees the source within |gamma| <= asin(R / F)
      const double tmax = 1.001 * R / sqrt(p.F * p.F - R * R);
      if (tmax <= 1.0) fit_atan(p, tmax * tmax);
    }
  }
  {
    // the quarter turn (x, y) -> (-y, x) maps pixel (i, j) to (n-1-j, i)
    // when ys == xs and xs[i] == -xs[n-1-i] (to a small fraction of a pixel)
    bool sym = nx == ny && nx % 2 == 0 && nx >= 2;
    const double tol = 1e-4 * fabs(p.xInc);
    for (int i = 0; sym && i < nx; i++)
      sym = fabs(static_cast<double>(p.xs[i]) + p.xs[nx - 1 - i]) <= tol &&
            fabs(static_cast<double>(p.ys[i]) - p.xs[i]) <= tol;
    p.symImage = sym;
    bool mx = true, my = true;
    for (int i = 0; mx && i < nx; i++) mx = fabs(static_cast<double>(p.xs[i]) + p.xs[nx - 1 - i]) <= 1e-4 * fabs(p.xInc);
    for (int i = 0; my && i < ny; i++) my = fabs(static_cast<double>(p.ys[i]) + p.ys[ny - 1 - i]) <= 1e-4 * fabs(p.yInc);
    p.mirX = mx;
    p.mirY = my;
  }
  p.mixed = bp_default_shape(p, 256);
  return CTSIM_OK;
}

// Views v + s*Q turned by s quarter turns (S = 4, Q = nView / 4) or by s
// quarter or half turns (S = 2), to 1e-9 rad for every view.
SymViews bp_sym_views(const double* angles, int nView) {
  SymViews out;
  const double quarter = M_PI / 2, tol = 1e-9;
  auto off = [](double d, double target) { return fabs(remainder(d - target, 2 * M_PI)); };
  for (int S : {4, 2}) {
    if (nView < S || nView % S != 0) continue;
    const int Q = nView / S;
    int k = 0;
    for (int kk = 1; kk <= 3; kk++)
      if (off(angles[Q] - angles[0], kk * quarter) < tol) k = kk;
    if (k == 0 || (S == 4 && k == 2)) continue;
    bool ok = true;
    for (int v = 0; v < Q && ok; v++)
      for (int s = 1; s < S && ok; s++) ok = off(angles[v + s * Q] - angles[v], s * k * quarter) < tol;
    if (!ok) continue;
    out.S = S;
    out.K = k == 3 ? 1 : k;
    out.dir = k == 3 ? -1.f : 1.f;
    return out;
  }
  return out;
}

// Per-view constants of the FP64 path; the same expressions as
// BackprojectDiff::BackprojectView (backprojectors.cpp:523-528) and
// BackprojectIntDiff::BackprojectView (593-597).
void bp_view_consts_f64(const BpPlan& p, const double* angles, int nView, ViewConstF64* out) {
  const long scale = 1L << 32;
  for (int v = 0; v < nView; v++) {
    double theta = angles[v];
    ViewConstF64& c = out[v];
    c.theta = theta;
    c.a = c.b = c.c = 0;
    if (p.kind == KIND_DIFF) {
      c.a = p.xInc * cos(theta) / p.detInc;
      c.b = p.yInc * sin(theta) / p.detInc;
      c.c = p.iDetCenter + p.start_r * cos(theta - p.start_phi) / p.detInc;
    } else if (p.kind == KIND_IDIFF) {
      long dx = nearest_long(p.xInc * cos(theta) / p.detInc * scale);
      long dy = nearest_long(p.yInc * sin(theta) / p.detInc * scale);
      long cs = nearest_long((p.start_r * cos(theta - p.start_phi) / p.detInc + p.iDetCenter) * scale);
      memcpy(&c.a, &dx, 8);
      memcpy(&c.b, &dy, 8);
      memcpy(&c.c, &cs, 8);
    }
  }
}

namespace {

// Partners of every view under one pixel mirror (0: Y, 1: X, 2: O), as
// view * 2 + reflected; false when some view has none or two views share
// one (angles equal to 1e-9 rad modulo 2 pi).
bool mirror_partners(const BpPlan& p, const double* angles, int nView, int type, std::vector<int>& out) {
  const double twoPi = 2 * kPi, tol = 1e-9;
  auto wrap = [&](double a) {
    a = fmod(a, twoPi);
    return a < 0 ? a + twoPi : a;
  };
  std::vector<std::pair<double, int>> sorted(nView);
  for (int v = 0; v < nView; v++) sorted[v] = {wrap(angles[v]), v};
  std::sort(sorted.begin(), sorted.end());
  auto find = [&](double target) -> int {
    const double t = wrap(target);
    auto it = std::lower_bound(sorted.begin(), sorted.end(), std::make_pair(t - tol, -1));
    if (it != sorted.end() && fabs(it->first - t) <= tol) return it->second;
    if (t < tol && !sorted.empty() && fabs(sorted.back().first - twoPi - t) <= tol) return sorted.back().second;
    if (t > twoPi - tol && !sorted.empty() && fabs(sorted.front().first + twoPi - t) <= tol)
      return sorted.front().second;
    return -1;
  };
  const bool par = p.geom == GEOM_PARALLEL;
  out.assign(nView, -1);
  std::vector<char> hit(nView, 0);
  for (int v = 0; v < nView; v++) {
    const double th = angles[v];
    // the reflecting target, then (parallel) the one at the same position
    const double refl = type == 0 ? -th : (type == 1 ? kPi - th : (par ? th : th + kPi));
    const double same = type == 0 ? kPi - th : (type == 1 ? -th : th + kPi);
    int w = -1, r = 0;
    if (!par && type == 2) {
      w = find(same);
    } else {
      w = find(refl);
      r = 1;
      if (w < 0 && par) {
        w = find(same);
        r = 0;
      }
    }
    if (w < 0 || hit[w]) return false;
    hit[w] = 1;
    out[v] = 2 * w + r;
  }
  return true;
}

}  // namespace

// Per-view constants of the mixed path: FP64 trig on the host, rounded once,
// and the views' mirror partners.
MirrorViews bp_view_consts_f32(const BpPlan& p, const double* angles, int nView, ViewConstF32* out) {
  for (int v = 0; v < nView; v++) {
    double c = cos(angles[v]), s = sin(angles[v]);
    if (p.geom == GEOM_PARALLEL) {
      out[v].ca = static_cast<float>(c / p.detInc);
      out[v].sa = static_cast<float>(s / p.detInc);
    } else {
      out[v].ca = static_cast<float>(c);
      out[v].sa = static_cast<float>(s);
    }
    out[v].mir[0] = out[v].mir[1] = out[v].mir[2] = 0;
    out[v].order = v;
    out[v].pad[0] = out[v].pad[1] = 0.f;
  }
  MirrorViews mv;
  // the mirror body reflects about the centre cell, which needs 2 * centre
  // to be a whole cell (it is: the centre is (nDet - 1) / 2 rounded down)
  std::vector<int> y, x, o;
  const bool hasY = p.mirX && mirror_partners(p, angles, nView, 0, y);
  const bool hasX = p.mirY && mirror_partners(p, angles, nView, 1, x);
  const bool hasO = p.mirX && p.mirY && mirror_partners(p, angles, nView, 2, o);
  if (hasY && hasX && hasO) {
    mv.NM = 4;
    for (int v = 0; v < nView; v++) {
      out[v].mir[0] = y[v];
      out[v].mir[1] = x[v];
      out[v].mir[2] = o[v];
    }
  } else if (hasY || hasO) {
    mv.NM = 2;
    mv.mapO = hasY ? 0 : 1;
    for (int v = 0; v < nView; v++) out[v].mir[0] = hasY ? y[v] : o[v];
  }
  if (mv.NM > 1) {  // orbits contiguous: each view followed by its partners
    std::vector<char> seen(nView, 0);
    int s = 0;
    for (int v = 0; v < nView; v++) {
      if (seen[v]) continue;
      int orbit[4] = {v, out[v].mir[0] >> 1, out[v].mir[1] >> 1, out[v].mir[2] >> 1};
      for (int m = 0; m < mv.NM; m++)
        if (!seen[orbit[m]]) {
          seen[orbit[m]] = 1;
          out[s++].order = orbit[m];
        }
    }
  }
  return mv;
}

// Upper bound on the detector positions one tile of tileX x tileY pixels can
// touch in one view, plus the interpolation taps and a rounding margin.
// Returns 0 when a full detector row must be staged instead.
int bp_window_bound(const BpPlan& p, int tileX, int tileY) {
  const double ax = (tileX - 1) * p.xInc, ay = (tileY - 1) * p.yInc;
  const double diag = sqrt(ax * ax + ay * ay);
  double span;
  if (p.geom == GEOM_PARALLEL) {
    span = diag / p.detInc;
  } else {
    double rx = fmax(fabs(p.xMin), fabs(p.xMax)), ry = fmax(fabs(p.yMin), fabs(p.yMax));
    double R = sqrt(rx * rx + ry * ry);
    if (!(p.F > 1.05 * R)) return 0;  // source near or inside the image: no monotone window
    double D = p.F - R;
    double ang = 2 * asin(fmin(1.0, diag / (2 * D)));
    if (p.geom == GEOM_EQUIANGULAR) {
      span = ang / p.detInc;
    } else {
      double s = R / p.F;
      span = p.sdl * ang / ((1 - s * s) * p.detInc);
    }
  }
  double w = ceil(span) + 10;
  if (!(w < p.nDet + 8)) return 0;
  return static_cast<int>(w);
}

// Default mixed launch shape: 16 x 64 pixel tiles (4 rows per lane); eight
// view groups (one per XCD) when there are enough views, so each XCD's L2
// holds its views' coefficients; chunks sized for about four workgroups'
// staging per CU in 160 KB of LDS.
MixedShape bp_default_shape(const BpPlan& p, int numCUs) {
  (void)numCUs;
  MixedShape m;
  m.lanesY = 16;
  m.P = p.ny >= 64 ? 4 : (p.ny >= 32 ? 2 : 1);
  const int tx = 256 / m.lanesY, ty = m.lanesY * m.P;
  m.split = p.nView >= 8 * 32 ? 8 : 1;
  m.W = bp_window_bound(p, tx, ty);
  const int Wv = m.W > 0 ? m.W : p.nDet + 4;
  const int elem = p.interp == INTERP_CUBIC ? 16 : 8;
  int chunk = (36 * 1024) / (Wv * elem);
  chunk = chunk > 32 ? 32 : (chunk < 1 ? 1 : chunk);
  m.chunk = chunk;
  return m;
}

}  // namespace ctsim
