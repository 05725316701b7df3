/* CPU oracle: plain-C restatement of the CTSim hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ctsim_oracle.h).  Each function names the
 * reference lines it restates.  Arithmetic is written in the reference's
 * evaluation order, in double precision with float storage where the
 * reference stores floats, and is compiled without FP contraction
 * (oracle/Makefile) so the results are bit-identical to the reference built
 * from source.
 */
#define _GNU_SOURCE
#include "ctsim_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>

/* include/ctsupport.h:167-177 */
#define PI 3.14159265358979323846
#define HALFPI 1.57079632679489661923
#define TWOPI 6.28318530717958647692
#define SQRT2 1.414213562373095049
#define F_EPSILON 1.0E-6
#define D_EPSILON 1.0E-10

static char g_err[256];
const char* orc_last_error(void) { return g_err; }
static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------------------------------------------------------------- helpers */

/* nearest<T>, include/ctsupport.h:192-194 (half away from zero, truncating cast) */
static int nearest_int(double x) { return x > 0 ? (int)(x + 0.5) : (int)(x - 0.5); }
static long nearest_long(double x) { return x > 0 ? (long)(x + 0.5) : (long)(x - 0.5); }

/* normalizeAngle, libctsupport/mathfuncs.cpp:72-81 */
static double normalize_angle(double t) {
  while (t < 0.) t += TWOPI;
  while (t >= TWOPI) t -= TWOPI;
  return t;
}

/* integrateSimpson, libctsupport/mathfuncs.cpp:36-60 */
double orc_integrate_simpson(double xmin, double xmax, const double* y, int np) {
  if (np < 2) return 0.;
  if (np == 2) return (xmax - xmin) * (y[0] + y[1]) / 2;
  double area = 0;
  int nDiv = (np - 1) / 2;
  double width = (xmax - xmin) / (double)(np - 1);
  for (int i = 1; i <= nDiv; i++) {
    int xr = 2 * i, xl = xr - 2, xm = xr - 1;
    area += (width / 3.0) * (y[xl] + 4.0 * y[xm] + y[xr]);
  }
  if ((np & 1) == 0) area += width * (y[np - 2] + y[np - 1]) / 2;
  return area;
}

/* 3x3 homogeneous matrices, libctsupport/xform.cpp:99-160 (row-vector form) */
typedef double mtx2[3][3];
static void ident_m(mtx2 m) {
  m[0][0] = 1; m[0][1] = 0; m[0][2] = 0;
  m[1][0] = 0; m[1][1] = 1; m[1][2] = 0;
  m[2][0] = 0; m[2][1] = 0; m[2][2] = 1;
}
static void xlat_m(mtx2 m, double x, double y) { ident_m(m); m[2][0] = x; m[2][1] = y; }
static void scale_m(mtx2 m, double sx, double sy) { ident_m(m); m[0][0] = sx; m[1][1] = sy; }
static void rot_m(mtx2 m, double t) {
  double c = cos(t), s = sin(t);
  ident_m(m);
  m[0][0] = c; m[0][1] = s;
  m[1][0] = -s; m[1][1] = c;
}
static void mult_m(const mtx2 a, const mtx2 b, mtx2 r) {
  mtx2 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      t[i][j] = 0;
      for (int k = 0; k < 3; k++) t[i][j] += a[i][k] * b[k][j];
    }
  memcpy(r, t, sizeof(mtx2));
}
static void xform_m(const mtx2 m, double* x, double* y) {
  double xt = *x * m[0][0] + *y * m[1][0] + m[2][0];
  double yt = *x * m[0][1] + *y * m[1][1] + m[2][1];
  *x = xt;
  *y = yt;
}

/* ------------------------------------------------------------- clipping */

/* rectcode + clip_rect, libctsupport/clip.cpp:332-395 */
static int rectcode(double x, double y, const double r[4]) {
  int c = 0;
  if (x < r[0]) c = 1;
  else if (x > r[2]) c = 2;
  if (y < r[1]) c += 4;
  else if (y > r[3]) c += 8;
  return c;
}
static int clip_rect(double* x1, double* y1, double* x2, double* y2, const double r[4]) {
  double x = 0, y = 0;
  int c1 = rectcode(*x1, *y1, r), c2 = rectcode(*x2, *y2, r);
  while (c1 || c2) {
    if (c1 & c2) return 0;
    int c = c1 ? c1 : c2;
    if (c & 1) { y = *y1 + (*y2 - *y1) * (r[0] - *x1) / (*x2 - *x1); x = r[0]; }
    else if (c & 2) { y = *y1 + (*y2 - *y1) * (r[2] - *x1) / (*x2 - *x1); x = r[2]; }
    else if (c & 4) { x = *x1 + (*x2 - *x1) * (r[1] - *y1) / (*y2 - *y1); y = r[1]; }
    else if (c & 8) { x = *x1 + (*x2 - *x1) * (r[3] - *y1) / (*y2 - *y1); y = r[3]; }
    if (c == c1) { *x1 = x; *y1 = y; c1 = rectcode(*x1, *y1, r); }
    else { *x2 = x; *y2 = y; c2 = rectcode(*x2, *y2, r); }
  }
  return 1;
}

/* clip_circle, libctsupport/clip.cpp:126-198 */
static int clip_circle(double* x1, double* y1, double* x2, double* y2, double cx, double cy,
                       double radius, double t1, double t2) {
  double xc1 = *x1, yc1 = *y1, xc2 = *x2, yc2 = *y2, ccx = cx, ccy = cy;
  double xtrans = -xc1, ytrans = -yc1;
  xc1 += xtrans; yc1 += ytrans; xc2 += xtrans; yc2 += ytrans; ccx += xtrans; ccy += ytrans;
  double theta = -atan2(yc2, xc2);
  mtx2 rm;
  rot_m(rm, theta);
  xform_m(rm, &xc2, &yc2);
  xform_m(rm, &ccx, &ccy);
  t1 += theta;
  t2 += theta;
  t1 = normalize_angle(t1);
  t2 = normalize_angle(t2);
  if (xc2 < -D_EPSILON || fabs(yc2) > F_EPSILON) return 0; /* reference: ERR_SEVERE + reject */
  if (fabs(ccy) > radius) return 0;
  double temp = sqrt(radius * radius - ccy * ccy);
  double xcmin = ccx - temp, xcmax = ccx + temp;
  if (fabs(t2 - t1) < D_EPSILON) {
    if (xc1 < xcmin) xc1 = xcmin;
    if (xc2 > xcmax) xc2 = xcmax;
  } else if (t1 < t2) {
    if (t1 < PI && t2 > PI)
      if (xc1 < xcmin) xc1 = xcmin;
  } else if (t1 > t2) {
    if (t1 < PI)
      if (xc1 < xcmin) xc1 = xcmin;
    if (xc2 > xcmax) xc2 = xcmax;
  }
  rot_m(rm, -theta);
  xform_m(rm, &xc1, &yc1);
  xform_m(rm, &xc2, &yc2);
  xc1 += -xtrans; yc1 += -ytrans; xc2 += -xtrans; yc2 += -ytrans;
  *x1 = xc1; *y1 = yc1; *x2 = xc2; *y2 = yc2;
  return 1;
}

/* tcode + clip_triangle, libctsupport/clip.cpp:243-316 */
static int tcode(double x, double y, double m, double b, int clip_xaxis) {
  int c = 0;
  if (clip_xaxis && y < 0.) c = 1;
  if (y > -m * x + b + D_EPSILON) c += 2;
  if (y > m * x + b + D_EPSILON) c += 4;
  return c;
}
static int clip_triangle(double* x1, double* y1, double* x2, double* y2, double u, double v,
                         int clip_xaxis) {
  double m = v / u, b = v;
  int c1 = tcode(*x1, *y1, m, b, clip_xaxis), c2 = tcode(*x2, *y2, m, b, clip_xaxis);
  while (c1 || c2) {
    if (c1 & c2) return 0;
    int c = c1 ? c1 : c2;
    double x = 0, y = 0;
    if (c & 1) {
      x = *x1 + (*x2 - *x1) * (0.0 - *y1) / (*y2 - *y1);
      y = 0.0;
    } else if (c & 2) {
      double dx = *x2 - *x1, dy = *y2 - *y1;
      if (fabs(dx) > D_EPSILON) x = (-*y1 + b + *x1 * dy / dx) / (m + dy / dx);
      else x = *x1;
      y = -m * x + b;
    } else if (c & 4) {
      double dx = *x2 - *x1, dy = *y2 - *y1;
      if (fabs(dx) > D_EPSILON) {
        x = (*y1 - b - *x1 * dy / dx);
        x /= (m - dy / dx);
      } else
        x = *x1;
      y = m * x + b;
    }
    if (c == c1) { *x1 = x; *y1 = y; c1 = tcode(*x1, *y1, m, b, clip_xaxis); }
    else { *x2 = x; *y2 = y; c2 = tcode(*x2, *y2, m, b, clip_xaxis); }
  }
  return 1;
}

/* clip_segment, libctsupport/clip.cpp:44-77 */
static int clip_segment(double* x1, double* y1, double* x2, double* y2, double u, double v) {
  double xc1 = *x1 * u, yc1 = *y1 * v, xc2 = *x2 * u, yc2 = *y2 * v;
  if (yc1 > 0. && yc2 > 0.) return 0;
  double radius = sqrt(u * u + v * v);
  if (!clip_circle(&xc1, &yc1, &xc2, &yc2, 0.0, v, radius, 0.0, 0.0)) return 0;
  if (yc1 > 0. && yc2 > 0.) return 0;
  if (yc1 > 0.) {
    xc1 = xc1 + (xc2 - xc1) * (0.0 - yc1) / (yc2 - yc1);
    yc1 = 0.0;
  } else if (yc2 > 0.) {
    xc2 = xc1 + (xc2 - xc1) * (0.0 - yc1) / (yc2 - yc1);
    yc2 = 0.0;
  }
  *x1 = xc1 / u; *y1 = yc1 / v; *x2 = xc2 / u; *y2 = yc2 / v;
  return 1;
}

/* clip_sector, libctsupport/clip.cpp:90-111 */
static int clip_sector(double* x1, double* y1, double* x2, double* y2, double u, double v) {
  double xc1 = *x1 * u, yc1 = *y1 * v, xc2 = *x2 * u, yc2 = *y2 * v;
  double radius = sqrt(u * u + v * v);
  if (!clip_circle(&xc1, &yc1, &xc2, &yc2, 0.0, v, radius, 0.0, 0.0)) return 0;
  if (!clip_triangle(&xc1, &yc1, &xc2, &yc2, u, v, 0)) return 0;
  *x1 = xc1 / u; *y1 = yc1 / v; *x2 = xc2 / u; *y2 = yc2 / v;
  return 1;
}

/* -------------------------------------------------------------- phantom */

enum { PE_RECT = 1, PE_TRI = 2, PE_ELLIPSE = 3, PE_SECTOR = 4, PE_SEGMENT = 5 };

typedef struct {
  int type;
  double cx, cy, u, v, rot, atten;
  double lim[4]; /* xmin, ymin, xmax, ymax */
  mtx2 objToPhm, phmToObj;
} elem_t;

typedef struct {
  int n, cap, unitPulse;
  elem_t* e; /* insertion order; the reference list is push_front'ed */
  double xmin, xmax, ymin, ymax;
} phantom_t;

/* numCirclePoints, phantom.cpp:821-827 */
static int num_circle_points(double theta) {
  theta = theta >= TWOPI ? TWOPI : (theta <= 0. ? 0. : theta);
  return (int)(360 * theta / TWOPI + 1.5);
}

/* calcArcPoints, phantom.cpp:771-793 (rotation recurrence) */
static void calc_arc(double* x, double* y, int pts, double xc, double yc, double r, double start,
                     double stop) {
  double th = (stop - start) / (pts - 1);
  double c = cos(th), s = sin(th);
  x[0] = r * cos(start) + xc;
  y[0] = r * sin(start) + yc;
  double xp = x[0] - xc, yp = y[0] - yc;
  for (int i = 1; i < pts; i++) {
    double xr = c * xp - s * yp;
    double yr = s * xp + c * yp;
    x[i] = xr + xc;
    y[i] = yr + yc;
    xp = xr;
    yp = yr;
  }
}

static void minmax(const double* a, int n, double* mn, double* mx) {
  *mx = *mn = a[0];
  for (int i = 1; i < n; i++)
    if (a[i] < *mn) *mn = a[i];
    else if (a[i] > *mx) *mx = a[i];
}

/* PhantomElement ctor, phantom.cpp:554-566; makeTransformMatrices 629-655;
 * makeVectorOutline 669-756 (SCALE_PELEM_EXTENT = 0, phantom.cpp:29) */
static void elem_init(elem_t* e, int type, double cx, double cy, double u, double v, double rotDeg,
                      double atten) {
  e->type = type; e->cx = cx; e->cy = cy; e->u = u; e->v = v; e->atten = atten;
  e->rot = rotDeg * (PI / 180.);
  mtx2 t;
  scale_m(e->objToPhm, u, v);
  rot_m(t, e->rot);
  mult_m(e->objToPhm, t, e->objToPhm);
  xlat_m(t, cx, cy);
  mult_m(e->objToPhm, t, e->objToPhm);
  xlat_m(e->phmToObj, -cx, -cy);
  rot_m(t, -e->rot);
  mult_m(e->phmToObj, t, e->phmToObj);
  scale_m(t, 1 / u, 1 / v);
  mult_m(e->phmToObj, t, e->phmToObj);

  double xs[400] = {0}, ys[400] = {0};
  int n = 0, cpts;
  double radius, theta, start, stop;
  switch (type) {
    case PE_RECT:
      n = 5;
      xs[0] = -u; ys[0] = -v; xs[1] = u; ys[1] = -v; xs[2] = u; ys[2] = v;
      xs[3] = -u; ys[3] = v; xs[4] = -u; ys[4] = -v;
      break;
    case PE_TRI:
      n = 4;
      xs[0] = -u; ys[0] = 0.0; xs[1] = u; ys[1] = 0.0; xs[2] = 0.0; ys[2] = v;
      xs[3] = -u; ys[3] = 0.0;
      break;
    case PE_ELLIPSE:
      n = num_circle_points(TWOPI);
      calc_arc(xs, ys, n, 0.0, 0.0, 1.0, 0.0, TWOPI);
      for (int i = 0; i < n; i++) { xs[i] *= u; ys[i] *= v; }
      break;
    case PE_SECTOR:
      radius = sqrt(u * u + v * v);
      theta = atan(u / v);
      start = 3.0 * HALFPI - theta;
      stop = 3.0 * HALFPI + theta;
      cpts = num_circle_points(stop - start);
      n = 3 + cpts;
      xs[0] = 0.0; ys[0] = v; xs[1] = -u; ys[1] = 0.0;
      calc_arc(xs + 2, ys + 2, cpts, 0.0, v, radius, start, stop);
      xs[cpts + 2] = 0.0; ys[cpts + 2] = v;
      break;
    case PE_SEGMENT:
      radius = sqrt(u * u + v * v);
      theta = atan(u / v);
      start = 3.0 * HALFPI - theta;
      stop = 3.0 * HALFPI + theta;
      cpts = num_circle_points(stop - start);
      n = cpts + 1;
      calc_arc(xs, ys, cpts, 0.0, v, radius, start, stop);
      xs[cpts] = -u; ys[cpts] = 0.0;
      break;
  }
  /* rotate2d / xlat2d, xform.cpp:32-68 */
  double ct = cos(e->rot), st = sin(e->rot);
  for (int i = 0; i < n; i++) {
    double xr = xs[i] * ct - ys[i] * st;
    double yr = xs[i] * st + ys[i] * ct;
    xs[i] = xr;
    ys[i] = yr;
  }
  for (int i = 0; i < n; i++) { xs[i] += cx; ys[i] += cy; }
  double xmn, xmx, ymn, ymx;
  minmax(xs, n, &xmn, &xmx);
  minmax(ys, n, &ymn, &ymx);
  double xf = (xmx - xmn) * 0.0, yf = (ymx - ymn) * 0.0;
  xmn -= xf; ymn -= yf; xmx += xf; ymx += yf;
  e->lim[0] = xmn; e->lim[1] = ymn; e->lim[2] = xmx; e->lim[3] = ymx;
}

static void phm_add(phantom_t* p, int type, double cx, double cy, double u, double v, double rot,
                    double atten) {
  if (type < PE_RECT || type > PE_SEGMENT) return; /* reference: warning + skip (phantom.cpp:240-244) */
  if (p->n == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 16;
    p->e = (elem_t*)realloc(p->e, sizeof(elem_t) * p->cap);
  }
  elem_t* e = &p->e[p->n++];
  elem_init(e, type, cx, cy, u, v, rot, atten);
  /* phantom.cpp:249-253 */
  if (p->xmin > e->lim[0]) p->xmin = e->lim[0];
  if (p->xmax < e->lim[2]) p->xmax = e->lim[2];
  if (p->ymin > e->lim[1]) p->ymin = e->lim[1];
  if (p->ymax < e->lim[3]) p->ymax = e->lim[3];
}

/* Phantom::init, createFromPhantom, addStdSheppLogan, addStdHerman
 * (phantom.cpp:67-77, 128-170, 390-433) */
static int phm_build(phantom_t* p, const int* types, const double* params, int nElem,
                     const char* builtin) {
  memset(p, 0, sizeof *p);
  p->xmin = 1E30; p->xmax = -1E30; p->ymin = 1E30; p->ymax = -1E30;
  if (builtin && builtin[0]) {
    if (!strcasecmp(builtin, "herman")) {
      static const double H[15][6] = {
          {0.000, 1.50, 0.375, 0.3000, 90.00, -0.003},  {0.675, -0.75, 0.225, 0.1500, 140.00, 0.010},
          {0.750, 1.50, 0.375, 0.2250, 50.00, 0.003},   {1.375, -7.50, 1.100, 0.6250, 19.20, -0.204},
          {1.375, -7.50, 1.100, 4.3200, 19.21, 0.204},  {0.000, -2.25, 1.125, 0.3750, 0.00, -0.003},
          {0.000, -2.25, 1.125, 3.0000, 0.00, 0.003},   {-1.000, 3.75, 1.000, 0.5000, 135.00, -0.003},
          {-1.000, 3.75, 1.000, 3.0000, 135.00, 0.003}, {1.000, 3.75, 1.000, 0.5000, 225.00, -0.003},
          {1.000, 3.75, 1.000, 3.0000, 225.00, 0.003},  {5.025, 3.75, 1.125, 0.5000, 110.75, 0.206},
          {-5.025, 3.75, 1.125, 0.9000, -110.75, 0.206}, {0.000, 0.00, 8.625, 6.4687, 90.00, 0.416},
          {0.000, 0.00, 7.875, 5.7187, 90.00, -0.206}};
      static const int HT[15] = {PE_ELLIPSE, PE_ELLIPSE, PE_ELLIPSE, PE_SEGMENT, PE_SEGMENT,
                                 PE_SEGMENT, PE_SEGMENT, PE_SEGMENT, PE_SEGMENT, PE_SEGMENT,
                                 PE_SEGMENT, PE_TRI,     PE_TRI,     PE_ELLIPSE, PE_ELLIPSE};
      for (int i = 0; i < 15; i++)
        phm_add(p, HT[i], H[i][0], H[i][1], H[i][2], H[i][3], H[i][4], H[i][5]);
    } else if (!strcasecmp(builtin, "shepp-logan")) {
      static const double S[11][6] = {
          {0.0000, 0.0000, 0.6900, 0.9200, 0.0, 1.00},    {0.0000, -0.0184, 0.6624, 0.8740, 0.0, -0.98},
          {0.2200, 0.0000, 0.1100, 0.3100, -18.0, -0.02}, {-0.2200, 0.0000, 0.1600, 0.4100, 18.0, -0.02},
          {0.0000, 0.3500, 0.2100, 0.2500, 0.0, 0.01},    {0.0000, 0.1000, 0.0460, 0.0460, 0.0, 0.01},
          {0.0000, -0.1000, 0.0460, 0.0460, 0.0, 0.01},   {-0.0800, -0.6050, 0.0460, 0.0230, 0.0, 0.01},
          {0.0000, -0.6050, 0.0230, 0.0230, 0.0, 0.01},   {0.0600, -0.6050, 0.0230, 0.0230, 0.0, 0.01},
          {0.5538, -0.3858, 0.0330, 0.2060, -18.0, 0.03}};
      for (int i = 0; i < 11; i++)
        phm_add(p, PE_ELLIPSE, S[i][0], S[i][1], S[i][2], S[i][3], S[i][4], S[i][5]);
    } else if (!strcasecmp(builtin, "unit-pulse")) {
      p->unitPulse = 1;
      phm_add(p, PE_RECT, 0., 0., 100., 100., 0., 0.);
      phm_add(p, PE_ELLIPSE, 0., 0., 1., 1., 0., 1.);
    } else {
      return fail(1, "Invalid phantom name");
    }
  }
  for (int i = 0; i < nElem; i++) {
    const double* q = params + 6 * i;
    phm_add(p, types[i], q[0], q[1], q[2], q[3], q[4], q[5]);
  }
  return 0;
}

/* PhantomElement::clipLineWorldCoords / clipLineNormalizedCoords,
 * phantom.cpp:830-898 */
static int elem_clip(const elem_t* e, double* x1, double* y1, double* x2, double* y2) {
  double cx1 = *x1, cy1 = *y1, cx2 = *x2, cy2 = *y2;
  if (!clip_rect(&cx1, &cy1, &cx2, &cy2, e->lim)) return 0;
  xform_m(e->phmToObj, x1, y1);
  xform_m(e->phmToObj, x2, y2);
  int ok = 0;
  static const double unit[4] = {-1.0, -1.0, 1.0, 1.0};
  switch (e->type) {
    case PE_RECT: ok = clip_rect(x1, y1, x2, y2, unit); break;
    case PE_ELLIPSE: ok = clip_circle(x1, y1, x2, y2, 0.0, 0.0, 1.0, 0.0, 0.0); break;
    case PE_TRI: ok = clip_triangle(x1, y1, x2, y2, 1.0, 1.0, 1); break;
    case PE_SEGMENT: ok = clip_segment(x1, y1, x2, y2, e->u, e->v); break;
    case PE_SECTOR: ok = clip_sector(x1, y1, x2, y2, e->u, e->v); break;
  }
  if (!ok) return 0;
  xform_m(e->objToPhm, x1, y1);
  xform_m(e->objToPhm, x2, y2);
  return 1;
}

/* Scanner::projectSingleLine + projectLineAgainstPElem, scanner.cpp:611-656;
 * lineLength, ctsupport.h:224-226.  The list is walked newest-first. */
static double project_line(const phantom_t* p, double x1, double y1, double x2, double y2) {
  double rsum = 0.0;
  for (int k = p->n - 1; k >= 0; k--) {
    double a = x1, b = y1, c = x2, d = y2;
    double s = 0.0;
    if (elem_clip(&p->e[k], &a, &b, &c, &d)) {
      double len = sqrt((c - a) * (c - a) + (d - b) * (d - b));
      s = len * p->e[k].atten;
    }
    rsum += s;
  }
  return rsum;
}

double orc_raysum_line(const int* types, const double* params, int nElem, const char* builtin,
                       double x1, double y1, double x2, double y2) {
  phantom_t p;
  if (phm_build(&p, types, params, nElem, builtin)) return NAN;
  double r = project_line(&p, x1, y1, x2, y2);
  free(p.e);
  return r;
}

int orc_phantom_extent(const int* types, const double* params, int nElem, const char* builtin,
                       double* ext) {
  phantom_t p;
  if (phm_build(&p, types, params, nElem, builtin)) return 1;
  ext[0] = p.xmin; ext[1] = p.xmax; ext[2] = p.ymin; ext[3] = p.ymax;
  free(p.e);
  return 0;
}

/* -------------------------------------------------------------- scanner */

enum { G_PARALLEL = 0, G_EQUIANGULAR = 1, G_EQUILINEAR = 2 };

static int geometry_id(const char* name) {
  static const char* names[] = {"parallel", "equiangular", "equilinear", "linogram"};
  for (int i = 0; i < 4; i++)
    if (!strcasecmp(name, names[i])) return i;
  return -1;
}

typedef struct {
  int geom, nDet, nView, offsetView;
  unsigned nSample;
  double viewDiameter, focalLength, centerDetLength, sdl, scanDiameter, xc, yc;
  double rotLen, rotInc, detLen, detStart, detInc, fanBeamAngle;
  double angDetLen, angDetInc;
  double xs1, ys1, xs2, ys2, xd1, yd1, xd2, yd2, angularDet;
  mtx2 rotInc_m;
} scanner_t;

/* Scanner::Scanner, scanner.cpp:86-224 */
static int scanner_init(scanner_t* s, const phantom_t* p, const char* geometry, int nDet, int nView,
                        int offsetView, int nSample, double rotLen, double fRatio, double cRatio,
                        double viewRatio, double scanRatio) {
  memset(s, 0, sizeof *s);
  s->geom = geometry_id(geometry);
  if (s->geom < 0) return fail(1, "Invalid geometry name");
  if (s->geom > G_EQUILINEAR) return fail(1, "Unsupported geometry");
  if (nView < 1 || nDet < 1) return fail(1, "nView & nDet must be greater than 0");
  s->nDet = nDet; s->nView = nView; s->offsetView = offsetView; s->nSample = (unsigned)nSample;
  double maxAxis = (p->xmax - p->xmin) > (p->ymax - p->ymin) ? (p->xmax - p->xmin) : (p->ymax - p->ymin);
  s->viewDiameter = SQRT2 * maxAxis * viewRatio;
  s->focalLength = (s->viewDiameter / 2) * fRatio;
  s->centerDetLength = (s->viewDiameter / 2) * cRatio;
  s->sdl = s->focalLength + s->centerDetLength;
  s->scanDiameter = s->viewDiameter * scanRatio;
  s->xc = p->xmin + (p->xmax - p->xmin) / 2;
  s->yc = p->ymin + (p->ymax - p->ymin) / 2;
  s->rotLen = rotLen;
  s->rotInc = s->rotLen / s->nView;
  int even = (nDet % 2) == 0;
  if (s->geom == G_PARALLEL) {
    s->fanBeamAngle = 0;
    s->detLen = s->scanDiameter;
    s->detStart = -s->detLen / 2;
    s->detInc = s->detLen / s->nDet;
    double endOff = 0;
    if (even) { s->detInc = s->detLen / (s->nDet - 1); endOff = s->detInc; }
    double h = s->detLen / 2;
    s->xs1 = s->xc - h; s->ys1 = s->yc + s->focalLength;
    s->xs2 = s->xc + h + endOff; s->ys2 = s->yc + s->focalLength;
    s->xd1 = s->xc - h; s->yd1 = s->yc - s->centerDetLength;
    s->xd2 = s->xc + h + endOff; s->yd2 = s->yc - s->centerDetLength;
    s->detLen += endOff;
  } else if (s->geom == G_EQUILINEAR) {
    if (s->scanDiameter / 2 >= s->focalLength)
      return fail(1, "Invalid geometry: Focal length must be larger than scan length");
    double a = asin((s->scanDiameter / 2) / s->focalLength);
    double h = s->sdl * tan(a);
    s->detLen = h * 2;
    s->detStart = -h;
    s->detInc = s->detLen / s->nDet;
    double endOff = 0;
    if (even) { s->detInc = s->detLen / (s->nDet - 1); endOff = s->detInc; s->detLen += endOff; }
    s->fanBeamAngle = a * 2;
    s->xs1 = s->xc; s->ys1 = s->yc + s->focalLength;
    s->xs2 = s->xc; s->ys2 = s->yc + s->focalLength;
    s->xd1 = s->xc - h; s->yd1 = s->yc - s->centerDetLength;
    s->xd2 = s->xc + h + endOff; s->yd2 = s->yc - s->centerDetLength;
  } else {
    if (s->scanDiameter / 2 > s->focalLength)
      return fail(1, "Invalid geometry: Focal length must be larger than scan length");
    double a = asin((s->scanDiameter / 2) / s->focalLength);
    s->detLen = 2 * a;
    s->detStart = -a;
    s->detInc = s->detLen / s->nDet;
    double endOff = 0;
    if (even) { s->detInc = s->detLen / (s->nDet - 1); endOff = s->detInc; }
    double a1 = acos((s->scanDiameter / 2) / s->centerDetLength);
    double angScale = 2 * (HALFPI + a - a1) / s->detLen;
    s->angDetLen = angScale * (s->detLen + endOff);
    s->angDetInc = angScale * s->detInc;
    s->angularDet = -s->angDetLen / 2;
    s->fanBeamAngle = a * 2;
    s->xs1 = s->xc; s->ys1 = s->yc + s->focalLength;
    s->xs2 = s->xc; s->ys2 = s->yc + s->focalLength;
    s->detLen += endOff;
  }
  mtx2 t;
  xlat_m(s->rotInc_m, -s->xc, -s->yc);
  rot_m(t, s->rotInc);
  mult_m(s->rotInc_m, t, s->rotInc_m);
  xlat_m(t, s->xc, s->yc);
  mult_m(s->rotInc_m, t, s->rotInc_m);
  return 0;
}

/* Scanner::projectSingleView, scanner.cpp:445-533 */
static void project_view(const scanner_t* s, const phantom_t* p, float* detval, double xd1,
                         double yd1, double xd2, double yd2, double xs1, double ys1, double xs2,
                         double ys2, double detAngle) {
  int nDet = s->nDet;
  double sdx = (xs2 - xs1) / nDet, sdy = (ys2 - ys1) / nDet;
  double xs_maj = xs1 + (sdx / 2), ys_maj = ys1 + (sdy / 2);
  double ddx = 0, ddy = 0, ddx2 = 0, ddy2 = 0, xd_maj = 0, yd_maj = 0;
  double aInc = 0, aSampInc = 0, aMaj = 0;
  if (s->geom == G_EQUIANGULAR) {
    aInc = s->angDetInc;
    aSampInc = aInc / s->nSample;
    double aOff = aSampInc / 2;
    aMaj = detAngle - (s->angDetLen / 2) + aOff;
  } else {
    ddx = (xd2 - xd1) / nDet;
    ddy = (yd2 - yd1) / nDet;
    ddx2 = ddx / s->nSample;
    ddy2 = ddy / s->nSample;
    xd_maj = xd1 + ddx2 / 2;
    yd_maj = yd1 + ddy2 / 2;
  }
  if (p->unitPulse) {
    for (int d = 0; d < nDet; d++) detval[d] = 0;
    detval[nDet / 2] = 1;
    return;
  }
  for (int d = 0; d < nDet; d++) {
    double xs = xs_maj, ys = ys_maj, xd = 0, yd = 0, ang = 0;
    if (s->geom == G_EQUIANGULAR) ang = aMaj;
    else { xd = xd_maj; yd = yd_maj; }
    double sum = 0.0;
    for (unsigned i = 0; i < s->nSample; i++) {
      if (s->geom == G_EQUIANGULAR) {
        xd = s->centerDetLength * cos(ang);
        yd = s->centerDetLength * sin(ang);
      }
      sum += project_line(p, xd, yd, xs, ys);
      if (s->geom == G_EQUIANGULAR) ang += aSampInc;
      else { xd += ddx2; yd += ddy2; }
    }
    detval[d] = (float)(sum / s->nSample);
    xs_maj += sdx;
    ys_maj += sdy;
    if (s->geom == G_EQUIANGULAR) aMaj += aInc;
    else { xd_maj += ddx; yd_maj += ddy; }
  }
}

/* Scanner::collectProjections, scanner.cpp:296-419 */
static void collect(const scanner_t* s, const phantom_t* p, int iStartView, int iNumViews,
                    int iOffsetView, int iStorageOffset, float* sino, double* angles) {
  double start = (iStartView + iOffsetView) * s->rotInc;
  mtx2 m0, t;
  xlat_m(m0, -s->xc, -s->yc);
  rot_m(t, start);
  mult_m(m0, t, m0);
  xlat_m(t, s->xc, s->yc);
  mult_m(m0, t, m0);
  double xd1 = 0, yd1 = 0, xd2 = 0, yd2 = 0;
  if (s->geom != G_EQUIANGULAR) {
    xd1 = s->xd1; yd1 = s->yd1; xd2 = s->xd2; yd2 = s->yd2;
    xform_m(m0, &xd1, &yd1);
    xform_m(m0, &xd2, &yd2);
  }
  double xs1 = s->xs1, ys1 = s->ys1, xs2 = s->xs2, ys2 = s->ys2;
  xform_m(m0, &xs1, &ys1);
  xform_m(m0, &xs2, &ys2);
  double va = start;
  for (int iv = 0; iv < iNumViews; iv++, va += s->rotInc) {
    int pos = iv + iStorageOffset;
    project_view(s, p, sino + (size_t)pos * s->nDet, xd1, yd1, xd2, yd2, xs1, ys1, xs2, ys2,
                 va + 3 * HALFPI);
    angles[pos] = va;
    xform_m(s->rotInc_m, &xs1, &ys1);
    xform_m(s->rotInc_m, &xs2, &ys2);
    if (s->geom != G_EQUIANGULAR) {
      xform_m(s->rotInc_m, &xd1, &yd1);
      xform_m(s->rotInc_m, &xd2, &yd2);
    }
  }
}

/* --------------------------------------------------------------- filters */

enum {
  F_ABS_BANDLIMIT = 0, F_ABS_G_HAMMING = 1, F_ABS_HANNING = 2, F_ABS_COSINE = 3,
  F_ABS_SINC = 4, F_SHEPP = 5, F_BANDLIMIT = 6, F_SINC = 7, F_G_HAMMING = 8,
  F_HANNING = 9, F_COSINE = 10, F_TRIANGLE = 11
};
/* filter.cpp:46-59: note "shepp" sits at index 4 (= FILTER_ABS_SINC) */
static const char* kFilterNames[] = {"abs_bandlimit", "abs_hamming", "abs_hanning", "abs_cosine",
                                     "shepp",         "abs_sinc",    "bandlimit",   "sinc",
                                     "hamming",       "hanning",     "cosine",      "triangle"};
static int lookup(const char* name, const char* const* table, int n) {
  for (int i = 0; i < n; i++)
    if (!strcasecmp(name, table[i])) return i;
  return -1;
}

/* SignalFilter::sinc, filter.h:120-124; integral_abscos, filter.h:163-164 */
static double sinc1(double x) { return fabs(x) > F_EPSILON ? sin(x) / x : 1.0; }
static double sinc2(double x, double m) { return fabs(x) > F_EPSILON ? sin(x * m) / x : 1.0; }
static double integral_abscos(double u, double w) {
  return fabs(u) > F_EPSILON ? (cos(u * w) - 1) / (u * u) + w / u * sin(u * w) : (w * w / 2);
}

/* SignalFilter::frequencyResponse, filter.cpp:403-475 */
double orc_frequency_response(int id, double bw, double u, double param) {
  double q, au = fabs(u), abw = fabs(bw);
  switch (id) {
    case F_BANDLIMIT: q = au >= (abw / 2) + F_EPSILON ? 0. : 1; break;
    case F_ABS_BANDLIMIT: q = au >= (abw / 2) + F_EPSILON ? 0. : au; break;
    case F_TRIANGLE: q = au >= (abw / 2) + F_EPSILON ? 0 : 1 - au / abw; break;
    case F_COSINE: q = au >= (abw / 2) + F_EPSILON ? 0 : cos(PI * au / abw); break;
    case F_ABS_COSINE: q = au >= (abw / 2) + F_EPSILON ? 0 : au * cos(PI * au / abw); break;
    case F_SINC: q = abw * sinc2(PI * abw * au, 1.); break;
    case F_ABS_SINC: q = au >= (abw / 2) + F_EPSILON ? 0 : au * abw * sinc2(PI * abw * au, 1.); break;
    case F_HANNING: param = 0.5; /* fall through */
    case F_G_HAMMING:
      q = au >= (abw / 2) + F_EPSILON ? 0 : param + (1 - param) * cos(TWOPI * au / abw);
      break;
    case F_ABS_HANNING: param = 0.5; /* fall through */
    case F_ABS_G_HAMMING:
      q = au >= (abw / 2) + F_EPSILON ? 0 : au * (param + (1 - param) * cos(TWOPI * au / abw));
      break;
    default: q = 0; break;
  }
  return q;
}

/* haveAnalyticSpatial, filter.cpp:489-512 */
static int have_analytic(int id) {
  switch (id) {
    case F_BANDLIMIT: case F_TRIANGLE: case F_COSINE: case F_G_HAMMING: case F_HANNING:
    case F_ABS_BANDLIMIT: case F_ABS_COSINE: case F_ABS_G_HAMMING: case F_ABS_HANNING:
    case F_SHEPP: case F_SINC:
      return 1;
  }
  return 0;
}

/* SignalFilter::spatialResponseAnalytic, filter.cpp:525-584 */
double orc_spatial_response_analytic(int id, double bw, double x, double param) {
  double q, temp, u = TWOPI * x, w = bw / 2, b = PI / bw, b2 = TWOPI / bw;
  switch (id) {
    case F_BANDLIMIT: q = bw * sinc2(u * w, 1.0); break;
    case F_TRIANGLE: temp = sinc2(u * w, 1.0); q = bw * temp * temp; break;
    case F_COSINE: q = sinc2(b - u, w) + sinc2(b + u, w); break;
    case F_HANNING: param = 0.5; /* fall through */
    case F_G_HAMMING:
      q = 2 * param * sin(u * w) / u + (1 - param) * (sinc2(b2 - u, w) + sinc2(b2 + u, w));
      break;
    case F_ABS_BANDLIMIT: q = 2 * integral_abscos(u, w); break;
    case F_ABS_COSINE: q = integral_abscos(b - u, w) + integral_abscos(b + u, w); break;
    case F_ABS_HANNING: param = 0.5; /* fall through */
    case F_ABS_G_HAMMING:
      q = 2 * param * integral_abscos(u, w) +
          (1 - param) * (integral_abscos(u - b2, w) + integral_abscos(u + b2, w));
      break;
    case F_SHEPP:
      if (fabs(u) < 1E-7) q = 4. / (PI * bw * bw);
      else q = fabs((2 / bw) * sin(u * w)) * sinc2(u * w, 1.) * sinc2(u * w, 1.);
      break;
    case F_SINC: q = fabs(x) < bw / 2 ? 1. : 0.; break;
    default: q = 0; break;
  }
  return q;
}

/* SignalFilter::spatialResponseCalc, filter.cpp:358-381 */
double orc_spatial_response_calc(int id, double bw, double x, double param, int n) {
  double zmin = 0, zmax = id == F_TRIANGLE ? bw : bw / 2;
  double zinc = (zmax - zmin) / (n - 1);
  double z = zmin;
  double* q = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++, z += zinc)
    q[i] = orc_frequency_response(id, bw, z, param) * cos(TWOPI * z * x);
  double y = 2 * orc_integrate_simpson(zmin, zmax, q, n);
  free(q);
  return y;
}

/* SignalFilter::init + createFrequencyFilter / createSpatialFilter,
 * filter.cpp:158-229 (N_INTEGRAL = 500, filter.cpp:28) */
static void signal_filter(double* out, int id, double fmin, double fmax, int n, double bw,
                          double param, int spatial) {
  double inc = (fmax - fmin) / (n - 1);
  if (!spatial) {
    double x = fmin;
    for (int i = 0; i < n; x += inc, i++) out[i] = orc_frequency_response(id, bw, x, param);
    return;
  }
  if (id == F_SHEPP) {
    double a = 2 * bw, c = -4. / (a * a);
    int center = (n - 1) / 2;
    out[center] = 4. / (a * a);
    for (int i = 1; i <= center; i++) out[center + i] = out[center - i] = c / (4 * (i * i) - 1);
  } else {
    double x = fmin;
    for (int i = 0; i < n; i++, x += inc)
      out[i] = have_analytic(id) ? orc_spatial_response_analytic(id, bw, x, param)
                                 : orc_spatial_response_calc(id, bw, x, param, 500);
  }
}

/* Fourier::shuffleNaturalToFourierOrder / shuffleFourierToNaturalOrder,
 * include/fourier.h:49-101 */
static void shuffle_nat_to_fourier(double* v, int n) {
  double* t = (double*)malloc(sizeof(double) * n);
  if (n % 2) {
    int h = (n - 1) / 2;
    t[0] = v[h];
    for (int i = 0; i < h; i++) t[i + 1] = v[i + 1 + h];
    for (int i = 0; i < h; i++) t[i + h + 1] = v[i];
  } else {
    int h = n / 2;
    t[0] = v[h];
    for (int i = 0; i < h - 1; i++) t[i + 1] = v[i + h + 1];
    for (int i = 0; i < h; i++) t[i + h] = v[i];
  }
  memcpy(v, t, sizeof(double) * n);
  free(t);
}
static void shuffle_fourier_to_nat(double* v, int n) {
  double* t = (double*)malloc(sizeof(double) * n);
  if (n % 2) {
    int h = (n - 1) / 2;
    t[h] = v[0];
    for (int i = 0; i < h; i++) t[i + 1 + h] = v[i + 1];
    for (int i = 0; i < h; i++) t[i] = v[i + h + 1];
  } else {
    int h = n / 2;
    t[h] = v[0];
    for (int i = 0; i < h; i++) t[i] = v[i + h];
    for (int i = 0; i < h - 1; i++) t[i + h + 1] = v[i + 1];
  }
  memcpy(v, t, sizeof(double) * n);
  free(t);
}

/* ProcessSignal::finiteFourierTransform, procsignal.cpp:695-774.
 * dir < 0 is FORWARD (normalised by 1/n), dir > 0 BACKWARD. */
static void dft_r2c(const double* in, double* re, double* im, int n, int dir) {
  dir = dir < 0 ? -1 : 1;
  double ainc = dir * 2 * PI / n;
  for (int i = 0; i < n; i++) {
    double sr = 0, si = 0;
    for (int j = 0; j < n; j++) {
      double a = i * j * ainc;
      sr += in[j] * cos(a);
      si += in[j] * sin(a);
    }
    if (dir < 0) { sr /= n; si /= n; }
    re[i] = sr;
    im[i] = si;
  }
}
static void dft_c2r(const double* re, const double* im, double* out, int n, int dir) {
  dir = dir < 0 ? -1 : 1;
  double ainc = dir * 2 * PI / n;
  for (int i = 0; i < n; i++) {
    double sr = 0;
    for (int j = 0; j < n; j++) {
      double a = i * j * ainc;
      sr += re[j] * cos(a) - im[j] * sin(a);
    }
    if (dir < 0) sr /= n;
    out[i] = sr;
  }
}
/* table-driven variants, procsignal.cpp:778-864 */
static void dft_r2c_tab(const double* in, double* re, double* im, int n, int dir,
                        const double* ct, const double* st) {
  dir = dir < 0 ? -1 : 1;
  for (int i = 0; i < n; i++) {
    double sr = 0, si = 0;
    for (int j = 0; j < n; j++) {
      int k = i * j;
      if (dir > 0) { sr += in[j] * ct[k]; si += in[j] * st[k]; }
      else { sr += in[j] * ct[k]; si -= in[j] * st[k]; }
    }
    if (dir < 0) { sr /= n; si /= n; }
    re[i] = sr;
    im[i] = si;
  }
}
static void dft_c2r_tab(const double* re, const double* im, double* out, int n, int dir,
                        const double* ct, const double* st) {
  dir = dir < 0 ? -1 : 1;
  for (int i = 0; i < n; i++) {
    double sr = 0;
    for (int j = 0; j < n; j++) {
      int k = i * j;
      if (dir > 0) sr += re[j] * ct[k] - im[j] * st[k];
      else sr += re[j] * ct[k] - im[j] * -st[k];
    }
    if (dir < 0) sr /= n;
    out[i] = sr;
  }
}

/* ProcessSignal::addZeropadFactor, procsignal.cpp:867-881 */
int orc_add_zeropad_factor(int n, int zeropad) {
  if (zeropad > 0) {
    double l = log(n) / log(2);
    int il = (int)floor(l);
    return 1 << (il + zeropad);
  }
  return n;
}

enum { M_CONVOLUTION = 0, M_FOURIER = 1, M_FOURIER_TABLE = 2, M_FFT = 3, M_FFTW = 4, M_RFFTW = 5 };

/* FFTW's real-to-halfcomplex transform and its inverse, by their published
 * definitions (FFTW 3 manual, "The Halfcomplex-format DFT"): R2HC stores
 * Re X_k at k (k <= N/2) and Im X_k at N-k (0 < k < N/2) for the unnormalised
 * forward DFT X_k = sum_j x_j e^{-2 pi i jk/N}; HC2R is its unnormalised
 * inverse.  Angles are reduced exactly, (j*k mod N) * 2 pi / N. */
static void fftw_r2hc(const double* x, double* hc, int N) {
  for (int k = 0; k <= N / 2; k++) {
    long double re = 0, im = 0;
    for (int j = 0; j < N; j++) {
      long m = ((long)j * k) % N;
      long double a = 6.283185307179586476925286766559L * m / N;
      re += x[j] * cosl(a);
      im -= x[j] * sinl(a);
    }
    hc[k] = (double)re;
    if (k > 0 && k < N - k) hc[N - k] = (double)im;
  }
}
static void fftw_hc2r(const double* hc, double* y, int nOut, int N) {
  for (int j = 0; j < nOut; j++) {
    long double s = hc[0];
    for (int k = 1; k < N - k; k++) {
      long m = ((long)j * k) % N;
      long double a = 6.283185307179586476925286766559L * m / N;
      s += 2.0L * (hc[k] * cosl(a) - hc[N - k] * sinl(a));
    }
    if (N % 2 == 0) s += (j % 2 ? -1.0L : 1.0L) * hc[N / 2];
    y[j] = (double)s;
  }
}
/* FFTW's complex DFT of a real input, multiplied by a real filter, inverse,
 * real part (procsignal.cpp:614-626). */
static void fftw_complex_filter(const double* x, const double* filt, double* y, int nOut, int N) {
  long double* re = (long double*)malloc(sizeof(long double) * N);
  long double* im = (long double*)malloc(sizeof(long double) * N);
  for (int k = 0; k < N; k++) {
    long double r = 0, i = 0;
    for (int j = 0; j < N; j++) {
      long m = ((long)j * k) % N;
      long double a = 6.283185307179586476925286766559L * m / N;
      r += x[j] * cosl(a);
      i -= x[j] * sinl(a);
    }
    re[k] = r * filt[k];
    im[k] = i * filt[k];
  }
  for (int j = 0; j < nOut; j++) {
    long double s = 0;
    for (int k = 0; k < N; k++) {
      long m = ((long)j * k) % N;
      long double a = 6.283185307179586476925286766559L * m / N;
      s += re[k] * cosl(a) - im[k] * sinl(a);
    }
    y[j] = (double)s;
  }
  free(re);
  free(im);
}

typedef struct {
  int filterId, method, generation, geom, n, nfp, nOut, preinterp;
  double inc, bw, F, sdl, fmin, fmax, finc;
  double* filter;
  double *ct, *st;
} procsig_t;

static void procsig_free(procsig_t* ps) {
  free(ps->filter);
  free(ps->ct);
  free(ps->st);
}

/* ProcessSignal::ProcessSignal + init, procsignal.cpp:82-442 */
static int procsig_init(procsig_t* ps, const char* filterName, const char* methodName, double bw,
                        double inc, int n, double param, const char* domainName,
                        const char* genName, int zeropad, int preinterp, int geom, double F,
                        double sdl) {
  /* "fftw"/"rfftw" exist in FFTW builds of the reference (procsignal.cpp:38-41) */
  static const char* methods[] = {"convolution", "fourier", "fouier-table", "fft", "fftw", "rfftw"};
  static const char* gens[] = {"direct", "inverse-fourier"};
  static const char* domains[] = {"frequency", "spatial"};
  memset(ps, 0, sizeof *ps);
  ps->method = lookup(methodName, methods, 6);
  if (ps->method < 0) return fail(1, "Invalid filter method name");
  ps->generation = lookup(genName, gens, 2);
  if (ps->generation < 0) return fail(1, "Invalid frequency filter name");
  ps->filterId = lookup(filterName, kFilterNames, 12);
  if (ps->filterId < 0) return fail(1, "Invalid Filter name");
  if (lookup(domainName, domains, 2) < 0) return fail(1, "Invalid domain name");
  ps->geom = geom; ps->F = F; ps->sdl = sdl; ps->bw = bw; ps->n = n; ps->inc = inc;
  ps->preinterp = preinterp;
  if (geom == G_EQUILINEAR) {
    double sc = sdl / F;
    ps->inc /= sc;
    ps->bw *= sc;
  }
  /* With FFTW the reference remaps "fft" to its real FFTW path
   * (procsignal.cpp:160-168); that is the behaviour restated here. */
  if (ps->method == M_FFT) ps->method = M_RFFTW;
  int freq = ps->method != M_CONVOLUTION;
  int i;
  if (!freq) {
    ps->nfp = 2 * (n - 1) + 1;
    if (ps->generation == 0) {
      ps->fmin = -ps->inc * (n - 1);
      ps->fmax = ps->inc * (n - 1);
      ps->finc = (ps->fmax - ps->fmin) / (ps->nfp - 1);
      ps->filter = (double*)malloc(sizeof(double) * ps->nfp);
      signal_filter(ps->filter, ps->filterId, ps->fmin, ps->fmax, ps->nfp, ps->bw, param, 1);
    } else {
      ps->fmin = -1. / (2 * ps->inc);
      ps->fmax = 1. / (2 * ps->inc);
      ps->finc = (ps->fmax - ps->fmin) / (ps->nfp - 1);
      double* fr = (double*)malloc(sizeof(double) * ps->nfp);
      signal_filter(fr, ps->filterId, ps->fmin, ps->fmax, ps->nfp, ps->bw, param, 0);
      shuffle_nat_to_fourier(fr, ps->nfp);
      ps->filter = (double*)malloc(sizeof(double) * ps->nfp);
      double* im = (double*)malloc(sizeof(double) * ps->nfp);
      dft_r2c(fr, ps->filter, im, ps->nfp, -1);
      free(im);
      free(fr);
      shuffle_fourier_to_nat(ps->filter, ps->nfp);
      for (i = 0; i < ps->nfp; i++) ps->filter[i] /= ps->inc;
    }
    if (geom == G_EQUILINEAR) {
      for (i = 0; i < ps->nfp; i++) ps->filter[i] *= 0.5;
    } else if (geom == G_EQUIANGULAR) {
      for (i = 0; i < ps->nfp; i++) {
        int k = i - ((ps->nfp - 1) / 2);
        double s = 1 / sinc1(k * ps->inc);
        ps->filter[i] *= 0.5 * s * s;
      }
    }
  } else if (ps->generation == 0) {
    ps->nfp = orc_add_zeropad_factor(n, zeropad);
    ps->nOut = ps->nfp * preinterp;
    ps->fmin = -1. / (2 * ps->inc);
    ps->fmax = 1. / (2 * ps->inc);
    if (ps->nfp % 2) {
      ps->finc = (ps->fmax - ps->fmin) / (ps->nfp - 1);
    } else {
      ps->finc = (ps->fmax - ps->fmin) / ps->nfp;
      ps->fmax -= ps->finc;
    }
    ps->filter = (double*)malloc(sizeof(double) * ps->nfp);
    signal_filter(ps->filter, ps->filterId, ps->fmin, ps->fmax, ps->nfp, ps->bw, param, 0);
    if (geom == G_EQUILINEAR) {
      for (i = 0; i < ps->nfp; i++) ps->filter[i] *= 0.5;
    } else if (geom == G_EQUIANGULAR) {
      for (i = 0; i < ps->nfp; i++) {
        int k = i - ((ps->nfp - 1) / 2);
        double s = 1 / sinc1(k * ps->inc);
        ps->filter[i] *= 0.5 * s * s;
      }
    }
    shuffle_nat_to_fourier(ps->filter, ps->nfp);
  } else {
    int nsp = 2 * (n - 1) + 1;
    ps->fmin = -ps->inc * (n - 1);
    ps->fmax = ps->inc * (n - 1);
    ps->finc = (ps->fmax - ps->fmin) / (nsp - 1);
    ps->nfp = nsp;
    if (zeropad > 0) {
      double l = log(nsp) / log(2);
      int np2 = (int)floor(l);
      if (l != floor(l)) np2++;
      np2 += zeropad - 1;
      ps->nfp = 1 << np2;
    }
    ps->nOut = ps->nfp * preinterp;
    double* sp = (double*)calloc(ps->nfp > nsp ? ps->nfp : nsp, sizeof(double));
    signal_filter(sp, ps->filterId, ps->fmin, ps->fmax, nsp, ps->bw, param, 1);
    if (geom == G_EQUILINEAR) {
      for (i = 0; i < nsp; i++) sp[i] *= 0.5;
    } else if (geom == G_EQUIANGULAR) {
      for (i = 0; i < nsp; i++) {
        int k = i - ((nsp - 1) / 2);
        double s = sin(k * ps->inc);
        if (fabs(s) < 1E-7) s = 1;
        else s = (k * ps->inc) / s;
        sp[i] *= 0.5 * s * s;
      }
    }
    for (i = nsp; i < ps->nfp; i++) sp[i] = 0;
    double* re = (double*)malloc(sizeof(double) * ps->nfp);
    double* im = (double*)malloc(sizeof(double) * ps->nfp);
    dft_r2c(sp, re, im, ps->nfp, 1);
    ps->filter = (double*)malloc(sizeof(double) * ps->nfp);
    for (i = 0; i < ps->nfp; i++) ps->filter[i] = hypot(re[i], im[i]) * ps->inc;
    free(re);
    free(im);
    free(sp);
  }
  if (ps->method == M_FFTW || ps->method == M_RFFTW) /* procsignal.cpp:413-416 */
    for (i = 0; i < ps->nfp; i++) ps->filter[i] /= ps->nfp;
  if (ps->method == M_FOURIER_TABLE) {
    int m = ps->nfp > ps->nOut ? ps->nfp : ps->nOut;
    int nF = m * m + 1;
    double ainc = (2. * PI) / ps->nfp, a = 0;
    ps->ct = (double*)malloc(sizeof(double) * nF);
    ps->st = (double*)malloc(sizeof(double) * nF);
    for (i = 0; i < nF; i++) {
      ps->ct[i] = cos(a);
      ps->st[i] = sin(a);
      a += ainc;
    }
  }
  return 0;
}

/* ProcessSignal::filterSignal + convolve, procsignal.cpp:544-673 */
static void procsig_filter(const procsig_t* ps, const float* cin, double* out) {
  int n = ps->n, i;
  double* in = (double*)malloc(sizeof(double) * n);
  for (i = 0; i < n; i++) in[i] = cin[i];
  if (ps->geom == G_EQUILINEAR) {
    for (i = 0; i < n; i++) {
      int k = i - (n / 2);
      in[i] *= ps->F / sqrt(ps->F * ps->F + k * k * ps->inc * ps->inc);
    }
  } else if (ps->geom == G_EQUIANGULAR) {
    for (i = 0; i < n; i++) {
      int k = i - (n / 2);
      in[i] *= ps->F * cos(k * ps->inc);
    }
  }
  if (ps->method == M_CONVOLUTION) {
    for (i = 0; i < n; i++) {
      double sum = 0.0;
      const double* f1 = in;
      const double* f2 = ps->filter + i + (n - 1);
      for (int j = 0; j < n; j++) sum += *f1++ * *f2--;
      out[i] = sum * ps->inc;
    }
  } else if (ps->method == M_RFFTW || ps->method == M_FFTW) {
    /* procsignal.cpp:600-626 */
    int N = ps->nfp;
    double* x = (double*)calloc(N, sizeof(double));
    double* y = (double*)malloc(sizeof(double) * N);
    for (i = 0; i < n; i++) x[i] = in[i];
    if (ps->method == M_RFFTW) {
      double* hc = (double*)malloc(sizeof(double) * N);
      fftw_r2hc(x, hc, N);
      for (i = 0; i < N; i++) hc[i] = ps->filter[i] * hc[i];
      fftw_hc2r(hc, y, n, N);
      free(hc);
    } else {
      fftw_complex_filter(x, ps->filter, y, n, N);
    }
    for (i = 0; i < n; i++) out[i] = y[i];
    free(x);
    free(y);
  } else {
    int N = ps->nfp;
    double* x = (double*)calloc(N, sizeof(double));
    double* re = (double*)malloc(sizeof(double) * N);
    double* im = (double*)malloc(sizeof(double) * N);
    double* y = (double*)malloc(sizeof(double) * N);
    for (i = 0; i < n; i++) x[i] = in[i];
    if (ps->method == M_FOURIER) dft_r2c(x, re, im, N, -1);
    else dft_r2c_tab(x, re, im, N, -1, ps->ct, ps->st);
    for (i = 0; i < N; i++) { re[i] *= ps->filter[i]; im[i] *= ps->filter[i]; }
    if (ps->method == M_FOURIER) dft_c2r(re, im, y, N, 1);
    else dft_c2r_tab(re, im, y, N, 1, ps->ct, ps->st);
    for (i = 0; i < n; i++) out[i] = y[i];
    free(x); free(re); free(im); free(y);
  }
  free(in);
}

int orc_filter(const char* filterName, const char* methodName, double bw, double inc, int n,
               double param, const char* domainName, const char* generationName, int zeropad,
               int preinterp, int geometry, double focalLength, double sdl, const float* in,
               int nView, double* out, double* filterOut, int filterCap, int* nFilterPoints,
               double* filterMinMaxInc) {
  procsig_t ps;
  int rc = procsig_init(&ps, filterName, methodName, bw, inc, n, param, domainName,
                        generationName, zeropad, preinterp, geometry, focalLength, sdl);
  if (rc) { procsig_free(&ps); return rc; }
  if (nFilterPoints) *nFilterPoints = ps.nfp;
  if (filterMinMaxInc) {
    filterMinMaxInc[0] = ps.fmin; filterMinMaxInc[1] = ps.fmax; filterMinMaxInc[2] = ps.finc;
  }
  if (filterOut)
    for (int i = 0; i < ps.nfp && i < filterCap; i++) filterOut[i] = ps.filter[i];
  for (int v = 0; v < nView; v++) procsig_filter(&ps, in + (size_t)v * n, out + (size_t)v * n);
  procsig_free(&ps);
  return 0;
}

/* --------------------------------------------------------- backprojection */

/* CubicPolyInterpolator::interpolate, libctsupport/interpolator.cpp:36-75 */
double orc_cubic_interpolate(const double* Y, int n, double x) {
  int lo = (int)floor(x) - 1, hi = lo + 3;
  if (lo < -1) return 0;
  else if (lo == -1) return Y[0] + x * (Y[1] - Y[0]);
  if (hi > n) return 0;
  else if (hi == n) {
    double frac = x - (lo + 1);
    return Y[n - 2] + frac * (Y[n - 1] - Y[n - 2]);
  }
  double xd0 = x - lo, xd1 = x - (lo + 1), xd2 = x - (lo + 2), xd3 = x - (lo + 3);
  static const double oneSixth = (1. / 6.);
  double y = xd1 * xd2 * xd3 * -oneSixth * Y[lo];
  y += xd0 * xd2 * xd3 * 0.5 * Y[lo + 1];
  y += xd0 * xd1 * xd3 * -0.5 * Y[lo + 2];
  y += xd0 * xd1 * xd2 * oneSixth * Y[lo + 3];
  return y;
}

enum { BP_TRIG = 0, BP_TABLE = 1, BP_DIFF = 2, BP_IDIFF = 3 };
enum { I_NEAREST = 0, I_LINEAR = 1, I_CUBIC = 2 };

typedef struct {
  int family, interp, geom, nDet, iDetCenter, nx, ny;
  double detInc, rotScale, xMin, xMax, yMin, yMax, xInc, yInc, F, sdl, start_r, start_phi;
  double *r, *phi;
} bp_t;

/* Backprojector::initBackprojector + Backproject::Backproject + table and
 * diff constructors, backprojectors.cpp:135-180, 264-319, 413-429, 491-502 */
static int bp_init(bp_t* b, int geom, int nDet, double detInc, int nView, double viewDiameter,
                   double F, double sdl, const char* bpName, const char* interpName, int nx,
                   int ny, const double* roi) {
  static const char* bps[] = {"trig", "table", "diff", "idiff"};
  static const char* ips[] = {"nearest", "linear", "cubic"};
  memset(b, 0, sizeof *b);
  b->family = lookup(bpName, bps, 4);
  if (b->family < 0) return fail(1, "Invalid backprojection name");
  b->interp = lookup(interpName, ips, 3);
  if (b->interp < 0) return fail(1, "Invalid interpolation name");
  if (geom == G_EQUILINEAR || geom == G_EQUIANGULAR) b->family = BP_TABLE; /* tables r/phi */
  b->geom = geom;
  b->detInc = detInc;
  b->nDet = nDet;
  b->iDetCenter = (nDet - 1) / 2;
  b->rotScale = geom == G_PARALLEL ? PI / nView : (2 * PI) / nView;
  b->nx = nx; b->ny = ny;
  double phmLen = viewDiameter / SQRT2;
  b->xMin = -phmLen / 2; b->xMax = b->xMin + phmLen;
  b->yMin = -phmLen / 2; b->yMax = b->yMin + phmLen;
  if (roi) { /* roi = {xmin, xmax, ymin, ymax} */
    if (roi[0] > b->xMin) b->xMin = roi[0];
    if (roi[1] < b->xMax) b->xMax = roi[1];
    if (roi[2] > b->yMin) b->yMin = roi[2];
    if (roi[3] < b->yMax) b->yMax = roi[3];
    if (b->xMin > b->xMax) { double t = b->xMin; b->xMin = b->xMax; b->xMax = t; }
    if (b->yMin > b->yMax) { double t = b->yMin; b->yMin = b->yMax; b->yMax = t; }
  }
  b->xInc = (b->xMax - b->xMin) / nx;
  b->yInc = (b->yMax - b->yMin) / ny;
  b->F = F;
  b->sdl = sdl;
  if (b->family == BP_TABLE) {
    b->r = (double*)malloc(sizeof(double) * nx * ny);
    b->phi = (double*)malloc(sizeof(double) * nx * ny);
    double x, y;
    int ix, iy;
    for (x = b->xMin + b->xInc / 2, ix = 0; ix < nx; x += b->xInc, ix++)
      for (y = b->yMin + b->yInc / 2, iy = 0; iy < ny; y += b->yInc, iy++) {
        b->r[(size_t)ix * ny + iy] = sqrt(x * x + y * y);
        b->phi[(size_t)ix * ny + iy] = atan2(y, x);
      }
  }
  if (b->family == BP_DIFF || b->family == BP_IDIFF) {
    double x = b->xMin + b->xInc / 2, y = b->yMin + b->yInc / 2;
    b->start_r = sqrt(x * x + y * y);
    b->start_phi = atan2(y, x);
  }
  return 0;
}

static void bp_free(bp_t* b) {
  free(b->r);
  free(b->phi);
}

/* BackprojectTrig::BackprojectView, backprojectors.cpp:365-404 */
static void bp_trig(const bp_t* b, float* v, const double* f, double theta) {
  int nDet = b->nDet;
  double x = b->xMin + b->xInc / 2;
  for (int ix = 0; ix < b->nx; x += b->xInc, ix++) {
    double y = b->yMin + b->yInc / 2;
    for (int iy = 0; iy < b->ny; y += b->yInc, iy++) {
      float* pv = &v[(size_t)ix * b->ny + iy];
      double r = sqrt(x * x + y * y);
      double phi = atan2(y, x);
      double L = r * cos(theta - phi);
      if (b->interp == I_NEAREST) {
        int i = b->iDetCenter + nearest_int(L / b->detInc);
        if (i >= 0 && i < nDet) *pv += b->rotScale * f[i];
      } else if (b->interp == I_LINEAR) {
        double p = L / b->detInc, pf = floor(p);
        int i = b->iDetCenter + (int)pf;
        double frac = p - pf;
        if (i >= 0 && i < nDet - 1) *pv += b->rotScale * ((1 - frac) * f[i] + frac * f[i + 1]);
      } else {
        double p = b->iDetCenter + (L / b->detInc);
        if (p >= 0 && p < nDet) *pv += b->rotScale * orc_cubic_interpolate(f, nDet, p);
      }
    }
  }
}

/* BackprojectTable::BackprojectView, backprojectors.cpp:444-481 */
static void bp_table(const bp_t* b, float* v, const double* f, double theta) {
  int nDet = b->nDet;
  for (int ix = 0; ix < b->nx; ix++)
    for (int iy = 0; iy < b->ny; iy++) {
      size_t k = (size_t)ix * b->ny + iy;
      double L = b->r[k] * cos(theta - b->phi[k]);
      if (b->interp == I_NEAREST) {
        int i = b->iDetCenter + nearest_int(L / b->detInc);
        if (i >= 0 && i < nDet) v[k] += f[i];
      } else if (b->interp == I_LINEAR) {
        double p = L / b->detInc, pf = floor(p);
        int i = b->iDetCenter + (int)pf;
        double frac = p - pf;
        if (i >= 0 && i < nDet - 1) v[k] += ((1 - frac) * f[i] + frac * f[i + 1]);
      } else {
        double p = b->iDetCenter + (L / b->detInc);
        if (p >= 0 && p < nDet) v[k] += orc_cubic_interpolate(f, nDet, p);
      }
    }
}

/* BackprojectDiff::BackprojectView, backprojectors.cpp:517-570 (the column
 * pointer only advances on in-range samples) */
static void bp_diff(const bp_t* b, float* v, const double* f, double theta) {
  int nDet = b->nDet;
  double det_dx = b->xInc * cos(theta) / b->detInc;
  double det_dy = b->yInc * sin(theta) / b->detInc;
  double colStart = b->iDetCenter + b->start_r * cos(theta - b->start_phi) / b->detInc;
  double* df = NULL;
  if (b->interp == I_LINEAR) {
    df = (double*)malloc(sizeof(double) * nDet);
    for (int i = 0; i < nDet - 1; i++) df[i] = f[i + 1] - f[i];
    df[nDet - 1] = 0;
  }
  int last = nDet - 1;
  for (int ix = 0; ix < b->nx; ix++, colStart += det_dx) {
    double cur = colStart;
    float* col = &v[(size_t)ix * b->ny];
    for (int iy = 0; iy < b->ny; iy++, cur += det_dy) {
      if (b->interp == I_NEAREST) {
        int i = nearest_int(cur);
        if (i >= 0 && i < nDet) *col++ += f[i];
      } else if (b->interp == I_LINEAR) {
        double pf = floor(cur);
        int i = (int)pf;
        double frac = cur - pf;
        if (i >= 0 && i <= last) *col++ += f[i] + (frac * df[i]);
      } else {
        double p = cur;
        if (p >= 0 && p < nDet) *col++ += orc_cubic_interpolate(f, nDet, p);
      }
    }
  }
  free(df);
}

/* BackprojectIntDiff::BackprojectView, backprojectors.cpp:579-653 (LP64:
 * 32.32 fixed point) */
static void bp_idiff(const bp_t* b, float* v, const double* f, double theta) {
  const int shift = 32;
  const long scale = 1L << shift, mask = scale - 1, half = scale / 2;
  const double inv = 1. / scale;
  int nDet = b->nDet;
  const long det_dx = nearest_long(b->xInc * cos(theta) / b->detInc * scale);
  const long det_dy = nearest_long(b->yInc * sin(theta) / b->detInc * scale);
  long colStart =
      nearest_long((b->start_r * cos(theta - b->start_phi) / b->detInc + b->iDetCenter) * scale);
  double* df = NULL;
  if (b->interp == I_LINEAR) {
    df = (double*)malloc(sizeof(double) * nDet);
    for (int i = 0; i < nDet - 1; i++) df[i] = (f[i + 1] - f[i]) * inv;
    df[nDet - 1] = 0;
  }
  int last = nDet - 1;
  for (int ix = 0; ix < b->nx; ix++, colStart += det_dx) {
    long cur = colStart;
    float* col = &v[(size_t)ix * b->ny];
    if (b->interp == I_NEAREST) {
      for (int iy = 0; iy < b->ny; iy++, cur += det_dy) {
        const int i = (int)((cur + half) >> shift);
        if (i >= 0 && i <= last) *col++ += f[i];
        else col++;
      }
    } else if (b->interp == I_LINEAR) {
      for (int iy = 0; iy < b->ny; iy++, cur += det_dy) {
        const long i = cur >> shift;
        if (i >= 0 && i <= last) {
          const long rem = cur & mask;
          *col++ += f[i] + (rem * df[i]);
        } else
          col++;
      }
    } else {
      for (int iy = 0; iy < b->ny; iy++, cur += det_dy)
        *col++ += orc_cubic_interpolate(f, nDet, (double)cur / scale);
    }
  }
  free(df);
}

/* BackprojectEquiangular / BackprojectEquilinear::BackprojectView,
 * backprojectors.cpp:656-743 */
static void bp_fan(const bp_t* b, float* v, const double* f, double beta) {
  int nDet = b->nDet;
  for (int ix = 0; ix < b->nx; ix++)
    for (int iy = 0; iy < b->ny; iy++) {
      size_t k = (size_t)ix * b->ny + iy;
      double d = beta - b->phi[k];
      double rc = b->r[k] * cos(d), rs = b->r[k] * sin(d);
      double pos, w;
      if (b->geom == G_EQUIANGULAR) {
        double fl = b->F + rs;
        double gamma = atan(rc / fl);
        pos = gamma / b->detInc;
        w = fl * fl + (rc * rc); /* divide by L^2 */
      } else {
        double U = (b->F + rs) / b->F;
        double dp = rc / U;
        dp *= b->sdl / b->F;
        pos = dp / b->detInc;
        w = U * U;
      }
      if (b->interp == I_NEAREST) {
        int i = b->iDetCenter + nearest_int(pos);
        if (i >= 0 && i < nDet) v[k] += f[i] / w;
      } else if (b->interp == I_LINEAR) {
        double pf = floor(pos);
        int i = b->iDetCenter + (int)pf;
        double frac = pos - pf;
        if (i >= 0 && i < nDet - 1) v[k] += (f[i] + frac * (f[i + 1] - f[i])) / w;
      } else {
        double p = b->iDetCenter + pos;
        if (p >= 0 && p < nDet) v[k] += orc_cubic_interpolate(f, nDet, p) / w;
      }
    }
}

static void bp_view(const bp_t* b, float* v, const double* f, double angle) {
  if (b->geom != G_PARALLEL) bp_fan(b, v, f, angle);
  else if (b->family == BP_TRIG) bp_trig(b, v, f, angle);
  else if (b->family == BP_TABLE) bp_table(b, v, f, angle);
  else if (b->family == BP_DIFF) bp_diff(b, v, f, angle);
  else bp_idiff(b, v, f, angle);
}

/* PostProcessing -> ScaleImageByRotIncrement, backprojectors.cpp:324-336,
 * 435-442, 508-515 (Trig scales per view instead) */
static void bp_post(const bp_t* b, float* v) {
  if (b->geom == G_PARALLEL && b->family == BP_TRIG) return;
  for (size_t k = 0; k < (size_t)b->nx * b->ny; k++) v[k] *= b->rotScale;
}

int orc_backproject(const int* types, const double* params, int nElem, const char* builtin,
                    const char* geometry, int nDet, int nView, int offsetView, int nSample,
                    double rotLen, double focalRatio, double centerRatio, double viewRatio,
                    double scanRatio, const char* backprojName, const char* interpName,
                    int interpFactor, int nx, int ny, const double* roi, const double* filtered,
                    const double* angles, int postProcess, float* image) {
  phantom_t p;
  scanner_t s;
  bp_t b;
  if (phm_build(&p, types, params, nElem, builtin)) return 1;
  int rc = scanner_init(&s, &p, geometry, nDet, nView, offsetView, nSample, rotLen, focalRatio,
                        centerRatio, viewRatio, scanRatio);
  if (rc) { free(p.e); return 2; }
  if (bp_init(&b, s.geom, nDet, s.detInc, nView, s.viewDiameter, s.focalLength, s.sdl,
              backprojName, interpName, nx, ny, roi)) {
    free(p.e);
    return 4;
  }
  memset(image, 0, sizeof(float) * nx * ny);
  int nf = nDet * interpFactor;
  for (int v = 0; v < nView; v++) bp_view(&b, image, filtered + (size_t)v * nf, angles[v]);
  if (postProcess) bp_post(&b, image);
  bp_free(&b);
  free(p.e);
  return 0;
}

int orc_pipeline(const int* types, const double* params, int nElem, const char* builtin,
                 const char* geometry, int nDet, int nView, int offsetView, int nSample,
                 double rotLen, double focalRatio, double centerRatio, double viewRatio,
                 double scanRatio, const char* filterName, double filterParam,
                 const char* methodName, int zeropad, const char* generationName,
                 const char* interpName, int interpFactor, const char* backprojName, int nx,
                 int ny, float* sino, double* angles, double* filtered, float* image,
                 float* imageRecon, double* info, double* timings) {
  phantom_t p;
  scanner_t s;
  if (phm_build(&p, types, params, nElem, builtin)) return 1;
  if (scanner_init(&s, &p, geometry, nDet, nView, offsetView, nSample, rotLen, focalRatio,
                   centerRatio, viewRatio, scanRatio)) {
    free(p.e);
    return 2;
  }
  double t0 = now_s();
  collect(&s, &p, 0, nView, offsetView, 0, sino, angles);
  double t1 = now_s();
  if (info) {
    info[0] = s.detInc; info[1] = s.detStart; info[2] = s.rotInc; info[3] = s.viewDiameter;
    info[4] = s.focalLength; info[5] = s.sdl; info[6] = s.fanBeamAngle;
    info[7] = s.viewDiameter / SQRT2; info[8] = p.xmin; info[9] = p.xmax; info[10] = p.ymin;
    info[11] = p.ymax; info[12] = s.detLen; info[13] = s.centerDetLength;
    info[14] = s.scanDiameter; info[15] = 0;
  }
  if (!filterName || nx <= 0) { free(p.e); return 0; }
  procsig_t ps;
  if (procsig_init(&ps, filterName, methodName, 1. / s.detInc, s.detInc, nDet, filterParam,
                   "spatial", generationName, zeropad, interpFactor, s.geom, s.focalLength,
                   s.sdl)) {
    procsig_free(&ps);
    free(p.e);
    return 3;
  }
  if (info) info[15] = ps.nfp;
  bp_t b;
  if (bp_init(&b, s.geom, nDet, s.detInc, nView, s.viewDiameter, s.focalLength, s.sdl,
              backprojName, interpName, nx, ny, NULL)) {
    procsig_free(&ps);
    free(p.e);
    return 4;
  }
  memset(image, 0, sizeof(float) * nx * ny);
  int nf = nDet * interpFactor;
  double* buf = (double*)calloc(nf, sizeof(double));
  double tf = 0, tb = 0;
  for (int v = 0; v < nView; v++) {
    double a = now_s();
    procsig_filter(&ps, sino + (size_t)v * nDet, buf);
    double c = now_s();
    if (filtered) memcpy(filtered + (size_t)v * nf, buf, sizeof(double) * nf);
    bp_view(&b, image, buf, angles[v]);
    tf += c - a;
    tb += now_s() - c;
  }
  bp_post(&b, image);
  if (imageRecon) memcpy(imageRecon, image, sizeof(float) * nx * ny);
  if (timings) { timings[0] = t1 - t0; timings[1] = tf; timings[2] = tb; }
  free(buf);
  bp_free(&b);
  procsig_free(&ps);
  free(p.e);
  return 0;
}

/* Filter + backproject the first nUse views of a given sinogram
 * (Reconstructor::reconstructView, reconstruct.cpp:156-239), timing the two
 * stages; the CPU-baseline leg of the benchmark when the reference library is
 * not available. */
int orc_fbp_sample(const int* types, const double* params, int nElem, const char* builtin,
                   const char* geometry, int nDet, int nView, int nSample, double rotLen,
                   double focalRatio, double centerRatio, const char* filterName,
                   const char* methodName, int zeropad, const char* interpName,
                   const char* backprojName, int nx, int ny, const float* sino,
                   const double* angles, int nUse, float* image, double* timings) {
  phantom_t p;
  scanner_t s;
  if (phm_build(&p, types, params, nElem, builtin)) return 1;
  if (scanner_init(&s, &p, geometry, nDet, nView, 0, nSample, rotLen, focalRatio, centerRatio,
                   1.0, 1.0)) {
    free(p.e);
    return 2;
  }
  procsig_t ps;
  if (procsig_init(&ps, filterName, methodName, 1. / s.detInc, s.detInc, nDet, -1.0, "spatial",
                   "direct", zeropad, 1, s.geom, s.focalLength, s.sdl)) {
    procsig_free(&ps);
    free(p.e);
    return 3;
  }
  bp_t b;
  if (bp_init(&b, s.geom, nDet, s.detInc, nView, s.viewDiameter, s.focalLength, s.sdl,
              backprojName, interpName, nx, ny, NULL)) {
    procsig_free(&ps);
    free(p.e);
    return 4;
  }
  memset(image, 0, sizeof(float) * nx * ny);
  double* buf = (double*)calloc(nDet, sizeof(double));
  double tf = 0, tb = 0;
  for (int v = 0; v < nUse; v++) {
    double a = now_s();
    procsig_filter(&ps, sino + (size_t)v * nDet, buf);
    double c = now_s();
    bp_view(&b, image, buf, angles[v]);
    tf += c - a;
    tb += now_s() - c;
  }
  bp_post(&b, image);
  timings[0] = tf;
  timings[1] = tb;
  free(buf);
  bp_free(&b);
  procsig_free(&ps);
  free(p.e);
  return 0;
}

/* ImageFile::comparativeStatistics, libctsim/imagefile.cpp:185-240 */
int orc_stats(int nx, int ny, const float* a, const float* b, double* dre) {
  if (nx == 0 || ny == 0) return 1;
  double mean = 0.;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) mean += a[(size_t)ix * ny + iy];
  mean /= ((unsigned)nx * (unsigned)ny);
  double sq = 0., ab = 0., sqm = 0., av = 0.;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++) {
      size_t k = (size_t)ix * ny + iy;
      double diff = a[k] - b[k];
      sq += diff * diff;
      ab += fabs(diff);
      double dm = a[k] - mean;
      sqm += dm * dm;
      av += fabs(a[k]);
    }
  dre[0] = sqrt(sq / sqm);
  dre[1] = ab / av;
  int hx = nx / 2, hy = ny / 2;
  double emax = -1;
  for (int ix = 0; ix < hx; ix++)
    for (int iy = 0; iy < hy; iy++) {
#define A(x, y) a[(size_t)(x) * ny + (y)]
#define B(x, y) b[(size_t)(x) * ny + (y)]
      double pa = 0.25 * (A(2 * ix, 2 * iy) + A(2 * ix + 1, 2 * iy) + A(2 * ix, 2 * iy + 1) +
                          A(2 * ix + 1, 2 * iy + 1));
      double pb = 0.25 * (B(2 * ix, 2 * iy) + B(2 * ix + 1, 2 * iy) + B(2 * ix, 2 * iy + 1) +
                          B(2 * ix + 1, 2 * iy + 1));
#undef A
#undef B
      double err = fabs(pa - pb);
      if (err > emax) emax = err;
    }
  dre[2] = emax;
  return 0;
}

/* ---------------------------------------------------------------- helical */

/* Fan angle of detector iDet for the helical routines.  Equiangular:
 * (iDet - (nDet-1)/2) * detInc with the reference's integer centre
 * (projections.cpp:279, 386).  Equilinear (no reference implementation; the
 * reference stops at projections.cpp:214-219): the same integer centre on the
 * flat detector at source distance sdl, gamma = atan(t / sdl). */
static double helix_gamma(int geom, int iDet, int nDet, double detInc, double sdl) {
  if (geom == 2) return atan((iDet - (nDet - 1) / 2) * detInc / sdl);
  return (iDet - (nDet - 1) / 2) * detInc;
}

/* Projections::Helical180LI_Equiangular, projections.cpp:231-358, as a
 * scatter in the reference's loop order into a zeroed float array.
 * in: [nView][nDet] float (views from 0), out: [(L+1)][nDet] with
 * L = int((pi + fan) / rotInc); anglesOut receives (i + offsetView) * rotInc.
 * geom: 1 equiangular (the reference), 2 equilinear (derived).  Returns L+1,
 * or -1 (not 2(pi + fan) of data) / -2 (interpView without pi + fan either side). */
int orc_helical180li(int geom, int nDet, int nView, double rotInc, double detInc,
                     double fanAngle, double sdl, int interpView, const float* in, float* out,
                     double* anglesOut) {
  const double dbeta = rotInc;
  int offsetView = 0;
  if (nView < (int)((2 * (PI + fanAngle)) / dbeta) - 1) {
    snprintf(g_err, sizeof g_err, "Data set does not include 360 +2*FanBeamAngle views");
    return -1;
  }
  if (interpView < 0) {
    interpView = (int)((PI + fanAngle) / dbeta);
  } else {
    if (interpView * dbeta < PI + fanAngle || interpView * dbeta + PI + fanAngle > nView * dbeta) {
      snprintf(g_err, sizeof g_err, "no pi+fanAngle of data on either side of the interpolation view");
      return -2;
    }
    offsetView = interpView - (int)((PI + fanAngle) / dbeta);
  }
  const int L = (int)((PI + fanAngle) / dbeta);
  for (int i = 0; i <= L; i++) {
    anglesOut[i] = (i + offsetView) * dbeta;
    for (int j = 0; j < nDet; j++) out[(size_t)i * nDet + j] = 0.f;
  }
  const int lastAcq = 2 * L;
  for (int iView = 0; iView <= lastAcq; iView++) {
    const double beta = iView * dbeta;
    for (int iDet = 0; iDet < nDet; iDet++) {
      const double gamma = helix_gamma(geom, iDet, nDet, detInc, sdl);
      int newiView, newiDet;
      if (beta < PI + fanAngle) {
        newiDet = iDet;
        newiView = iView;
      } else {
        newiDet = -iDet + (nDet - 1);
        newiView = nearest_int(((iView * dbeta) + 2 * gamma - PI) / dbeta);
      }
      /* the reference can index one view past its data at the far edge of
       * the last acquired view (projections.cpp:303); such samples count as 0 */
      if (iView + offsetView >= nView) continue;
      if (beta > fanAngle - 2 * gamma && beta < 2 * PI + fanAngle - 2 * gamma) {
        /* the mirrored view rounds to L + 1 when (pi + fan) / rotInc has a
         * fractional part >= 1/2; the reference then writes past its array
         * (projections.cpp:281, 304); such samples are dropped */
        if (newiView > L) continue;
        const float src = in[(size_t)(iView + offsetView) * nDet + iDet];
        float* dst = &out[(size_t)newiView * nDet + newiDet];
        if (beta > fanAngle - 2 * gamma && beta <= 2 * fanAngle)
          *dst += (beta + 2 * gamma - fanAngle) / (PI + 2 * gamma) * src;
        else if (beta > 2 * fanAngle && beta <= PI - 2 * gamma)
          *dst += (beta + 2 * gamma - fanAngle) / (PI + 2 * gamma) * src;
        else if (beta > PI - 2 * gamma && beta <= PI + fanAngle)
          *dst += (beta + 2 * gamma - fanAngle) / (PI + 2 * gamma) * src;
        else if (beta > PI + fanAngle && beta <= PI + 2 * fanAngle - 2 * gamma)
          *dst += (2 * PI - beta - 2 * gamma + fanAngle) / (PI - 2 * gamma) * src;
        else if (beta > PI + 2 * fanAngle - 2 * gamma && beta <= 2 * PI)
          *dst += (2 * PI - beta - 2 * gamma + fanAngle) / (PI - 2 * gamma) * src;
        else if (beta > 2 * PI && beta <= 2 * PI + fanAngle - 2 * gamma)
          *dst += (2 * PI - beta - 2 * gamma + fanAngle) / (PI - 2 * gamma) * src;
      }
    }
  }
  return L + 1;
}

/* Projections::HalfScanFeather, projections.cpp:369-439, in place on
 * data [nView][nDet] (nView must be int((pi + fan) / rotInc) + 1).  Returns
 * 0, or -1 (not a half-scan data set) / -4 (outside every weighting region). */
int orc_halfscan_feather(int geom, int nDet, int nView, double rotInc, double detInc,
                         double fanAngle, double sdl, float* data) {
  const double dbeta = rotInc;
  if (nView != (int)((PI + fanAngle) / dbeta) + 1) {
    snprintf(g_err, sizeof g_err, "not a half-scan data set");
    return -1;
  }
  for (int iView2 = 0; iView2 < nView; iView2++) {
    const double beta2 = iView2 * dbeta;
    for (int iDet2 = 0; iDet2 < nDet; iDet2++) {
      const double gamma2 = helix_gamma(geom, iDet2, nDet, detInc, sdl);
      if (beta2 >= PI - 2 * gamma2) {
        const int iDet1 = (nDet - 1) - iDet2;
        const int iView1 = nearest_int(((iView2 * dbeta) + 2 * gamma2 - PI) / dbeta);
        float* d2 = &data[(size_t)iView2 * nDet + iDet2];
        float* d1 = &data[(size_t)iView1 * nDet + iDet1];
        *d1 = *d2;
        const double beta1 = iView1 * dbeta;
        const double gamma1 = -gamma2;
        double x;
        if (beta1 <= (fanAngle - 2 * gamma1))
          x = beta1 / (fanAngle - 2 * gamma1);
        else if ((fanAngle - 2 * gamma1 <= beta1) && beta1 <= PI - 2 * gamma1)
          x = 1;
        else if ((PI - 2 * gamma1 <= beta1) && (beta1 <= PI + fanAngle))
          x = (PI + fanAngle - beta1) / (fanAngle + 2 * gamma1);
        else {
          snprintf(g_err, sizeof g_err, "Shouldn't be here!");
          return -4;
        }
        const double w1 = (3 * x - 2 * x * x) * x;
        const double w2 = 1 - w1;
        *d1 *= w1;
        *d2 *= w2;
      }
    }
  }
  const double scalefactor = nView * rotInc / PI;
  for (size_t k = 0; k < (size_t)nView * nDet; k++) data[k] *= scalefactor;
  return 0;
}

/* -------------------------------------------------------------- rebinning */

typedef struct {
  double t, theta, raysum;
  int index;
} RebinCoord;

/* ParallelRaysumCoordinate::compareByT (include/projections.h:50-52), with
 * the original (view-major) position breaking exact (t, theta) ties */
static int cmp_by_t(const void* pa, const void* pb) {
  const RebinCoord* a = (const RebinCoord*)pa;
  const RebinCoord* b = (const RebinCoord*)pb;
  if (a->t != b->t) return a->t < b->t ? -1 : 1;
  if (a->theta != b->theta) return a->theta < b->theta ? -1 : 1;
  return a->index - b->index;
}

/* LinearInterpolator<double>::interpolate with abscissae, zero outside off
 * (include/interpolator.h:203-262); the search hint does not change the
 * result for increasing abscissae and is not kept. */
static double lin_interp_x(const double* X, const double* Y, int n, double dX) {
  int iLower = -1, iUpper = n;
  while (iUpper - iLower > 1) {
    const int iMiddle = (iUpper + iLower) >> 1;
    if (dX >= X[iMiddle]) iLower = iMiddle;
    else iUpper = iMiddle;
  }
  if (dX == X[0]) return Y[0];
  if (dX < X[0]) return Y[0];
  if (dX == X[n - 1]) return Y[n - 1];
  if (dX > X[n - 1]) return Y[n - 1];
  return Y[iLower] + (Y[iUpper] - Y[iLower]) * ((dX - X[iLower]) / (X[iUpper] - X[iLower]));
}

/* Projections::interpolateToParallel, libctsim/projections.cpp:1250-1328,
 * with ParallelRaysums (1339-1469, theta normalised to [0, 2 pi)).
 * in/out [nView][nDet] float; info6 receives the new rotStart(0), rotInc,
 * detStart, detInc and the first and last new view angles. */
int orc_interpolate_to_parallel(int geom, int nDet, int nView, double detStart, double detInc,
                                double focalLength, double sdl, double viewDiameter,
                                double rotStart, const double* angles, const float* in,
                                float* out, double* anglesOut, double* info6) {
  if (geom == 0) {
    memcpy(out, in, sizeof(float) * (size_t)nDet * nView);
    memcpy(anglesOut, angles, sizeof(double) * nView);
    return 0;
  }
  const double newRotInc = TWOPI / nView;
  const double newDetStart = -viewDiameter / 2;
  double newDetInc = viewDiameter / nDet;
  if (nDet % 2 == 0) newDetInc = viewDiameter / (nDet - 1);
  const size_t nc = (size_t)nView * nDet;
  RebinCoord* c = (RebinCoord*)malloc(nc * sizeof(RebinCoord));
  size_t k = 0;
  for (int iV = 0; iV < nView; iV++) {
    const double dViewAngle = angles[iV];
    double dDetPos = detStart;
    for (int iD = 0; iD < nDet; iD++, k++) {
      if (geom == 2) {
        const double dFanAngle = atan(dDetPos / sdl);
        c[k].theta = dViewAngle + dFanAngle;
        c[k].t = focalLength * sin(dFanAngle);
      } else {
        c[k].theta = dViewAngle + dDetPos;
        c[k].t = focalLength * sin(dDetPos);
      }
      c[k].theta = normalize_angle(c[k].theta);
      c[k].raysum = in[k];
      c[k].index = (int)k;
      dDetPos += detInc;
    }
  }
  qsort(c, nc, sizeof(RebinCoord), cmp_by_t);
  double* X = (double*)malloc(sizeof(double) * (nView > nDet ? nView : nDet));
  double* Y = (double*)malloc(sizeof(double) * (nView > nDet ? nView : nDet));
  /* interpolate to evenly spaced theta (views) */
  for (int iD = 0; iD < nDet; iD++) {
    for (int i = 0; i < nView; i++) {
      X[i] = c[(size_t)iD * nView + i].theta;
      Y[i] = c[(size_t)iD * nView + i].raysum;
    }
    double dViewAngle = rotStart;
    for (int iV = 0; iV < nView; iV++, dViewAngle += newRotInc)
      out[(size_t)iV * nDet + iD] = (float)lin_interp_x(X, Y, nView, dViewAngle);
  }
  /* interpolate to evenly spaced t (detectors) */
  double* T = (double*)malloc(sizeof(double) * nDet);
  for (int i = 0; i < nDet; i++) T[i] = c[(size_t)i * nView].t;
  double dViewAngle = rotStart;
  for (int iV = 0; iV < nView; iV++, dViewAngle += newRotInc) {
    anglesOut[iV] = dViewAngle;
    for (int i = 0; i < nDet; i++) Y[i] = out[(size_t)iV * nDet + i];
    double dDetPos = newDetStart;
    for (int iD = 0; iD < nDet; iD++, dDetPos += newDetInc)
      out[(size_t)iV * nDet + iD] = (float)lin_interp_x(T, Y, nDet, dDetPos);
  }
  info6[0] = 0;
  info6[1] = newRotInc;
  info6[2] = newDetStart;
  info6[3] = newDetInc;
  info6[4] = anglesOut[0];
  info6[5] = anglesOut[nView - 1];
  free(T);
  free(X);
  free(Y);
  free(c);
  return 0;
}

/* ------------------------------------------------------------ rasteriser */

/* PhantomElement::isPointInside with PHM_COORD, phantom.cpp:913-975 */
static int point_inside(const elem_t* e, double x, double y) {
  xform_m(e->phmToObj, &x, &y);
  switch (e->type) {
    case PE_RECT:
      return !(x > 1. || x < -1. || y > 1. || y < -1.);
    case PE_TRI:
      return !(y < 0. || y > 1. - x || y > 1. + x);
    case PE_ELLIPSE:
      if (x > 1. || x < -1. || y > 1. || y < -1.) return 0;
      return !(x * x + y * y > 1.);
    case PE_SEGMENT:
      if (x > 1. || x < -1. || y > 0.) return 0;
      x *= e->u;
      y *= e->v;
      return !(x * x + (y - e->v) * (y - e->v) > e->u * e->u + e->v * e->v);
    case PE_SECTOR:
      if (x > 1. || x < -1. || y > 1.) return 0;
      if (y > 1. - x || y > 1. + x) return 0;
      x *= e->u;
      y *= e->v;
      return !(x * x + (y - e->v) * (y - e->v) > e->u * e->u + e->v * e->v);
  }
  return 0;
}

/* Phantom::convertToImagefile, phantom.cpp:463-538, whole image: img is
 * [nx][ny] column-major float (pixel (ix, iy) at ix*ny+iy); ext4 receives the
 * axis extent xmin, xmax, ymin, ymax.  Returns 0, or 1 when nx or ny < 2
 * (the reference leaves the image untouched). */
int orc_rasterize(const int* types, const double* params, int nElem, const char* builtin,
                  int nx, int ny, double viewRatio, int nsample, float* img, double* ext4) {
  if (nx < 2 || ny < 2) return 1;
  phantom_t p;
  if (phm_build(&p, types, params, nElem, builtin)) return 2;
  if (nsample < 1) nsample = 1;
  const double dx = p.xmax - p.xmin, dy = p.ymax - p.ymin;
  const double xcent = p.xmin + dx / 2, ycent = p.ymin + dy / 2;
  const double maxAxis = dx > dy ? dx : dy;
  const double dHalflen = viewRatio * ((SQRT2 * maxAxis) / SQRT2 / 2);
  const double xmin = xcent - dHalflen, xmax = xcent + dHalflen;
  const double ymin = ycent - dHalflen, ymax = ycent + dHalflen;
  const double xinc = (xmax - xmin) / nx, yinc = (ymax - ymin) / ny;
  const double kxinc = xinc / nsample, kyinc = yinc / nsample;
  const double kxofs = kxinc / 2, kyofs = kyinc / 2;
  ext4[0] = xmin; ext4[1] = xmax; ext4[2] = ymin; ext4[3] = ymax;
  for (size_t k = 0; k < (size_t)nx * ny; k++) img[k] = 0;
  for (int el = p.n - 1; el >= 0; el--) { /* the reference's list is newest first */
    const elem_t* e = &p.e[el];
    double x = xmin;
    for (int ix = 0; ix < nx; ix++, x += xinc) {
      double y = ymin;
      for (int iy = 0; iy < ny; iy++, y += yinc) {
        double dAtten = 0;
        double xi = x + kxofs;
        for (int kx = 0; kx < nsample; kx++, xi += kxinc) {
          double yi = y + kyofs;
          for (int ky = 0; ky < nsample; ky++, yi += kyinc)
            if (point_inside(e, xi, yi)) dAtten += e->atten;
        }
        img[(size_t)ix * ny + iy] += dAtten;
      }
    }
  }
  if (nsample > 1) {
    const double factor = 1.0 / (double)(nsample * nsample);
    for (size_t k = 0; k < (size_t)nx * ny; k++) img[k] *= factor;
  }
  free(p.e);
  return 0;
}
