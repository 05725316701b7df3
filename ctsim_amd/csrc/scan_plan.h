// Phantom and Scanner plans: host-side geometry of the ray-sum stage.
#pragma once

#include "ctsim_internal.h"

namespace ctsim {

// Element types, include/phantom.h:33-38.
enum ElemType { PE_INVALID = 0, PE_RECT = 1, PE_TRI = 2, PE_ELLIPSE = 3, PE_SECTOR = 4, PE_SEGMENT = 5 };

// One phantom element as the kernels read it (160 bytes).  The affine maps
// hold the reference's 3x3 row-vector matrices (libctsupport/xform.cpp:
// x' = x m00 + y m10 + m20, y' = x m01 + y m11 + m21) as
// {m00, m01, m10, m11, m20, m21}.
struct ElemGPU {
  double lim[4];  // polygon-outline bounding box: xmin, ymin, xmax, ymax
  double p2o[6];  // phantom -> normalised element coordinates
  double o2p[6];  // normalised element -> phantom coordinates
  double u, v, atten;
  int type;
  int pad;
};

// An element's definition as PhantomElement keeps it (phantom.h:41-128):
// what Phantom::print and printDefinitions report.
struct ElemInfo {
  int type;
  int nOutline;  // outline points (PhantomElement::nOutlinePoints)
  double cx, cy, u, v, rotRad, atten;
  double rotDeg;  // the rotation as defined
};

struct PhantomPlan {
  std::vector<ElemGPU> elems;  // definition order (the reference's list is newest-first)
  std::vector<ElemInfo> infos; // same order as elems
  double xmin = 1E30, xmax = -1E30, ymin = 1E30, ymax = -1E30;
  bool unitPulse = false;
  DevBuf d_elems;
  bool dirty = true;
};

struct ScannerPlan {
  int geom = 0, nDet = 0, nView = 0, offsetView = 0, startView = 0;
  unsigned nSample = 1;
  double viewDiameter = 0, focalLength = 0, centerDetLength = 0, sdl = 0, scanDiameter = 0;
  double xc = 0, yc = 0, rotLen = 0, rotInc = 0, detLen = 0, detStart = 0, detInc = 0;
  double fanBeamAngle = 0, angDetLen = 0, angDetInc = 0;
  double xs1 = 0, ys1 = 0, xs2 = 0, ys2 = 0, xd1 = 0, yd1 = 0, xd2 = 0, yd2 = 0;
  double rotIncM[3][3];
};

// Per-view endpoints after the incremental rotations of
// Scanner::collectProjections (scanner.cpp:302-419).
struct ViewEnds {
  double xd1, yd1, xd2, yd2, xs1, ys1, xs2, ys2, detAngle, viewAngle;
};

int phantom_add(PhantomPlan& p, const char* type, double cx, double cy, double u, double v,
                double rotDeg, double atten);
int phantom_builtin(PhantomPlan& p, const char* name);
int scanner_init(ScannerPlan& s, const PhantomPlan& p, const char* geometry, int nDet, int nView,
                 int offsetView, int nSample, double rotLen, double focalRatio, double centerRatio,
                 double viewRatio, double scanRatio);
void scanner_view_ends(const ScannerPlan& s, int iStartView, int iNumViews, int iOffsetView,
                       ViewEnds* out);

int raysum_launch(ctsim_ctx* ctx, const ScannerPlan& s, PhantomPlan& p, int precision,
                  int iStartView, int iNumViews, int iOffsetView, int iStorageOffset,
                  float* d_sino, double* h_angles);

}  // namespace ctsim

struct ctsim_phantom {
  ctsim::PhantomPlan plan;
};
struct ctsim_scanner {
  ctsim::ScannerPlan plan;
};
