// Backprojector plan: host-side geometry of one (Projections, ImageFile,
// backprojector, interpolation) combination and its device tables.
#pragma once

#include "ctsim_internal.h"

namespace ctsim {

// Which kernel body a plan runs (the reference's class, with the fan-beam
// geometries split by geometry).
enum BpKind {
  KIND_TRIG = 0,
  KIND_TABLE = 1,
  KIND_DIFF = 2,
  KIND_IDIFF = 3,
  KIND_EQUIANGULAR = 4,
  KIND_EQUILINEAR = 5,
};

// Per-view constants of the FP64 path (one struct per view, 32 bytes).
struct ViewConstF64 {
  double theta;     // view angle (Trig, Table, fan)
  double a, b;      // Diff: det_dx, det_dy       IntDiff: bit patterns of int64 det_dx, det_dy
  double c;         // Diff: detPosColStart       IntDiff: bit pattern of int64 colStart
};

// Per-view constants of the mixed path (32 bytes).  mir: the view's partners
// under the pixel-grid mirrors of the mirror body (MirrorViews), each
// partner view * 2 + 1 when the partner sees the mirrored pixel at the
// reflected detector position, + 0 when at the same position.  order: the
// mirror body takes its base views in the order vc[0].order, vc[1].order,
// ..., which lists each view's partners next to it, so that a view group
// holds whole orbits and reads no other group's rows.
struct ViewConstF32 {
  float ca, sa;     // parallel: cos/detInc, sin/detInc; fan: cos, sin
  int mir[3];
  int order;
  float pad[2];
};

// Launch shape of the mixed kernel: a workgroup of 256 lanes covers
// (256 / lanesY) columns x (lanesY * P) rows; views are split into `split`
// groups (separate workgroups whose partial sums are combined in a fixed
// order) and staged `chunk` views at a time.
// sym = 1 selects the rotationally symmetric kernel when the image and the
// views allow it (P of 1 or 2; views staged up to chunk base views at a time
// in symKB KB of LDS, 0: 62 KB); sym = 2 its software-pipelined form (one
// base view at a time, windows of at most 128 cells); sym = 3 the mirror
// body (the plain body's staging, 2 or 4 updates per position) when the
// views have mirror partners.
struct MixedShape {
  int lanesY = 16, P = 4, split = 1, chunk = 16, W = 0, sym = 0, symKB = 0;
};

// Rotational symmetry of a view set: views v + s*Q (s < S, Q = nView / S)
// are view v turned by s*K quarter turns, counter-clockwise (dir = 1) or
// clockwise (dir = -1); S = 0 when the views have no such symmetry.
// NM, mapO: the mirror members of the same views (MirrorViews).
struct SymViews {
  int S = 0, K = 0;
  float dir = 1.f;
  int NM = 1, mapO = 0;
};

// Mirror symmetry of a view set.  The mirrors Y (x -> -x), X (y -> -y) and
// O (both) map the pixel grid onto itself when it is centred.  A pixel's
// mirror image M p in a partner view sees the detector where p sees it in
// the view itself, reflected about the centre cell or not:
//   fan beam:  Y with theta' = -theta, X with pi - theta (reflected),
//              O with theta + pi (same position);
//   parallel:  Y with -theta (reflected) or pi - theta (same), X with
//              pi - theta (reflected) or -theta (same), O with theta itself
//              (reflected) or theta + pi (same).
// NM = 4: every view has Y, X and O partners (mir[0..2]); NM = 2: every view
// has a Y partner (mapO = 0) or else an O partner (mapO = 1), in mir[0];
// NM = 1: no mirror body.
struct MirrorViews {
  int NM = 1, mapO = 0;
};

struct BpPlan {
  int geom = 0, family = 0, kind = 0, interp = 0, interpFactor = 1;
  int nDet = 0, nView = 0, nx = 0, ny = 0, iDetCenter = 0;
  double detInc = 0, rotScale = 0, xMin = 0, xMax = 0, yMin = 0, yMax = 0, xInc = 0, yInc = 0;
  double F = 0, sdl = 0, start_r = 0, start_phi = 0, viewDiameter = 0;
  // Host tables (built exactly as the reference builds them).
  std::vector<double> r, phi;       // [nx*ny] column-major, Table/fan/Trig
  std::vector<float> xs, ys;        // pixel-centre coordinates (mixed path)
  // Device copies.
  DevBuf d_rphi;                    // r then phi, doubles
  DevBuf d_xy;                      // xs then ys, floats
  // Launch shape.
  int tileX = 16, tileY = 16;       // pixels per block along ix, iy
  int windowF64 = 0, windowF32 = 0; // detector window per (tile, view); 0 = full row
  bool smallTan = false;            // fan beam: |tan gamma| <= 1 for every pixel
  // Mixed equiangular path: atan(t) * posScale = t * P(t^2), P of degree
  // atanDeg (0: range-reduced generic polynomial) fitted on [0, tan^2 gmax].
  int atanDeg = 0;
  float atanCoef[9] = {};
  // Square, even-sized pixel grid centred on the origin: the quarter turn
  // maps pixel centres onto pixel centres.
  bool symImage = false;
  // xs (ys) symmetric about 0: the Y (X) mirror maps pixel centres onto
  // pixel centres.
  bool mirX = false, mirY = false;
  MixedShape mixed;                 // current launch shape (tuned or default)
  bool mixedTuned = false;
  // Per-view constants of the last angle set, kept on the device so repeated
  // reconstructions with the same views upload nothing.
  mutable std::vector<double> cachedAngles;
  mutable int cachedKind = -1;      // 0: FP64 constants, 1: FP32 constants
  mutable DevBuf d_vc;
  mutable SymViews cachedSym;       // symmetry of the cached angles (FP32 constants)
};

// Host side (bp_host.cpp).
int bp_plan_init(BpPlan& p, int geometry, int nDet, double detInc, int nView,
                 double viewDiameter, double F, double sdl, const char* backprojName,
                 const char* interpName, int interpFactor, int nx, int ny, const double* roi);
void bp_view_consts_f64(const BpPlan& p, const double* angles, int nView, ViewConstF64* out);
// Also fills the mirror partners (mir) and returns which mirror members the
// views allow on this plan's grid.
MirrorViews bp_view_consts_f32(const BpPlan& p, const double* angles, int nView, ViewConstF32* out);
int bp_window_bound(const BpPlan& p, int tileX, int tileY);
SymViews bp_sym_views(const double* angles, int nView);
MixedShape bp_default_shape(const BpPlan& p, int numCUs);

// Device side (backproject.hip).
int bp_upload_tables(BpPlan& p);
int bp_launch(ctsim_ctx* ctx, const BpPlan& p, int precision, const void* d_filtered,
              const double* h_angles, int nView, float* d_image);
int bp_launch_post(ctsim_ctx* ctx, const BpPlan& p, float* d_image);
int bp_launch_mixed(ctsim_ctx* ctx, const BpPlan& p, const MixedShape& m, const float* d_filtered,
                    const ViewConstF32* d_vc, int nView, float* d_image);
int bp_mixed_coef(ctsim_ctx* ctx, const BpPlan& p, const float* d_filtered, int nView);
// The same into a caller buffer of nView * nDet coefficient records.
int bp_mixed_coef_into(ctsim_ctx* ctx, const BpPlan& p, const float* d_filtered, int nView,
                       void* d_coef);
// views: the symmetry of these angles (NULL: the plan's cached one);
// partial: view-group partial images, split * nx * ny floats (NULL: the
// context's scratch).
int bp_mixed_tiles(ctsim_ctx* ctx, const BpPlan& p, const MixedShape& m, const void* coef,
                   const ViewConstF32* d_vc, int nView, float* d_image, double endScale,
                   bool overwrite, const SymViews* views = nullptr, float* partial = nullptr);
int bp_view_consts_mixed(ctsim_ctx* ctx, const BpPlan& p, const double* h_angles, int nView,
                         const ViewConstF32** out);
// True when the mixed precision of this plan runs the tiled coefficient
// kernel (not the FP64 index path) and one view's staging fits LDS.
bool bp_mixed_tiled(const BpPlan& p);
bool bp_sym_ok(const BpPlan& p, const MixedShape& m, int nView);
bool bp_sym_ok_for(const BpPlan& p, const MixedShape& m, const SymViews& sv);
// The mirror body runs for shape sym = 3 when the views have mirror partners
// (NM >= 2): the views per chunk it stages (NM windows each, in one of its
// two LDS buffers), the largest power of two up to min(m.chunk, 32) that fits; 0
// when one view does not fit or the coefficient rows span 2 GiB or more.
int bp_mirror_chunk(const BpPlan& p, const MixedShape& m, int NM, int nView);
size_t sym_lds(const MixedShape& m);  // LDS bytes the symmetric kernel stages into

}  // namespace ctsim

struct ctsim_backprojector {
  ctsim::BpPlan plan;
};
