// Device code of the mixed-precision path shared by the single-slice
// kernels (backproject.hip) and the batched persistent kernel (batch.hip):
// interval coefficients of the reference's interpolants, the fan-beam
// position and weight in FP32, and the tiled backprojection body.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "bp_plan.h"

namespace ctsim {
namespace dev {

constexpr int kMaxChunk = 64;
// LDS row stride (cells) of the pipelined symmetric body: its largest window.
constexpr int kPipeRow = 128;

// Remap block ids so that neighbouring tiles run on the same XCD.  The
// dispatcher deals workgroups round-robin over the 8 XCDs, so XCD k receives
// block ids k, k+8, ...; give it the contiguous tile range
// [k*q + min(k, r), ...) where n = 8q + r (a bijection on [0, n)).
__device__ __forceinline__ int xcd_tile(int bid, int n) {
  const int q = n >> 3, r = n & 7;
  const int k = bid & 7, local = bid >> 3;
  return k * q + min(k, r) + local;
}

#pragma clang fp contract(fast)

// The mixed path interpolates from per-interval coefficients: on detector
// interval i (position p in [i, i+1), u = p - i) the reference's interpolant
// is the polynomial a0 + u (a1 + u (a2 + u a3)), zero outside the detector,
// with its end rules folded in (CubicPolyInterpolator::interpolate,
// interpolator.cpp:36-75; linear with the IntDiff last-cell rule,
// backprojectors.cpp:635-638).  coef_kernel writes them once per view, in
// FP32, so the backprojection inner loop is a window lookup and three FMAs.
template <int INTERP>
using CoefT = typename std::conditional<INTERP == INTERP_CUBIC, float4, float2>::type;

// Store with the non-temporal hint (streaming: the line is not kept for reuse
// in this XCD's L2), for results only a later kernel reads.
template <class T>
__device__ __forceinline__ void store_stream(T* p, const T& v) {
  if constexpr (sizeof(T) == 16) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(__builtin_bit_cast(f4, v), reinterpret_cast<f4*>(p));
  } else if constexpr (sizeof(T) == 8) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(__builtin_bit_cast(f2, v), reinterpret_cast<f2*>(p));
  } else {
    __builtin_nontemporal_store(v, p);
  }
}

// Coefficients of detector interval i of one filtered row (n cells), read
// through row(j) (any index the rules below use is in [0, n)).
template <int INTERP, bool LAST_INCL, class Row>
__device__ __forceinline__ CoefT<INTERP> interval_coef(const Row& row, int i, int n) {
  // FP32 with fused multiply-adds: the coefficients carry a few units in the
  // last place of the row values, the order of the rows' own FP32 rounding
  if constexpr (INTERP == INTERP_CUBIC) {
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    if (i == 0) {
      if (n >= 2) { a0 = row(0); a1 = row(1) - row(0); }
    } else if (i + 2 > n) {
      // beyond the last full interval: zero
    } else if (i + 2 == n) {
      a0 = row(n - 2);
      a1 = row(n - 1) - row(n - 2);
    } else {
      const float y0 = row(i - 1), y1 = row(i), y2 = row(i + 1), y3 = row(i + 2);
      const float third = 1.f / 3, sixth = 1.f / 6;
      a0 = y1;
      a1 = __builtin_fmaf(-third, y0, __builtin_fmaf(-0.5f, y1, __builtin_fmaf(-sixth, y3, y2)));
      a2 = __builtin_fmaf(0.5f, y0 + y2, -y1);
      a3 = __builtin_fmaf(sixth, y3 - y0, 0.5f * (y1 - y2));
    }
    return make_float4(a0, a1, a2, a3);
  } else {
    const int hiOk = LAST_INCL ? n - 1 : n - 2;
    float a0 = 0.f, a1 = 0.f;
    if (i <= hiOk) {
      a0 = row(i);
      a1 = i < n - 1 ? row(i + 1) - row(i) : 0.f;
    }
    return make_float2(a0, a1);
  }
}

// Both rows of a view pair at once (row(j).x and row(j).y), each neighbour
// read once: the same values as interval_coef on each row, with the cases
// selected instead of branched on.
template <int INTERP, bool LAST_INCL, class Row2>
__device__ __forceinline__ void interval_coef2(const Row2& row, int i, int n, CoefT<INTERP>& ca,
                                               CoefT<INTERP>& cb) {
  const auto v1 = row(i), v2 = row(i + 1 < n ? i + 1 : n - 1);
  if constexpr (INTERP == INTERP_CUBIC) {
    const auto v0 = row(i > 0 ? i - 1 : 0), v3 = row(i + 2 < n ? i + 2 : n - 1);
    const bool interior = i > 0 && i + 2 < n;
    const bool edge = (i == 0 && n >= 2) || i + 2 == n;  // a0 = y1, a1 = y2 - y1
    auto one = [interior, edge](float y0, float y1, float y2, float y3) {
      const float third = 1.f / 3, sixth = 1.f / 6;
      const float b1 = __builtin_fmaf(-third, y0, __builtin_fmaf(-0.5f, y1, __builtin_fmaf(-sixth, y3, y2)));
      const float b2 = __builtin_fmaf(0.5f, y0 + y2, -y1);
      const float b3 = __builtin_fmaf(sixth, y3 - y0, 0.5f * (y1 - y2));
      const float d = y2 - y1;
      return make_float4(interior || edge ? y1 : 0.f, interior ? b1 : (edge ? d : 0.f), interior ? b2 : 0.f,
                         interior ? b3 : 0.f);
    };
    ca = one(v0.x, v1.x, v2.x, v3.x);
    cb = one(v0.y, v1.y, v2.y, v3.y);
  } else {
    const int hiOk = LAST_INCL ? n - 1 : n - 2;
    const bool ok = i <= hiOk, diff = ok && i < n - 1;
    ca = make_float2(ok ? v1.x : 0.f, diff ? v2.x - v1.x : 0.f);
    cb = make_float2(ok ? v1.y : 0.f, diff ? v2.y - v1.y : 0.f);
  }
}

struct MixedArgs {
  int nx, ny, nDet, nView, lanesY, tilesX, tilesY, W, chunk, split, viewsPerGroup;
  float center, F, twoF, posScale, endScale;
  float atan[9];
  const float* xs;
  const float* ys;
  const ViewConstF32* vc;
  const void* coef;
  float* image;     // split == 1: image += endScale * sum (overwrite: image = endScale * sum)
  float* partial;   // split > 1: partial[group][pixel] = sum
  int overwrite;
  // Rotational symmetry (bp_sym_tile): views v + s*Q (s < S) are view v
  // turned by s*K quarter turns; tiles cover the quadrant [0, half)^2;
  // capW coefficients of LDS; dirf = +1 (counter-clockwise steps) or -1.
  int Q, half, capW;
  float dirf;
  // Mirror body (bp_mirror_tile): tiles cover columns [0, domX) and rows
  // [0, domY); with two members, mapO = 1 when the second is O (else Y).
  int domX, domY, mapO;
};

// A pointer known to be the same in every lane of the wave, moved to scalar
// registers (the compiler cannot prove it for pointers loaded from memory).
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long u = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(u >> 32));
  return reinterpret_cast<T*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// Loads of one row of interval coefficients through a buffer resource: the
// hardware range check returns zeros for cells outside [0, n) (a negative
// cell's byte offset wraps to a large unsigned value), which is the zero
// padding the windows need, without a compare per cell.  The row base must be
// wave-uniform; n = 0 makes every load return zeros.
template <class CT>
struct RowLoad {
  __amdgpu_buffer_rsrc_t r;
  __device__ RowLoad(const void* coef, size_t rowStart, int n)
      : r(__builtin_amdgcn_make_buffer_rsrc(
            const_cast<CT*>(static_cast<const CT*>(coef) + rowStart), 0,
            n * static_cast<int>(sizeof(CT)), 0x00020000)) {}
  __device__ CT operator()(int cell) const {
    if constexpr (sizeof(CT) == 16)
      return __builtin_bit_cast(CT, __builtin_amdgcn_raw_buffer_load_b128(r, cell * 16, 0, 0));
    else
      return __builtin_bit_cast(CT, __builtin_amdgcn_raw_buffer_load_b64(r, cell * 8, 0, 0));
  }
};

// One window of staged coefficients in the symmetric kernel.
struct SymWin {
  float sizeM1;  // cells staged - 1 (the window-relative clamp)
  int size;      // cells staged
  int base;      // LDS offset of the window: cell lo + j of row v + s*Q at base + s*size + j
  float cshift;  // centre - lo
};

// atan on the whole line for FP32: odd polynomial t*P(t^2) on [0,1]
// (weighted least-squares fit of atan(t)/t in t^2, max abs error 1.3e-7),
// with atan(t) = pi/2 - atan(1/t) beyond 1.  Used when the fan is wider than
// the fitted range.
__device__ __forceinline__ float atan_f32(float t) {
  const float at = fabsf(t);
  const bool inv = at > 1.f;
  const float u = inv ? __builtin_amdgcn_rcpf(at) : at;
  const float z = u * u;
  float p = 0.0028499013278633356f;
  p = p * z - 0.01606869138777256f;
  p = p * z + 0.042691655457019806f;
  p = p * z - 0.07504309713840485f;
  p = p * z + 0.10640943795442581f;
  p = p * z - 0.14203648269176483f;
  p = p * z + 0.1999261975288391f;
  p = p * z - 0.3333307206630707f;
  p = p * z + 1.0f;
  float r = u * p;
  r = inv ? 1.5707963267948966f - r : r;
  return copysignf(r, t);
}

// Detector position of the equiangular geometry: atan(t) / detInc + add
// with t = tan(gamma).
template <int DEG>
__device__ __forceinline__ float fan_angle(const float* c, float t, float add) {
  if constexpr (DEG > 0) {
    const float z = t * t;
    float p = c[DEG];
#pragma unroll
    for (int k = DEG - 1; k >= 0; k--) p = fmaf(p, z, c[k]);
    return fmaf(t, p, add);
  } else {
    return fmaf(atan_f32(t), c[0], add);
  }
}

// Position of pixel (x, y) in one view (window-relative when cshift is the
// centre minus the window start) and its weight.  Parallel: ca, sa carry
// 1/detInc.  Fan: ca, sa = cos, sin of the view angle.
template <int GK, int DEG>
__device__ __forceinline__ void pos_mixed(const MixedArgs& a, float ca, float sa, float cshift,
                                          float x, float y, float r2mF2, float& pos, float& wgt) {
  if constexpr (GK == 0) {
    pos = fmaf(y, sa, fmaf(x, ca, cshift));
    wgt = 1.f;
  } else {
    // the x terms are common to a lane's rows: x ca and F + x sa once per view
    const float rc = fmaf(y, sa, x * ca);             // r cos(beta - phi)
    const float fl = fmaf(-y, ca, fmaf(x, sa, a.F));  // F + r sin(beta - phi)
    if constexpr (GK == 1) {
      // one reciprocal for both 1/fl and 1/L^2, L^2 = fl^2 + rc^2 = 2 F fl + r^2 - F^2
      const float L2 = fmaf(a.twoF, fl, r2mF2);
      const float rr = __builtin_amdgcn_rcpf(fl * L2);
      pos = fan_angle<DEG>(a.atan, rc * (L2 * rr), cshift);
      wgt = fl * rr;
    } else {
      const float q = __builtin_amdgcn_rcpf(fl);
      pos = fmaf(rc * q, a.posScale, cshift);
      const float u = a.F * q;
      wgt = u * u;
    }
  }
}

// One workgroup's work item: a tile of (256 / lanesY) columns x
// (lanesY * P) rows and one group of views.  s_v must hold kMaxChunk + 2
// entries.  Lane (cx, ly) owns column ix0 + cx and rows
// iy0 + P*ly .. + P-1 (contiguous, so the final read-modify-write is a
// vector access).  Per chunk of views: lanes < nv compute each view's window
// start from the tile corners, the workgroup stages the window's interval
// coefficients in LDS, then every lane walks the chunk's views for its P
// pixels, summing the chunk in FP32 and folding the chunk sum into a
// Kahan-compensated total.
template <int GK, int INTERP, int P, int DEG>
__device__ __forceinline__ void bp_mixed_tile(const MixedArgs& a, int tile, int grp,
                                              CoefT<INTERP>* s_w, float4* s_v, int* s_lo) {
  using CT = CoefT<INTERP>;
  const int t = threadIdx.x;
  const int tx = 256 / a.lanesY, ty = a.lanesY * P;
  const int ix0 = (tile / a.tilesY) * tx, iy0 = (tile % a.tilesY) * ty;
  const int ix = ix0 + t / a.lanesY, iyb = iy0 + (t % a.lanesY) * P;
  const int cx1 = min(ix0 + tx - 1, a.nx - 1), cy1 = min(iy0 + ty - 1, a.ny - 1);
  const int W = a.W;
  const float x = a.xs[min(ix, a.nx - 1)];
  float y[P], r2mF2[P];
#pragma unroll
  for (int k = 0; k < P; k++) {
    y[k] = a.ys[min(iyb + k, a.ny - 1)];
    r2mF2[k] = fmaf(x, x, fmaf(y[k], y[k], -a.F * a.F));
  }
  const int vBegin = grp * a.viewsPerGroup;
  const int vEnd = min(a.nView, vBegin + a.viewsPerGroup);
  float sum[P], comp[P];
#pragma unroll
  for (int k = 0; k < P; k++) sum[k] = comp[k] = 0.f;
  const CT* coef = static_cast<const CT*>(a.coef);

  for (int v0 = vBegin; v0 < vEnd; v0 += a.chunk) {
    const int nv = min(a.chunk, vEnd - v0);
    if (t < nv) {
      const ViewConstF32 c = a.vc[v0 + t];
      int lo = -2;
      if (a.W < a.nDet + 4) {
        float p0, p1, p2, p3, wd;
        const float xa = a.xs[ix0], xb = a.xs[cx1], ya = a.ys[iy0], yb = a.ys[cy1];
        pos_mixed<GK, DEG>(a, c.ca, c.sa, a.center, xa, ya, fmaf(xa, xa, fmaf(ya, ya, -a.F * a.F)), p0, wd);
        pos_mixed<GK, DEG>(a, c.ca, c.sa, a.center, xb, ya, fmaf(xb, xb, fmaf(ya, ya, -a.F * a.F)), p1, wd);
        pos_mixed<GK, DEG>(a, c.ca, c.sa, a.center, xa, yb, fmaf(xa, xa, fmaf(yb, yb, -a.F * a.F)), p2, wd);
        pos_mixed<GK, DEG>(a, c.ca, c.sa, a.center, xb, yb, fmaf(xb, xb, fmaf(yb, yb, -a.F * a.F)), p3, wd);
        float pmin = fminf(fminf(p0, p1), fminf(p2, p3));
        if (!(pmin > -1e9f && pmin < 1e9f)) pmin = 0.f;
        lo = static_cast<int>(floorf(pmin)) - 2;
      }
      s_lo[t] = lo;
      s_v[t] = make_float4(c.ca, c.sa, a.center - static_cast<float>(lo), 0.f);
    }
    __syncthreads();
    // stage the windows, four views and two cells per lane at a time (eight
    // loads in flight; cells outside the detector read as zero)
    for (int vq = 0; vq < nv; vq += 4) {
      for (int jb = 0; jb < W; jb += 512) {
        CT val[4][2];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int vv = vq + u;
          const RowLoad<CT> row(coef, static_cast<size_t>(v0 + min(vv, nv - 1)) * a.nDet, vv < nv ? a.nDet : 0);
          const int lo = s_lo[min(vv, nv - 1)];
          val[u][0] = row(lo + jb + t);
          val[u][1] = row(lo + jb + t + 256);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const int j = jb + t + 256 * h;
            if (vq + u < nv && j < W) s_w[(vq + u) * W + j] = val[u][h];
          }
      }
    }
    __syncthreads();
    float part[P];
#pragma unroll
    for (int k = 0; k < P; k++) part[k] = 0.f;
    // Two views per iteration in three phases (all positions, then all
    // window reads, then all interpolations), so 2P independent LDS reads are
    // in flight together; the next pair's constants are read ahead (s_v holds
    // kMaxChunk + 2 entries, so the read-ahead stays in bounds).
    float4 c0 = s_v[0], c1 = s_v[1];
    const float wmax = static_cast<float>(W - 1);
    int vv = 0;
    for (; vv + 1 < nv; vv += 2) {
      const float4 n0 = s_v[vv + 2], n1 = s_v[vv + 3];
      float u[2 * P], wg[2 * P];
      int j[2 * P];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const float4& c = h ? c1 : c0;
#pragma unroll
        for (int k = 0; k < P; k++) {
          float pos;
          pos_mixed<GK, DEG>(a, c.x, c.y, c.z, x, y[k], r2mF2[k], pos, wg[h * P + k]);
          // interval index clamped in float (one med3; the clamp keeps NaN in
          // range too), truncated (non-negative), offset within the interval
          u[h * P + k] = __builtin_amdgcn_fractf(pos);
          j[h * P + k] = (vv + h) * W + static_cast<int>(__builtin_amdgcn_fmed3f(pos, 0.f, wmax));
        }
      }
      CT q[2 * P];
#pragma unroll
      for (int i = 0; i < 2 * P; i++) q[i] = s_w[j[i]];
#pragma unroll
      for (int i = 0; i < 2 * P; i++) {
        float val;
        if constexpr (INTERP == INTERP_CUBIC)
          val = fmaf(u[i], fmaf(u[i], fmaf(u[i], q[i].w, q[i].z), q[i].y), q[i].x);
        else
          val = fmaf(u[i], q[i].y, q[i].x);
        part[i % P] = fmaf(val, wg[i], part[i % P]);  // view vv before view vv+1
      }
      c0 = n0;
      c1 = n1;
    }
    if (vv < nv) {
      const CT* w = s_w + vv * W;
#pragma unroll
      for (int k = 0; k < P; k++) {
        float pos, wgt;
        pos_mixed<GK, DEG>(a, c0.x, c0.y, c0.z, x, y[k], r2mF2[k], pos, wgt);
        const float uu = __builtin_amdgcn_fractf(pos);
        const CT qq = w[static_cast<int>(__builtin_amdgcn_fmed3f(pos, 0.f, wmax))];
        float val;
        if constexpr (INTERP == INTERP_CUBIC) val = fmaf(uu, fmaf(uu, fmaf(uu, qq.w, qq.z), qq.y), qq.x);
        else val = fmaf(uu, qq.y, qq.x);
        part[k] = fmaf(val, wgt, part[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < P; k++) {  // Kahan step with the chunk's sum
      const float yk = part[k] - comp[k];
      const float tk = sum[k] + yk;
      comp[k] = (tk - sum[k]) - yk;
      sum[k] = tk;
    }
    __syncthreads();
  }
  if (ix >= a.nx) return;
  const size_t base = static_cast<size_t>(ix) * a.ny + iyb;
  if (a.split == 1) {
#pragma unroll
    for (int k = 0; k < P; k++)
      if (iyb + k < a.ny)
        a.image[base + k] = a.overwrite ? a.endScale * (sum[k] - comp[k])
                                        : fmaf(a.endScale, sum[k] - comp[k], a.image[base + k]);
  } else {
    float* out = a.partial + static_cast<size_t>(grp) * a.nx * a.ny;
#pragma unroll
    for (int k = 0; k < P; k++)
      if (iyb + k < a.ny) out[base + k] = sum[k] - comp[k];
  }
}

// The mirror body's staging buffers: two static LDS arrays of this many
// bytes (separate objects, so that the compiler sees that reads of one do
// not wait for the direct-to-LDS loads into the other).
constexpr int kMirBufBytes = 16 * 1024;

// Staging of one window of W cells in the mirror body: the direct-to-LDS
// load unit (16 bytes) is one cubic cell or two linear cells (a pair
// starting at an even cell).  Units per window:
__host__ __device__ constexpr int mirror_window_units(int W, bool cubic) { return cubic ? W : (W + 1) / 2; }

typedef __attribute__((address_space(3))) void lds_void;

// The staging buffers and view-block records of the mirror body, shared by
// all of its instantiations in a kernel: per block of up to 64 base views,
// the view constants (ca, sa, centre - lo) and, per window (view vv, member m) at vv * NM + m, the
// detector cell of its slot 0 (low word) and its row * 2 + reflected (high
// word; slot j then holds cell first + j, or first - j).  Plain arrays indexed from their own base,
// so that the compiler can tell every LDS access's object.
static __shared__ __attribute__((aligned(16))) unsigned char s_mirBuf0[kMirBufBytes];
static __shared__ __attribute__((aligned(16))) unsigned char s_mirBuf1[kMirBufBytes];
static __shared__ float4 s_mirView[64];
static __shared__ long long s_mirSeg[64 * 4];
// reflected flags of the block's windows: bit 4 (vv % 16) + m of word vv / 16
static __shared__ unsigned long long s_mirRefl[4];

// The mirror body: the plain body with NM updates per position.  Member 0 is
// the lane's pixel p in view v; member m > 0 is the mirror image M_m p in the
// partner view vc[v].mir[m - 1] (NM = 4: members Y, X, O; NM = 2: Y, or O
// when mapO).  M_m p sees the partner's detector at p's position or at its
// reflection 2c - pos about the centre cell c, so each partner window is
// staged so that its slot j holds the interval that p's position falls in
// when it falls in slot j of view v's window [lo, lo + W): cell lo + j of
// the partner row, or, reflected, cell 2c - 1 - lo - j, read at fraction
// 1 - u.  Every member thus reads the slot of the base pixel.
// Base views come in orbit order (vc[s].order), so a view group reads only
// its own views' rows.  Tiles cover columns [0, domX) (and rows [0, domY)
// with four members); a member whose mirror image is the pixel itself (the
// middle column or row of an odd size) or another member's is not stored,
// so every pixel takes every view once.
// Windows are found for 64 base views at a time (four lanes per view, one
// tile corner each).  Within a block, staging is software-pipelined: the
// windows of the next chunk of views (a power of two, so chunks never
// straddle a block) go straight from global memory to one LDS buffer (one
// flat copy, 256 units a round; cells outside the detector read as zero by
// the buffer range check) while the lanes interpolate the current chunk from
// the other.  The two buffers are separate objects used at fixed call sites
// (chunks unrolled by two), so the compiler sees that reads of one do not
// wait for the copy into the other.  The FP32 sums are folded into
// Kahan-compensated totals every 8 or more views.
template <int GK, int INTERP, int P, int DEG, int NM>
__device__ __forceinline__ void bp_mirror_tile(const MixedArgs& a, int tile, int grp) {
  using CT = CoefT<INTERP>;
  static_assert(NM == 2 || NM == 4, "two or four members");
  CT* const s_buf0 = reinterpret_cast<CT*>(s_mirBuf0);
  CT* const s_buf1 = reinterpret_cast<CT*>(s_mirBuf1);
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int tx = 256 / a.lanesY, ty = a.lanesY * P;
  const int ix0 = (tile / a.tilesY) * tx, iy0 = (tile % a.tilesY) * ty;
  const int ix = ix0 + t / a.lanesY, iyb = iy0 + (t % a.lanesY) * P;
  const int cx1 = min(ix0 + tx - 1, a.domX - 1), cy1 = min(iy0 + ty - 1, a.domY - 1);
  const int W = a.W;
  const float x = a.xs[min(ix, a.domX - 1)];
  float y[P], r2mF2[P];
#pragma unroll
  for (int k = 0; k < P; k++) {
    y[k] = a.ys[min(iyb + k, a.domY - 1)];
    r2mF2[k] = fmaf(x, x, fmaf(y[k], y[k], -a.F * a.F));
  }
  const int vBegin = grp * a.viewsPerGroup;
  const int vEnd = min(a.nView, vBegin + a.viewsPerGroup);
  float sum[NM][P], comp[NM][P], part[NM][P];
#pragma unroll
  for (int m = 0; m < NM; m++)
#pragma unroll
    for (int k = 0; k < P; k++) sum[m][k] = comp[m][k] = part[m][k] = 0.f;
  const int nDet = __builtin_amdgcn_readfirstlane(a.nDet);
  const int twoC = static_cast<int>(2.f * a.center);  // the centre is a whole cell
  // one buffer resource over every row (the host keeps nView * nDet cells
  // below 2^31 bytes); offsets at or beyond it read zero
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<CT*>(uniform_ptr(static_cast<const CT*>(a.coef))), 0,
      __builtin_amdgcn_readfirstlane(a.nView * nDet * static_cast<int>(sizeof(CT))), 0x00020000);
  // staging: windows of upw 16-byte units packed one after another (slot
  // stride Wp cells; linear cells in pairs from an even cell); lane t's
  // first unit is unit k0 of window seg0 and each round steps 256 units
  constexpr bool kCubic = INTERP == INTERP_CUBIC;
  const int upw = mirror_window_units(W, kCubic);
  const int Wp = kCubic ? upw : 2 * upw;
  const int seg0 = t / upw, k0 = t % upw, qU = 256 / upw, rU = 256 % upw;
  constexpr int kPairFlip = kCubic ? 0 : 1;  // a reflected pair holds its two slots swapped
  const int lane = t & 63;

  // windows of base views [vb, vb + nb), nb <= 64: lane (vv, q) = (t / 4,
  // t % 4) takes tile corner q; the four lanes' minimum gives lo
  auto meta = [&](int vb, int nb) {
    const int vv = t >> 2, q = t & 3;
    bool refl = false;
    if (vv < nb) {
      const int view = a.vc[vb + vv].order;  // base views in orbit order
      const ViewConstF32 c = a.vc[view];
      int lo = -2;
      if (a.W < a.nDet + 4) {
        const float cx = (q & 1) ? a.xs[cx1] : a.xs[ix0], cy = (q & 2) ? a.ys[cy1] : a.ys[iy0];
        float pos, wd;
        pos_mixed<GK, DEG>(a, c.ca, c.sa, a.center, cx, cy, fmaf(cx, cx, fmaf(cy, cy, -a.F * a.F)), pos, wd);
        pos = fminf(pos, __shfl_xor(pos, 1, 64));
        pos = fminf(pos, __shfl_xor(pos, 2, 64));
        if (!(pos > -1e9f && pos < 1e9f)) pos = 0.f;
        lo = static_cast<int>(floorf(pos)) - 2;
        if (!kCubic) lo &= ~1;  // pairs start at even cells (rounds down)
      }
      if (q == 0) s_mirView[vv] = make_float4(c.ca, c.sa, a.center - static_cast<float>(lo), 0.f);
      if (q < NM) {
        const int e = q == 0 ? 2 * view : c.mir[q - 1];
        const unsigned first = static_cast<unsigned>((e & 1) ? twoC - 1 - lo : lo);
        s_mirSeg[vv * NM + q] = static_cast<long long>((static_cast<unsigned long long>(e) << 32) | first);
        refl = (e & 1) != 0;
      }
    }
    // wave w holds views 16 w ... 16 w + 15 (lane 4 (vv % 16) + q)
    const unsigned long long bits = __ballot(refl);
    if (lane == 0) s_mirRefl[wave] = bits;
  };
  // the flat copy of windows [s0, s0 + nseg) of the block into buf: unit
  // f = r * 256 + t of round r is unit f % upw of window f / upw and lands at
  // byte 16 f (a wave's 64 units are contiguous).  Slot j of a direct window
  // holds cell first + j, of a reflected one cell first - j (a reflected
  // pair lands with its two slots swapped, read at j ^ 1).  With an odd nDet
  // the pair (nDet - 1, nDet) straddles the row end: its lane copies the
  // cell's two words and two zero words alone, one 4-byte copy each.
  constexpr int kRounds = kMirBufBytes / (256 * 16);
  auto issue = [&](int s0, int nseg, CT* buf) {
    const int total = nseg * upw;
    int seg = seg0, k = k0;
#pragma unroll
    for (int r = 0; r < kRounds; r++) {
      if (r * 256 >= total) break;
      unsigned off = 0xFFFFFFF0u;  // beyond the resource: reads zero
      int edge = -1;               // byte offset of cell nDet - 1 of a straddling pair
      if (r * 256 + t < total) {
        const long long rec = s_mirSeg[s0 + seg];
        const int first = static_cast<int>(rec), e = static_cast<int>(rec >> 32);
        const bool refl = (e & 1) != 0;
        const int rowBase = (e >> 1) * nDet;
        if (kCubic) {
          const int cm = refl ? first - k : first + k;
          if (static_cast<unsigned>(cm) < static_cast<unsigned>(nDet))
            off = static_cast<unsigned>((rowBase + cm) * 16);
        } else {
          const int cm = refl ? first - 2 * k - 1 : first + 2 * k;  // the pair's lower cell
          if (cm >= 0 && cm + 1 < nDet) off = static_cast<unsigned>((rowBase + cm) * 8);
          else if (cm + 1 == nDet) edge = (rowBase + cm) * 8;
        }
      }
      unsigned char* base = reinterpret_cast<unsigned char*>(buf) + (r * 256 + wave * 64) * 16;
      if (kCubic || edge < 0)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)base, 16, static_cast<int>(off), 0, 0, 0);
      if constexpr (!kCubic) {
        // straddling pairs, one lane at a time: lane l writes at M0 + 4 l
        unsigned long long todo = __ballot(edge >= 0);
        while (todo) {
          const int l = static_cast<int>(__builtin_ctzll(todo));
          todo &= todo - 1;
          const int eo = __builtin_amdgcn_readlane(edge, l);
          if (lane == l) {
            unsigned char* d = base + 16 * l - 4 * l;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)d, 4, eo, 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)(d + 4), 4, eo + 4, 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)(d + 8), 4, static_cast<int>(0xFFFFFFF0u), 0, 0, 0);
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_void*)(d + 12), 4, static_cast<int>(0xFFFFFFF0u), 0, 0, 0);
          }
        }
      }
      seg += qU;
      k += rU;
      if (k >= upw) {
        k -= upw;
        seg++;
      }
    }
  };
  // views [b0, b0 + nv) of block M from buf (window vv * NM + m of the
  // chunk at (vv * NM + m) * W)
  auto compute = [&](int b0, int nv, const CT* buf) {
    const float wmax = static_cast<float>(W - 1);
    // the chunk's reflected flags (a chunk of at most 32 views starts at a
    // multiple of its size, so two words hold it)
    const int w0 = b0 >> 4;
    // (readfirstlane returns a signed int: widen through unsigned)
    auto uniform64 = [](unsigned long long v) {
      const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(v >> 32)));
      const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(v)));
      return (static_cast<unsigned long long>(hi) << 32) | lo;
    };
    const unsigned long long fl0 = uniform64(s_mirRefl[w0]), fl1 = uniform64(s_mirRefl[min(w0 + 1, 3)]);
    for (int vv = 0; vv < nv; vv++) {
      const int bv = b0 + vv;
      const unsigned vflags = static_cast<unsigned>(((bv >> 4) == w0 ? fl0 : fl1) >> (4 * (bv & 15)));
      // (a uniform index through readfirstlane and no read-ahead: no pointer
      // induction variable, whose object the compiler could not tell)
      const float4 c = s_mirView[__builtin_amdgcn_readfirstlane(b0 + vv)];
      float u[P], u1[P], wg[P];
      int j[P];
#pragma unroll
      for (int k = 0; k < P; k++) {
        float pos;
        pos_mixed<GK, DEG>(a, c.x, c.y, c.z, x, y[k], r2mF2[k], pos, wg[k]);
        u[k] = __builtin_amdgcn_fractf(pos);
        u1[k] = 1.f - u[k];
        j[k] = static_cast<int>(__builtin_amdgcn_fmed3f(pos, 0.f, wmax));
      }
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const int wb = __builtin_amdgcn_readfirstlane((vv * NM + m) * Wp);
        CT q[P];
        if ((vflags >> m) & 1) {  // reflected window: the interval's polynomial at 1 - u
#pragma unroll
          for (int k = 0; k < P; k++) q[k] = buf[wb + (j[k] ^ kPairFlip)];
#pragma unroll
          for (int k = 0; k < P; k++) {
            float val;
            if constexpr (INTERP == INTERP_CUBIC)
              val = fmaf(u1[k], fmaf(u1[k], fmaf(u1[k], q[k].w, q[k].z), q[k].y), q[k].x);
            else
              val = fmaf(u1[k], q[k].y, q[k].x);
            part[m][k] = fmaf(val, wg[k], part[m][k]);
          }
        } else {
#pragma unroll
          for (int k = 0; k < P; k++) q[k] = buf[wb + j[k]];
#pragma unroll
          for (int k = 0; k < P; k++) {
            float val;
            if constexpr (INTERP == INTERP_CUBIC)
              val = fmaf(u[k], fmaf(u[k], fmaf(u[k], q[k].w, q[k].z), q[k].y), q[k].x);
            else
              val = fmaf(u[k], q[k].y, q[k].x);
            part[m][k] = fmaf(val, wg[k], part[m][k]);
          }
        }
      }
    }
  };
  auto kahan = [&]() {
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
      for (int k = 0; k < P; k++) {
        const float yk = part[m][k] - comp[m][k];
        const float tk = sum[m][k] + yk;
        comp[m][k] = (tk - sum[m][k]) - yk;
        sum[m][k] = tk;
        part[m][k] = 0.f;
      }
  };
  // one pipeline step: chunk (v0, nv) of the block at vb from cur while the
  // block's next chunk goes to nxt; false when the block has no next chunk
  int v0 = vBegin, nv = 0, vb = vBegin, pending = 0;
  auto step = [&](const CT* cur, CT* nxt) -> bool {
    const int v1 = v0 + nv, nv1 = v1 - vb < 64 ? min(a.chunk, vEnd - v1) : 0;
    if (nv1 > 0) issue((v1 - vb) * NM, nv1 * NM, nxt);
    compute(v0 - vb, nv, cur);
    pending += nv;
    if (pending >= 8 || v1 >= vEnd) {
      kahan();
      pending = 0;
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
    __syncthreads();
    v0 = v1;
    nv = nv1;
    return nv1 > 0;
  };
  for (vb = vBegin; vb < vEnd; vb += 64) {
    // the block's windows, then its first chunk (not overlapped)
    meta(vb, min(64, vEnd - vb));
    __syncthreads();
    v0 = vb;
    nv = min(a.chunk, vEnd - vb);
    issue(0, nv * NM, s_buf0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    while (step(s_buf0, s_buf1) && step(s_buf1, s_buf0)) {
    }
  }
  if (ix >= a.domX) return;
  float* out = a.split == 1 ? a.image : a.partial + static_cast<size_t>(grp) * a.nx * a.ny;
  const int mx = a.nx - 1 - ix;
#pragma unroll
  for (int m = 0; m < NM; m++) {
    const bool flipX = NM == 4 ? (m == 1 || m == 3) : m == 1;
    const bool flipY = NM == 4 ? m >= 2 : (m == 1 && a.mapO);
    if (flipX && mx == ix) continue;  // the middle column: member 0 has it
    const int col = flipX ? mx : ix;
#pragma unroll
    for (int k = 0; k < P; k++) {
      const int iy = iyb + k;
      if (iy >= a.domY) break;
      const int row = flipY ? a.ny - 1 - iy : iy;
      if (NM == 4 && flipY && row == iy) continue;  // the middle row
      const size_t idx = static_cast<size_t>(col) * a.ny + row;
      const float v = sum[m][k] - comp[m][k];
      if (a.split > 1) out[idx] = v;
      else out[idx] = a.overwrite ? a.endScale * v : fmaf(a.endScale, v, out[idx]);
    }
  }
}

// Position and weight of orbit member m (pixel R^m (x, y), R the quarter
// turn (x, y) -> (-y, x)) in one view, from A = x ca + y sa and
// B = x sa - y ca of the base pixel: the member's r cos and F + r sin are
// +-A, +-B and F -+ A, F -+ B.
template <int GK, int DEG, int M>
__device__ __forceinline__ void pos_member(const MixedArgs& a, float A, float B, float r2mF2,
                                           float cshift, float& pos, float& wgt) {
  const float rc = M == 0 ? A : (M == 1 ? B : (M == 2 ? -A : -B));
  if constexpr (GK == 0) {
    pos = rc + cshift;
    wgt = 1.f;
  } else {
    const float fl = M == 0 ? a.F + B : (M == 1 ? a.F - A : (M == 2 ? a.F - B : a.F + A));
    if constexpr (GK == 1) {
      const float L2 = fmaf(a.twoF, fl, r2mF2);
      const float rr = __builtin_amdgcn_rcpf(fl * L2);
      pos = fan_angle<DEG>(a.atan, rc * (L2 * rr), cshift);
      wgt = fl * rr;
    } else {
      const float q = __builtin_amdgcn_rcpf(fl);
      pos = fmaf(rc * q, a.posScale, cshift);
      const float u = a.F * q;
      wgt = u * u;
    }
  }
}

// The symmetric body.  A pixel p of the quadrant and its images R p, R^2 p,
// R^3 p form an orbit; when views v + s*Q are view v turned by s*K quarter
// turns (s < S), pixel R^t p in view v + s*Q sees the detector exactly where
// R^(t - sK) p sees it in view v.  So per base view v < Q a lane evaluates
// four positions and weights (one per orbit member) and each serves S
// updates: member m in row v + s*Q lands in orbit pixel (m + sK) mod 4.
// With dirf = -1 the steps turn clockwise and the odd members swap roles.
// Per chunk of base views, wave 0 computes each member tile's window
// (corner positions, two cells of margin), lays the windows out in LDS by a
// prefix sum and takes as many base views as fit in capW coefficients (one
// always fits: capW >= 4*S*W); the workgroup stages rows v + s*Q of each
// window (rows of one window adjacent; interleaving them so that a member's
// S reads differ only in an immediate offset saves an add per update but
// spreads a wave's reads over four times the banks, and measured slower);
// every lane then accumulates its 4P pixels in FP32 and folds the sum into
// Kahan-compensated totals every 8 or more base views.
template <int GK, int INTERP, int P, int DEG, int S, int K>
__device__ __forceinline__ void bp_sym_tile(const MixedArgs& a, int tile, int grp,
                                            CoefT<INTERP>* s_w, float2* s_ab, SymWin* s_win,
                                            int* s_n) {
  using CT = CoefT<INTERP>;
  static_assert(kMaxChunk >= 16, "window table holds 16 base views");
  const int t = threadIdx.x;
  const int h = a.half, n = 2 * a.half;
  const int tx = 256 / a.lanesY, ty = a.lanesY * P;
  const int ix0 = (tile / a.tilesY) * tx, iy0 = (tile % a.tilesY) * ty;
  const int ix = ix0 + t / a.lanesY, iyb = iy0 + (t % a.lanesY) * P;
  const int cx1 = min(ix0 + tx - 1, h - 1), cy1 = min(iy0 + ty - 1, h - 1);
  const float x = a.xs[min(ix, h - 1)];
  const float F2 = a.F * a.F;
  float y[P], r2mF2[P];
#pragma unroll
  for (int k = 0; k < P; k++) {
    y[k] = a.ys[min(iyb + k, h - 1)];
    r2mF2[k] = fmaf(x, x, fmaf(y[k], y[k], -F2));
  }
  const int vBegin = grp * a.viewsPerGroup;
  const int vEnd = min(a.Q, vBegin + a.viewsPerGroup);
  const int maxChunk = min(a.chunk, 16);
  // part: FP32 sums of up to about 8 base views (32 updates per pixel when
  // S = 4), folded into the Kahan-compensated sum, comp
  float sum[4][P], comp[4][P], part[4][P];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int k = 0; k < P; k++) sum[m][k] = comp[m][k] = part[m][k] = 0.f;
  int pending = 0;
  const CT* coef = static_cast<const CT*>(a.coef);
  const float dirf = a.dirf;

  for (int v0 = vBegin; v0 < vEnd;) {
    const int cmax = min(maxChunk, vEnd - v0);
    if (t < 64) {
      // wave 0: lane (vv, m) = (t / 4, t % 4) computes member m's window in
      // base view v0 + vv, then the lanes scan the staged sizes
      const int vv = t >> 2, m = t & 3;
      int lo = -2, size = 0;
      if (t < 4 * cmax) {
        const ViewConstF32 c = a.vc[v0 + vv];
        if (t % 4 == 0) s_ab[vv] = make_float2(c.ca, c.sa);
        size = a.W;
        if (a.W < a.nDet + 4) {
          const float xa = a.xs[ix0], xb = a.xs[cx1], ya = a.ys[iy0], yb = a.ys[cy1];
          const float sg = (m & 1) ? dirf : 1.f;
          float pmin = 3.0e38f, pmax = -3.0e38f;
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const float cx = (q & 1) ? xb : xa, cy = (q & 2) ? yb : ya;
            const float A = sg * fmaf(cy, c.sa, cx * c.ca), B = sg * fmaf(-cy, c.ca, cx * c.sa);
            const float r2 = fmaf(cx, cx, fmaf(cy, cy, -F2));
            float pos, w;
            switch (m) {
              case 0: pos_member<GK, DEG, 0>(a, A, B, r2, a.center, pos, w); break;
              case 1: pos_member<GK, DEG, 1>(a, A, B, r2, a.center, pos, w); break;
              case 2: pos_member<GK, DEG, 2>(a, A, B, r2, a.center, pos, w); break;
              default: pos_member<GK, DEG, 3>(a, A, B, r2, a.center, pos, w); break;
            }
            pmin = fminf(pmin, pos);
            pmax = fmaxf(pmax, pos);
          }
          if (pmin > -1e9f && pmax < 1e9f) {
            lo = static_cast<int>(floorf(pmin)) - 2;
            size = min(static_cast<int>(floorf(pmax)) + 3 - lo, a.W);
          } else {
            lo = static_cast<int>(floorf(pmin > -1e9f && pmin < 1e9f ? pmin : 0.f)) - 2;
          }
        }
      }
      const int need = S * size;
      int cum = need;  // inclusive prefix sum over the wave
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(cum, d, 64);
        if (t >= d) cum += o;
      }
      // base view vv fits when the total through its last member does; the
      // totals grow with the lane, so the fitting views are a prefix
      const unsigned long long fit = __ballot(m == 3 && t < 4 * cmax && cum <= a.capW);
      if (t < 4 * cmax) {
        SymWin w;
        w.sizeM1 = static_cast<float>(size - 1);
        w.size = size;
        w.base = cum - need;
        w.cshift = a.center - static_cast<float>(lo);
        s_win[t] = w;
      }
      if (t == 0) *s_n = max(1, __popcll(fit));
    }
    __syncthreads();
    const int nv = __builtin_amdgcn_readfirstlane(*s_n);
    {  // stage: wave w takes segments (vv, m, s) = w, w + 4, ..., two at a
       // time with two cells per lane each, so four loads are in flight
      const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
      const int nseg = nv * 4 * S;
      for (int seg = wave; seg < nseg; seg += 8) {
        const bool two = seg + 4 < nseg;
        const SymWin w0 = s_win[seg / S];
        const SymWin w1 = s_win[two ? (seg + 4) / S : seg / S];
        // first cell staged: lo = centre - cshift (integers below 2^24, exact)
        const int lo0 = __builtin_amdgcn_readfirstlane(static_cast<int>(a.center - w0.cshift));
        const int n0 = __builtin_amdgcn_readfirstlane(w0.size);
        const int lo1 = __builtin_amdgcn_readfirstlane(static_cast<int>(a.center - w1.cshift));
        const int n1 = two ? __builtin_amdgcn_readfirstlane(w1.size) : 0;
        const RowLoad<CT> r0(coef, static_cast<size_t>(v0 + (seg / S >> 2) + (seg % S) * a.Q) * a.nDet, a.nDet);
        const RowLoad<CT> r1(coef, static_cast<size_t>(v0 + ((seg + 4) / S >> 2) + ((seg + 4) % S) * a.Q) * a.nDet,
                             two ? a.nDet : 0);
        CT* d0 = s_w + __builtin_amdgcn_readfirstlane(w0.base) + (seg % S) * n0;
        CT* d1 = s_w + __builtin_amdgcn_readfirstlane(w1.base) + ((seg + 4) % S) * n1;
        for (int j0 = 0; j0 < max(n0, n1); j0 += 128) {
          const int ja = j0 + lane, jb = ja + 64;
          const CT a0 = r0(lo0 + ja), a1 = r0(lo0 + jb), b0 = r1(lo1 + ja), b1 = r1(lo1 + jb);
          if (ja < n0) d0[ja] = a0;
          if (jb < n0) d0[jb] = a1;
          if (ja < n1) d1[ja] = b0;
          if (jb < n1) d1[jb] = b1;
        }
      }
    }
    __syncthreads();
    for (int vv = 0; vv < nv; vv++) {
      const float2 ab = s_ab[vv];
      const float xca = x * ab.x, xsa = x * ab.y;
      SymWin w[4];
#pragma unroll
      for (int m = 0; m < 4; m++) w[m] = s_win[vv * 4 + m];
#pragma unroll
      for (int k = 0; k < P; k++) {
        const float A = fmaf(y[k], ab.y, xca), B = fmaf(-y[k], ab.x, xsa);
        const float Ao = dirf * A, Bo = dirf * B;
        float u[4], wg[4];
        int j[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
          float pos;
          const float Am = (m & 1) ? Ao : A, Bm = (m & 1) ? Bo : B;
          if (m == 0) pos_member<GK, DEG, 0>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else if (m == 1) pos_member<GK, DEG, 1>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else if (m == 2) pos_member<GK, DEG, 2>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else pos_member<GK, DEG, 3>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          u[m] = __builtin_amdgcn_fractf(pos);
          j[m] = w[m].base +
                 static_cast<int>(__builtin_amdgcn_fmed3f(pos, 0.f, w[m].sizeM1));
        }
#pragma unroll
        for (int s = 0; s < S; s++) {
          CT q[4];
#pragma unroll
          for (int m = 0; m < 4; m++) q[m] = s_w[j[m] + s * w[m].size];
#pragma unroll
          for (int m = 0; m < 4; m++) {
            float val;
            if constexpr (INTERP == INTERP_CUBIC)
              val = fmaf(u[m], fmaf(u[m], fmaf(u[m], q[m].w, q[m].z), q[m].y), q[m].x);
            else
              val = fmaf(u[m], q[m].y, q[m].x);
            const int tgt = (m + s * K) & 3;
            part[tgt][k] = fmaf(val, wg[m], part[tgt][k]);
          }
        }
      }
    }
    v0 += nv;
    pending += nv;
    if (pending >= 8 || v0 >= vEnd) {  // Kahan step with the last few chunks' sum
#pragma unroll
      for (int m = 0; m < 4; m++)
#pragma unroll
        for (int k = 0; k < P; k++) {
          const float yk = part[m][k] - comp[m][k];
          const float tk = sum[m][k] + yk;
          comp[m][k] = (tk - sum[m][k]) - yk;
          sum[m][k] = tk;
          part[m][k] = 0.f;
        }
      pending = 0;
    }
    __syncthreads();
  }
  if (ix >= h) return;
  float* out = a.split == 1 ? a.image : a.partial + static_cast<size_t>(grp) * n * n;
#pragma unroll
  for (int k = 0; k < P; k++) {
    const int iy = iyb + k;
    if (iy >= h) break;
    // orbit pixels (column, row): R (i, j) = (n-1-j, i); clockwise steps
    // swap the images of R and R^3
    const int c1 = dirf > 0.f ? n - 1 - iy : iy, r1 = dirf > 0.f ? ix : n - 1 - ix;
    const size_t idx[4] = {static_cast<size_t>(ix) * n + iy, static_cast<size_t>(c1) * n + r1,
                           static_cast<size_t>(n - 1 - ix) * n + (n - 1 - iy),
                           static_cast<size_t>(n - 1 - c1) * n + (n - 1 - r1)};
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const float v = sum[m][k] - comp[m][k];
      if (a.split > 1) out[idx[m]] = v;
      else out[idx[m]] = a.overwrite ? a.endScale * v : fmaf(a.endScale, v, out[idx[m]]);
    }
  }
}

// Window of orbit member m of a base tile (corners xa..xb, ya..yb) in one
// view: the first cell and the number of cells (at most W) its pixels can
// read, with two cells of margin on each side.
template <int GK, int DEG>
__device__ __forceinline__ void member_window(const MixedArgs& a, int m, float ca, float sa, float xa,
                                              float xb, float ya, float yb, float dirf, int& lo,
                                              int& size) {
  lo = -2;
  size = a.W;
  if (a.W >= a.nDet + 4) return;  // full rows
  const float F2 = a.F * a.F;
  const float sg = (m & 1) ? dirf : 1.f;
  float pmin = 3.0e38f, pmax = -3.0e38f;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const float cx = (q & 1) ? xb : xa, cy = (q & 2) ? yb : ya;
    const float A = sg * fmaf(cy, sa, cx * ca), B = sg * fmaf(-cy, ca, cx * sa);
    const float r2 = fmaf(cx, cx, fmaf(cy, cy, -F2));
    float pos, w;
    switch (m) {
      case 0: pos_member<GK, DEG, 0>(a, A, B, r2, a.center, pos, w); break;
      case 1: pos_member<GK, DEG, 1>(a, A, B, r2, a.center, pos, w); break;
      case 2: pos_member<GK, DEG, 2>(a, A, B, r2, a.center, pos, w); break;
      default: pos_member<GK, DEG, 3>(a, A, B, r2, a.center, pos, w); break;
    }
    pmin = fminf(pmin, pos);
    pmax = fmaxf(pmax, pos);
  }
  if (pmin > -1e9f && pmax < 1e9f) {
    lo = static_cast<int>(floorf(pmin)) - 2;
    size = min(static_cast<int>(floorf(pmax)) + 3 - lo, a.W);
  } else {
    lo = static_cast<int>(floorf(pmin > -1e9f && pmin < 1e9f ? pmin : 0.f)) - 2;
  }
}

// The symmetric body of bp_sym_tile, software-pipelined one base view deep:
// while the lanes interpolate base view v from LDS, the global loads of
// view v + 1's windows are in flight into registers; they go to LDS after a
// barrier.  Wave w stages member w's S rows (at most 128 cells each, two per
// lane), so each LDS slot (member m, row s) sits at a fixed offset
// (m S + s) kPipeRow and needs no layout pass; windows and view constants
// are computed for 64 base views at a time, one (view, member) per lane.
// The row stride is the compile-time maximum window (128 cells), so every
// LDS read of an update is one address plus an immediate offset; the LDS
// holds 4 S kPipeRow coefficients (capW = 4 S kPipeRow, W <= kPipeRow).
template <int GK, int INTERP, int P, int DEG, int S, int K>
__device__ __forceinline__ void bp_sym_tile_pipe(const MixedArgs& a, int tile, int grp,
                                                 CoefT<INTERP>* s_w, float2* s_ab, SymWin* s_win) {
  using CT = CoefT<INTERP>;
  const int t = threadIdx.x;
  const int h = a.half, n = 2 * a.half;
  const int tx = 256 / a.lanesY, ty = a.lanesY * P;
  const int ix0 = (tile / a.tilesY) * tx, iy0 = (tile % a.tilesY) * ty;
  const int ix = ix0 + t / a.lanesY, iyb = iy0 + (t % a.lanesY) * P;
  const int cx1 = min(ix0 + tx - 1, h - 1), cy1 = min(iy0 + ty - 1, h - 1);
  const float xa = a.xs[ix0], xb = a.xs[cx1], ya = a.ys[iy0], yb = a.ys[cy1];
  const float x = a.xs[min(ix, h - 1)];
  const float F2 = a.F * a.F;
  float y[P], r2mF2[P];
#pragma unroll
  for (int k = 0; k < P; k++) {
    y[k] = a.ys[min(iyb + k, h - 1)];
    r2mF2[k] = fmaf(x, x, fmaf(y[k], y[k], -F2));
  }
  const int vBegin = grp * a.viewsPerGroup;
  const int vEnd = min(a.Q, vBegin + a.viewsPerGroup);
  float sum[4][P], comp[4][P], part[4][P];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int k = 0; k < P; k++) sum[m][k] = comp[m][k] = part[m][k] = 0.f;
  int pending = 0;
  const float dirf = a.dirf;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  CT* slot = s_w + wave * S * kPipeRow;  // this wave's member rows

  // global loads of member `wave`'s S rows of base view v0 + vv; the row
  // addresses are wave-uniform, and said so, so that the buffer descriptors
  // stay in scalar registers when the arguments come from memory (a batch
  // slice table) rather than from the kernel arguments
  const CT* coefU = uniform_ptr(static_cast<const CT*>(a.coef));
  const int nDetU = __builtin_amdgcn_readfirstlane(a.nDet), QU = __builtin_amdgcn_readfirstlane(a.Q);
  auto fetch = [&](int vv, int v0, CT (&r)[S][2]) {
    const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(a.center - s_win[vv * 4 + wave].cshift));
#pragma unroll
    for (int s = 0; s < S; s++) {
      const int rowIdx = __builtin_amdgcn_readfirstlane(v0 + vv + s * QU);
      const RowLoad<CT> row(coefU, static_cast<size_t>(rowIdx) * nDetU, nDetU);
      r[s][0] = row(lo + lane);
      r[s][1] = row(lo + lane + 64);
    }
  };
  auto place = [&](int vv, const CT (&r)[S][2]) {
    const int size = __builtin_amdgcn_readfirstlane(s_win[vv * 4 + wave].size);
#pragma unroll
    for (int s = 0; s < S; s++) {
      if (lane < size) slot[s * kPipeRow + lane] = r[s][0];
      if (lane + 64 < size) slot[s * kPipeRow + lane + 64] = r[s][1];
    }
  };

  for (int v0 = vBegin; v0 < vEnd; v0 += 64) {
    const int nb = min(64, vEnd - v0);
    __syncthreads();  // the previous block's readers of s_win and s_w are done
    if (t < 4 * nb) {
      const int vv = t >> 2, m = t & 3;
      const ViewConstF32 c = a.vc[v0 + vv];
      if (m == 0) s_ab[vv] = make_float2(c.ca, c.sa);
      int lo, size;
      member_window<GK, DEG>(a, m, c.ca, c.sa, xa, xb, ya, yb, dirf, lo, size);
      SymWin w;
      w.sizeM1 = static_cast<float>(size - 1);
      w.size = size;
      w.base = 0;
      w.cshift = a.center - static_cast<float>(lo);
      s_win[t] = w;
    }
    __syncthreads();
    CT r[S][2];
    fetch(0, v0, r);
    place(0, r);
    __syncthreads();
    for (int vv = 0; vv < nb; vv++) {
      const bool more = vv + 1 < nb;
      if (more) fetch(vv + 1, v0, r);  // in flight during the interpolation below
      const float2 ab = s_ab[vv];
      const float xca = x * ab.x, xsa = x * ab.y;
      SymWin w[4];
#pragma unroll
      for (int m = 0; m < 4; m++) w[m] = s_win[vv * 4 + m];
#pragma unroll
      for (int k = 0; k < P; k++) {
        const float A = fmaf(y[k], ab.y, xca), B = fmaf(-y[k], ab.x, xsa);
        const float Ao = dirf * A, Bo = dirf * B;
        float u[4], wg[4];
        int j[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
          float pos;
          const float Am = (m & 1) ? Ao : A, Bm = (m & 1) ? Bo : B;
          if (m == 0) pos_member<GK, DEG, 0>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else if (m == 1) pos_member<GK, DEG, 1>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else if (m == 2) pos_member<GK, DEG, 2>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          else pos_member<GK, DEG, 3>(a, Am, Bm, r2mF2[k], w[m].cshift, pos, wg[m]);
          u[m] = __builtin_amdgcn_fractf(pos);
          j[m] = m * S * kPipeRow + static_cast<int>(__builtin_amdgcn_fmed3f(pos, 0.f, w[m].sizeM1));
        }
#pragma unroll
        for (int s = 0; s < S; s++) {
          CT q[4];
#pragma unroll
          for (int m = 0; m < 4; m++) q[m] = s_w[j[m] + s * kPipeRow];
#pragma unroll
          for (int m = 0; m < 4; m++) {
            float val;
            if constexpr (INTERP == INTERP_CUBIC)
              val = fmaf(u[m], fmaf(u[m], fmaf(u[m], q[m].w, q[m].z), q[m].y), q[m].x);
            else
              val = fmaf(u[m], q[m].y, q[m].x);
            const int tgt = (m + s * K) & 3;
            part[tgt][k] = fmaf(val, wg[m], part[tgt][k]);
          }
        }
      }
      if (++pending >= 8 || vv + 1 == nb) {  // Kahan step with the last few views' sum
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
          for (int k = 0; k < P; k++) {
            const float yk = part[m][k] - comp[m][k];
            const float tk = sum[m][k] + yk;
            comp[m][k] = (tk - sum[m][k]) - yk;
            sum[m][k] = tk;
            part[m][k] = 0.f;
          }
        pending = 0;
      }
      __syncthreads();  // every lane is done reading view vv
      if (more) {
        place(vv + 1, r);
        __syncthreads();
      }
    }
  }
  if (ix < h) {
    float* out = a.split == 1 ? a.image : a.partial + static_cast<size_t>(grp) * n * n;
#pragma unroll
    for (int k = 0; k < P; k++) {
      const int iy = iyb + k;
      if (iy >= h) break;
      const int c1 = dirf > 0.f ? n - 1 - iy : iy, r1 = dirf > 0.f ? ix : n - 1 - ix;
      const size_t idx[4] = {static_cast<size_t>(ix) * n + iy, static_cast<size_t>(c1) * n + r1,
                             static_cast<size_t>(n - 1 - ix) * n + (n - 1 - iy),
                             static_cast<size_t>(n - 1 - c1) * n + (n - 1 - r1)};
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const float v = sum[m][k] - comp[m][k];
        if (a.split > 1) store_stream(&out[idx[m]], v);
        else out[idx[m]] = a.overwrite ? a.endScale * v : fmaf(a.endScale, v, out[idx[m]]);
      }
    }
  }
}

#pragma clang fp contract(off)

}  // namespace dev
}  // namespace ctsim
