// crmath.h -- correctly rounded FP64 cos, sin, atan and atan2.
//
// The FP64 backprojectors and ray sums evaluate cos, sin, atan and atan2 per
// pixel or per ray where the reference calls the C library.  These
// routines return the correctly rounded (round-to-nearest) value, so the
// device result is a property of the argument alone, not of a vendor
// library; where the host C library is correctly rounded too, device and
// host agree bit for bit.
//
// Method (Ziv, "Fast evaluation of elementary mathematical functions with
// correctly rounded last bit", ACM TOMS 17(3), 1991, as used by CORE-MATH):
//   - argument reduction: Cody-Waite with pi/2 split into four parts (the
//     first two of 33 bits, so k*C1 and k*C2 are exact for |k| < 2^20), the
//     reduced argument kept as a double-double (Dekker 1971 / Knuth
//     TwoSum and an FMA TwoProduct) with a running bound of its error;
//   - Tang's table-driven evaluation (P. T. P. Tang, "Table-lookup
//     algorithms for elementary functions and their error analysis", 1991):
//     sin/cos(j/128 + t) from tabulated double-double sin(j/128),
//     cos(j/128) and Taylor series of t, |t| <= 2^-8; atan(u) = atan(j/128)
//     + atan(z), z = (u - j/128) / (1 + u j/128), |z| <= 2^-8;
//   - stage 1 evaluates the leading products exactly and the rest in
//     double, about 2^-68 relative; Ziv's rounding test accepts its result
//     when every value within the error bound rounds to the same double;
//   - otherwise stage 2 evaluates everything in double-double (about 2^-96
//     relative plus the reduction's bound) and tests again.  When that test
//     cannot decide either (the exact value lies within ~2^-96 of a
//     rounding boundary: probability about 2^-42 per call), the
//     double-double's rounding is returned and crm_undecided() is called.
//
// Domain: cos/sin of |x| < 2^20 * pi/2 (1.6e6); larger arguments, which no
// caller in this library produces, fall back to the platform's cos/sin.
// atan and atan2 take every double.  Special values follow C99.  One
// documented exception: an atan2 result in the subnormal range is the
// rounded quotient y/x, which is one subnormal too large when y/x lies
// exactly halfway between two subnormals.
//
// The same source compiles for the device (hipcc, FMA = v_fma_f64) and for
// the host (g++ -ffp-contract=off, for the CPU checks against a 113-bit
// reference and the C library); both must run without contraction of
// a*b+c, which the pragmas below enforce under clang.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CRM_FN static __device__ __forceinline__
#define CRM_TABLE static __constant__ const
#else
#define CRM_FN static inline
#define CRM_TABLE static const
#endif

// Hooks for the CPU checks: stage 2 entered, and stage 2 undecided.
#ifndef CRM_STAGE2
#define CRM_STAGE2() ((void)0)
#endif
#ifndef CRM_UNDECIDED
#define CRM_UNDECIDED() ((void)0)
#endif

#include "crmath_tables.h"

#if defined(__clang__)
#define CRM_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define CRM_NO_CONTRACT
#endif

struct crm_dd {
  double h, l;
};

// Exact sum a + b = h + l.
CRM_FN crm_dd crm_two_sum(double a, double b) {
  CRM_NO_CONTRACT
  const double s = a + b;
  const double bp = s - a;
  return {s, (a - (s - bp)) + (b - bp)};
}

// Exact sum when |a| >= |b| (or a == 0).
CRM_FN crm_dd crm_fast_two_sum(double a, double b) {
  CRM_NO_CONTRACT
  const double s = a + b;
  return {s, b - (s - a)};
}

// Exact product a * b = h + l.
CRM_FN crm_dd crm_two_prod(double a, double b) {
  CRM_NO_CONTRACT
  const double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}

CRM_FN crm_dd crm_mul(crm_dd a, crm_dd b) {
  CRM_NO_CONTRACT
  crm_dd p = crm_two_prod(a.h, b.h);
  const double c = a.h * b.l + a.l * b.h;
  return crm_fast_two_sum(p.h, p.l + c);
}

CRM_FN crm_dd crm_add(crm_dd a, crm_dd b) {
  CRM_NO_CONTRACT
  const crm_dd s = crm_two_sum(a.h, b.h);
  return crm_fast_two_sum(s.h, s.l + (a.l + b.l));
}

CRM_FN crm_dd crm_add_d(crm_dd a, double b) {
  CRM_NO_CONTRACT
  const crm_dd s = crm_two_sum(a.h, b);
  return crm_fast_two_sum(s.h, s.l + a.l);
}

CRM_FN crm_dd crm_neg(crm_dd a) { return {-a.h, -a.l}; }

// Ziv's rounding test: h + l approximates the result within err (absolute).
// True when every value within err of h + l rounds to the same double,
// which is then the correctly rounded result.  err must exceed the true
// bound by a few ulps of l (the bounds below carry that margin).
CRM_FN bool crm_round(double h, double l, double err, double* out) {
  CRM_NO_CONTRACT
  const double a = h + (l - err);
  const double b = h + (l + err);
  *out = a;
  return a == b;
}

// ------------------------------------------------------------ cos / sin

// x = k pi/2 + r with |r| <= pi/4 (+ a few ulps); r as a double-double,
// *quad = k mod 4, *rerr a bound of |r - (rh + rl)|.  |x| < 2^20 pi/2.
CRM_FN crm_dd crm_reduce_pio2(double x, int* quad, double* rerr) {
  CRM_NO_CONTRACT
  const double kd = __builtin_rint(x * CRM_TWO_OVER_PI);
  const crm_dd a = crm_two_sum(x, -(kd * CRM_PIO2_C1));  // kd*C1 exact
  const crm_dd b = crm_two_sum(a.h, -(kd * CRM_PIO2_C2));  // kd*C2 exact
  const crm_dd p3 = crm_two_prod(kd, CRM_PIO2_C3);
  const crm_dd c = crm_two_sum(b.h, -p3.h);
  const double lo = (((a.l + b.l) + c.l) - p3.l) - kd * CRM_PIO2_C4;
  // the low-order sum rounds at most four times, each within 2^-53 of
  // partial sums below 2^-52 (|a| + |b| + |c| + |p3|); pi/2 - (C1+..+C4)
  // is below 2^-176
  *rerr = 0x1p-101 * (fabs(a.h) + fabs(b.h) + fabs(c.h) + fabs(p3.h)) + 0x1p-150;
  *quad = static_cast<int>(static_cast<long long>(kd) & 3);
  return crm_two_sum(c.h, lo);
}

// Stage 1: sin and cos of a = ah + al, 0 <= ah <= pi/4 + 2^-40, from the
// table entry j = rint(128 ah) and t = a - j/128 (|t| <= 2^-8).  The
// products S t, C t, S t^2, C t^2 are exact; the remainder is summed in
// double: error below 2^-68 |sin| and 2^-70 |cos| (plus the reduction's).
struct crm_sc1 {
  crm_dd s, c;
};

CRM_FN crm_sc1 crm_sincos_stage1(double ah, double al, int j) {
  CRM_NO_CONTRACT
  const double th = ah - j * 0x1p-7;  // exact (Sterbenz, or j == 0)
  const double tl = al;
  const double Sh = crm_sincos_tab[j][0], Sl = crm_sincos_tab[j][1];
  const double Ch = crm_sincos_tab[j][2], Cl = crm_sincos_tab[j][3];
  const double q = th * th;
  const double ql = __builtin_fma(th, th, -q);
  // cos t - 1 + t^2/2 and sin t - t, |t| <= 2^-8
  const double c4 = (q * q) * (CRM_INV24_H - q * 0x1.6c16c16c16c17p-10);
  const double s3 = (th * q) * (-CRM_INV6_H + q * (CRM_INV120_H - q * 0x1.a01a01a01a01ap-13));
  const double tt = 0.5 * ql + th * tl;  // t^2/2 beyond q/2
  crm_sc1 r;
  {  // sin(j/128 + t) = S + C t - S t^2/2 + S c4 + C s3
    const crm_dd P = crm_two_prod(Ch, th);
    const crm_dd M = crm_two_prod(Sh, q);
    const crm_dd s1 = crm_two_sum(Sh, P.h);
    const crm_dd s2 = crm_two_sum(s1.h, -0.5 * M.h);
    const double low = (s1.l + s2.l) + (P.l - 0.5 * M.l) + (Sl - 0.5 * (Sl * q)) +
                       ((Ch * tl + Cl * th) - Sh * tt) + (Sh * c4 + Ch * s3);
    r.s = crm_fast_two_sum(s2.h, low);
  }
  {  // cos(j/128 + t) = C - S t - C t^2/2 + C c4 - S s3
    const crm_dd P = crm_two_prod(Sh, th);
    const crm_dd M = crm_two_prod(Ch, q);
    const crm_dd s1 = crm_two_sum(Ch, -P.h);
    const crm_dd s2 = crm_two_sum(s1.h, -0.5 * M.h);
    const double low = (s1.l + s2.l) - (P.l + 0.5 * M.l) + (Cl - 0.5 * (Cl * q)) -
                       ((Sh * tl + Sl * th) + Ch * tt) + (Ch * c4 - Sh * s3);
    r.c = crm_fast_two_sum(s2.h, low);
  }
  return r;
}

// Stage 2: the same in double-double throughout, about 2^-98 relative.
CRM_FN crm_sc1 crm_sincos_stage2(double ah, double al, int j) {
  CRM_NO_CONTRACT
  const crm_dd T = crm_two_sum(ah - j * 0x1p-7, al);  // t may be tinier than al
  const crm_dd S = {crm_sincos_tab[j][0], crm_sincos_tab[j][1]};
  const crm_dd C = {crm_sincos_tab[j][2], crm_sincos_tab[j][3]};
  crm_dd q = crm_two_prod(T.h, T.h);
  q = crm_fast_two_sum(q.h, q.l + 2.0 * (T.h * T.l));
  // sin t = t + t^3 (-1/6 + t^2 (1/120 + t^2 (-1/5040 + t^2/362880)))
  const double tail_s = q.h * (-0x1.a01a01a01a01ap-13 + q.h * 0x1.71de3a556c734p-19);
  const crm_dd ps = crm_add(crm_dd{-CRM_INV6_H, -CRM_INV6_L},
                            crm_mul(q, crm_add_d(crm_dd{CRM_INV120_H, CRM_INV120_L}, tail_s)));
  const crm_dd sint = crm_add(T, crm_mul(crm_mul(q, T), ps));
  // cos t = 1 - t^2/2 + t^4 (1/24 + t^2 (-1/720 + t^2/40320))
  const double tail_c = q.h * (-0x1.6c16c16c16c17p-10 + q.h * 0x1.a01a01a01a01ap-16);
  const crm_dd pc = crm_add_d(crm_dd{CRM_INV24_H, CRM_INV24_L}, tail_c);
  crm_dd c1 = crm_two_sum(1.0, -0.5 * q.h);
  c1 = crm_fast_two_sum(c1.h, c1.l - 0.5 * q.l);
  const crm_dd cost = crm_add(c1, crm_mul(crm_mul(q, q), pc));
  crm_sc1 r;
  r.s = crm_add(crm_mul(S, cost), crm_mul(C, sint));
  r.c = crm_add(crm_mul(C, cost), crm_neg(crm_mul(S, sint)));
  return r;
}

// Relative error bounds of the stages (with margin), see above.
#define CRM_SIN_EPS1 0x1p-66
#define CRM_COS_EPS1 0x1p-69
#define CRM_SC_EPS2 0x1p-95

// sin and cos of x, correctly rounded; want_s / want_c select which.
CRM_FN void crm_sincos_impl(double x, double* s, double* c, bool want_s, bool want_c) {
  CRM_NO_CONTRACT
  const double ax = fabs(x);
  if (!(ax < 0x1.921fb54442d18p+20)) {  // NaN, infinity or outside the reduction's range
    if (want_s) *s = sin(x);
    if (want_c) *c = cos(x);
    return;
  }
  if (ax == 0.0) {
    if (want_s) *s = x;
    if (want_c) *c = 1.0;
    return;
  }
  int quad;
  double rerr;
  const crm_dd r = crm_reduce_pio2(x, &quad, &rerr);
  const bool neg = r.h < 0.0;
  const double ah = neg ? -r.h : r.h, al = neg ? -r.l : r.l;
  const int j = static_cast<int>(__builtin_rint(ah * 128.0));
  // sin(x) = +-sin(a) or +-cos(a) by quadrant; cos likewise
  const bool sin_uses_cos = quad & 1;
  const double sin_sign = (((quad >> 1) & 1) ^ (sin_uses_cos ? 0 : (neg ? 1 : 0))) ? -1.0 : 1.0;
  const double cos_sign = ((((quad + 1) >> 1) & 1) ^ (sin_uses_cos ? (neg ? 1 : 0) : 0)) ? -1.0 : 1.0;
  const crm_sc1 v1 = crm_sincos_stage1(ah, al, j);
  double rs = 0, rc = 0;
  bool ok_s = true, ok_c = true;
  // the reduction error moves sin and cos by at most rerr
  if (want_s) {
    const crm_dd f = sin_uses_cos ? v1.c : v1.s;
    ok_s = crm_round(f.h, f.l, (sin_uses_cos ? CRM_COS_EPS1 : CRM_SIN_EPS1) * fabs(f.h) + rerr, &rs);
  }
  if (want_c) {
    const crm_dd f = sin_uses_cos ? v1.s : v1.c;
    ok_c = crm_round(f.h, f.l, (sin_uses_cos ? CRM_SIN_EPS1 : CRM_COS_EPS1) * fabs(f.h) + rerr, &rc);
  }
  if (!(ok_s && ok_c)) {
    CRM_STAGE2();
    const crm_sc1 v2 = crm_sincos_stage2(ah, al, j);
    if (want_s && !ok_s) {
      const crm_dd f = sin_uses_cos ? v2.c : v2.s;
      if (!crm_round(f.h, f.l, CRM_SC_EPS2 * fabs(f.h) + rerr, &rs)) {
        CRM_UNDECIDED();
        rs = f.h + f.l;
      }
    }
    if (want_c && !ok_c) {
      const crm_dd f = sin_uses_cos ? v2.s : v2.c;
      if (!crm_round(f.h, f.l, CRM_SC_EPS2 * fabs(f.h) + rerr, &rc)) {
        CRM_UNDECIDED();
        rc = f.h + f.l;
      }
    }
  }
  if (want_s) *s = sin_sign * rs;
  if (want_c) *c = cos_sign * rc;
}

CRM_FN void crm_sincos(double x, double* s, double* c) { crm_sincos_impl(x, s, c, true, true); }

CRM_FN double crm_sin(double x) {
  double s;
  crm_sincos_impl(x, &s, nullptr, true, false);
  return s;
}

CRM_FN double crm_cos(double x) {
  double c;
  crm_sincos_impl(x, nullptr, &c, false, true);
  return c;
}

// ----------------------------------------------------------- atan / atan2

// atan(u) for u = uh + ul, 0 <= uh <= 1, as a double-double with relative
// error below *eps (stage 1) -- stage 2 is a separate call.
struct crm_at {
  double zh, zl;  // z = (u - c) / (1 + u c), c = j/128
  int j;
};

CRM_FN crm_at crm_atan_reduce(double uh, double ul) {
  CRM_NO_CONTRACT
  crm_at r;
  r.j = static_cast<int>(__builtin_rint(uh * 128.0));
  const double cj = r.j * 0x1p-7;
  const double nh = uh - cj;  // exact (Sterbenz, or j == 0)
  const crm_dd p = crm_two_prod(uh, cj);
  const crm_dd d = crm_two_sum(1.0, p.h);
  const double dl = d.l + (p.l + ul * cj);
  r.zh = nh / d.h;
  const double res = __builtin_fma(-r.zh, d.h, nh);  // nh - zh dh
  r.zl = (res + (ul - r.zh * dl)) / d.h;
  return r;
}

// Stage 1: atan(j/128) + z + z^3 (-1/3 + z^2/5 - z^4/7 + z^6/9); the
// table sum is exact, the cubic term in double: about 2^-69 relative.
CRM_FN crm_dd crm_atan_stage1(const crm_at& a) {
  CRM_NO_CONTRACT
  const double q = a.zh * a.zh;
  const double cub = (a.zh * q) * (-CRM_INV3_H + q * (CRM_INV5_H + q * (-CRM_INV7_H + q * 0x1.c71c71c71c71cp-4)));
  const crm_dd s = crm_two_sum(crm_atan_tab[a.j][0], a.zh);
  const double low = s.l + (crm_atan_tab[a.j][1] + (a.zl - q * a.zl)) + cub;
  return crm_fast_two_sum(s.h, low);
}

// Stage 2: the same in double-double through z^5, about 2^-99 relative.
CRM_FN crm_dd crm_atan_stage2(const crm_at& a) {
  CRM_NO_CONTRACT
  const crm_dd Z = crm_fast_two_sum(a.zh, a.zl);
  crm_dd q = crm_two_prod(Z.h, Z.h);
  q = crm_fast_two_sum(q.h, q.l + 2.0 * (Z.h * Z.l));
  // -1/3 + z^2 (1/5 + z^2 (-1/7 + z^2 (1/9 - z^2/11)))
  const double tail = q.h * (-CRM_INV7_H + q.h * (0x1.c71c71c71c71cp-4 - q.h * 0x1.745d1745d1746p-4));
  const crm_dd p = crm_add(crm_dd{-CRM_INV3_H, -CRM_INV3_L},
                           crm_mul(q, crm_add_d(crm_dd{CRM_INV5_H, CRM_INV5_L}, tail)));
  const crm_dd at = crm_add(Z, crm_mul(crm_mul(q, Z), p));
  return crm_add(crm_dd{crm_atan_tab[a.j][0], crm_atan_tab[a.j][1]}, at);
}

#define CRM_ATAN_EPS1 0x1p-66
#define CRM_ATAN_EPS2 0x1p-95

// base + sgn * atan(u), base one of 0, pi/2, pi; correctly rounded.
CRM_FN double crm_atan_combine(double uh, double ul, int base, double sgn) {
  CRM_NO_CONTRACT
  const crm_at a = crm_atan_reduce(uh, ul);
  crm_dd v = crm_atan_stage1(a);
  crm_dd bd = {0.0, 0.0};
  if (base == 1) bd = crm_dd{CRM_PIO2_H, CRM_PIO2_L};
  if (base == 2) bd = crm_dd{CRM_PI_H, CRM_PI_L};
  // |base + sgn atan(u)| >= |atan(u)| in every case used, so the relative
  // bound of atan(u) carries over; the base constants add 2^-106
  crm_dd f = base ? crm_add(bd, sgn > 0 ? v : crm_neg(v)) : v;
  double out;
  if (crm_round(f.h, f.l, CRM_ATAN_EPS1 * fabs(f.h), &out)) return out;
  CRM_STAGE2();
  v = crm_atan_stage2(a);
  f = base ? crm_add(bd, sgn > 0 ? v : crm_neg(v)) : v;
  if (crm_round(f.h, f.l, CRM_ATAN_EPS2 * fabs(f.h), &out)) return out;
  CRM_UNDECIDED();
  return f.h + f.l;
}

CRM_FN double crm_atan(double x) {
  CRM_NO_CONTRACT
  const double ax = fabs(x);
  if (x != x) return x + x;
  if (ax == 0.0) return x;
  if (ax > 0x1p+54) return copysign(CRM_PIO2_H, x);  // pi/2 - 1/x rounds to pi/2
  double r;
  if (ax <= 1.0) {
    r = crm_atan_combine(ax, 0.0, 0, 1.0);
  } else {
    const double uh = 1.0 / ax;
    const double ul = __builtin_fma(-uh, ax, 1.0) * uh;  // 1/ax - uh to 2^-104
    r = crm_atan_combine(uh, ul, 1, -1.0);
  }
  return copysign(r, x);
}

CRM_FN double crm_atan2(double y, double x) {
  CRM_NO_CONTRACT
  if (x != x || y != y) return x + y;
  const double ax = fabs(x), ay = fabs(y);
  const bool xneg = signbit(x);
  if (ay == 0.0) return copysign(xneg ? CRM_PI_H : 0.0, y);
  if (ax == 0.0) return copysign(CRM_PIO2_H, y);
  if (isinf(ax) || isinf(ay)) {
    double r;
    if (isinf(ax) && isinf(ay)) r = xneg ? 0x1.2d97c7f3321d2p+1 : 0x1.921fb54442d18p-1;  // 3pi/4, pi/4
    else if (isinf(ax)) r = xneg ? CRM_PI_H : 0.0;
    else r = CRM_PIO2_H;
    return copysign(r, y);
  }
  const bool swap = ay > ax;
  const double num = swap ? ax : ay, den = swap ? ay : ax;
  const double uh = num / den;  // (0, 1]
  double r;
  if (uh < 0x1p-60) {
    // atan(u) = u (1 - u^2/3 + ...) is u to 2^-120; a quotient of doubles is
    // never within that of a rounding boundary, so RN(atan u) = RN(u) = uh
    if (!swap && !xneg) r = uh;
    else if (!swap) r = CRM_PI_H;    // pi - u rounds to pi
    else r = CRM_PIO2_H;             // pi/2 -+ u rounds to pi/2
  } else {
    const double ul = __builtin_fma(-uh, den, num) / den;
    if (!swap) r = xneg ? crm_atan_combine(uh, ul, 2, -1.0) : crm_atan_combine(uh, ul, 0, 1.0);
    else r = xneg ? crm_atan_combine(uh, ul, 1, 1.0) : crm_atan_combine(uh, ul, 1, -1.0);
  }
  return copysign(r, y);
}
