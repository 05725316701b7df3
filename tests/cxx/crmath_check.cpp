// CPU check of crmath.h (the device's correctly rounded FP64 cos, sin, atan,
// atan2, compiled here for the host) against a 113-bit reference
// (libquadmath, rounded to double) and against the C library.
//
//   crmath_check <fn> <n> <seed>      fn: cos sin atan atan2
//   crmath_check eval <fn> <in> <out> evaluate on the doubles of file <in>
//                                     (pairs y, x for atan2), write <out>
//
// The first form prints one line of counts:
//   fn n <n> wrong <ours != CR> ambiguous <reference too close to call>
//   libc_wrong <libc != CR> differ <ours != libc> stage2 <n> undecided <n>
#include <quadmath.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static long g_stage2 = 0, g_undecided = 0;
#define CRM_STAGE2() (++g_stage2)
#define CRM_UNDECIDED() (++g_undecided)
#include "crmath.h"

static uint64_t g_s = 0;
static uint64_t next_u64() {  // splitmix64
  uint64_t z = (g_s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double uni() { return (next_u64() >> 11) * 0x1p-53; }

// Argument mix of each function: uniform over the kernels' ranges, log-uniform
// magnitudes, and arguments near multiples of pi/2 (cos/sin) or near 1 (atan).
static double arg_trig() {
  const uint64_t k = next_u64() % 4;
  if (k == 0) return (uni() * 2 - 1) * 10.0;
  if (k == 1) return (uni() * 2 - 1) * 4 * M_PI;
  // log-uniform up to 2^20, inside the domain |x| < 2^20 pi/2
  if (k == 2) return ldexp(1.0 + uni(), -60 + static_cast<int>(next_u64() % 80)) * (next_u64() & 1 ? -1 : 1);
  const double m = static_cast<double>(static_cast<int64_t>(next_u64() % 2001) - 1000) * M_PI_2;
  return nextafter(m, 0.0) + (uni() - 0.5) * ldexp(1.0, -static_cast<int>(next_u64() % 40));
}
static double arg_atan() {
  const uint64_t k = next_u64() % 3;
  if (k == 0) return (uni() * 2 - 1) * 4.0;
  if (k == 1) return ldexp(1.0 + uni(), -60 + static_cast<int>(next_u64() % 121)) * (next_u64() & 1 ? -1 : 1);
  return (1.0 + (uni() - 0.5) * 1e-3) * (next_u64() & 1 ? -1 : 1);
}

// cos or sin of x in 113 bits.  libquadmath reduces large arguments with a
// 113-bit pi, which leaves results near zero (x near a multiple of pi/2)
// short of 113 bits; here x - k pi/2 is formed with the 172-bit split of
// pi/2 (C1..C4), every product exact in 113 bits, and the quadrant applied.
static __float128 ref_trig(double x, bool c) {
  const double kd = __builtin_rint(x * CRM_TWO_OVER_PI);
  const __float128 k = kd;
  const __float128 r = (((static_cast<__float128>(x) - k * CRM_PIO2_C1) - k * CRM_PIO2_C2) -
                        k * CRM_PIO2_C3) - k * CRM_PIO2_C4;
  const int quad = static_cast<int>(static_cast<long long>(kd) & 3);
  const __float128 s = sinq(r), co = cosq(r);
  if (c) return quad == 0 ? co : quad == 1 ? -s : quad == 2 ? -co : s;
  return quad == 0 ? s : quad == 1 ? co : quad == 2 ? -s : -co;
}

// Correct rounding of a 113-bit value; false when the value is too close to a
// double's rounding boundary for 113 bits to decide.
static bool cr_of(__float128 q, double* out) {
  const double d = static_cast<double>(q);
  const double up = nextafter(d, INFINITY), dn = nextafter(d, -INFINITY);
  const __float128 mu = (static_cast<__float128>(d) + up) / 2, md = (static_cast<__float128>(d) + dn) / 2;
  const __float128 tol = fabsq(q) * static_cast<__float128>(0x1p-108);
  *out = d;
  return fabsq(q - mu) > tol && fabsq(q - md) > tol;
}

// Largest ratio of a stage's actual error to the bound its rounding test
// uses, over n arguments: both must stay below 1.
static void bounds(const char* fn, long n) {
  double r1 = 0, r2 = 0, rr = 0;
  for (long i = 0; i < n; i++) {
    crm_dd v1, v2;
    double e1, e2;
    __float128 q;
    if (!strcmp(fn, "cos") || !strcmp(fn, "sin")) {
      const double x = arg_trig();
      if (x == 0) continue;
      int quad;
      double rerr;
      const crm_dd r = crm_reduce_pio2(x, &quad, &rerr);
      const bool neg = r.h < 0.0;
      const double ah = neg ? -r.h : r.h, al = neg ? -r.l : r.l;
      const int j = static_cast<int>(__builtin_rint(ah * 128.0));
      const crm_sc1 s1 = crm_sincos_stage1(ah, al, j), s2 = crm_sincos_stage2(ah, al, j);
      // the stage error against the 113-bit function at the stages' own
      // argument ah + al; the reduction's error is checked separately
      const __float128 aq = static_cast<__float128>(ah) + al;
      // x - k pi/2 with pi/2 = C1 + .. + C4 (172 bits, checked against
      // acosq(0) in main): every product is exact in 113 bits
      const __float128 k = __builtin_rint(x * CRM_TWO_OVER_PI);
      const __float128 exact = fabsq((((static_cast<__float128>(x) - k * CRM_PIO2_C1) - k * CRM_PIO2_C2) -
                                      k * CRM_PIO2_C3) - k * CRM_PIO2_C4);
      const double dr = static_cast<double>(fabsq(exact - aq));
      if (dr / rerr > rr) rr = dr / rerr;
      const bool c = fn[0] == 'c';
      q = c ? cosq(aq) : sinq(aq);
      v1 = c ? s1.c : s1.s;
      v2 = c ? s2.c : s2.s;
      e1 = (c ? CRM_COS_EPS1 : CRM_SIN_EPS1) * fabs(v1.h);
      e2 = CRM_SC_EPS2 * fabs(v2.h);
    } else {
      double u = fabs(arg_atan());
      double uh = u, ul = 0;
      if (u > 1) {
        uh = 1.0 / u;
        ul = __builtin_fma(-uh, u, 1.0) * uh;
      }
      if (uh == 0) continue;
      const crm_at a = crm_atan_reduce(uh, ul);
      v1 = crm_atan_stage1(a);
      v2 = crm_atan_stage2(a);
      q = atanq(u > 1 ? 1 / static_cast<__float128>(u) : static_cast<__float128>(u));
      e1 = CRM_ATAN_EPS1 * fabs(v1.h);
      e2 = CRM_ATAN_EPS2 * fabs(v2.h);
    }
    const double d1 = static_cast<double>(fabsq(static_cast<__float128>(v1.h) + v1.l - q));
    const double d2 = static_cast<double>(fabsq(static_cast<__float128>(v2.h) + v2.l - q));
    if (d1 / e1 > r1) r1 = d1 / e1;
    if (d2 / e2 > r2) r2 = d2 / e2;
  }
  printf("%s bounds n %ld stage1 max error/bound %.3g stage2 max error/bound %.3g reduction %.3g\n",
         fn, n, r1, r2, rr);
}

int main(int argc, char** argv) {
  if (argc == 5 && !strcmp(argv[1], "bounds")) {
    const __float128 pio2 = static_cast<__float128>(CRM_PIO2_C1) + CRM_PIO2_C2 + CRM_PIO2_C3 + CRM_PIO2_C4;
    if (fabsq(pio2 - acosq(static_cast<__float128>(0))) > ldexpq(1, -111)) {
      printf("pi/2 constants disagree with acosq(0)\n");
      return 1;
    }
    g_s = strtoull(argv[4], nullptr, 10);
    bounds(argv[2], atol(argv[3]));
    return 0;
  }
  if (argc == 5 && !strcmp(argv[1], "eval")) {
    const char* fn = argv[2];
    FILE* f = fopen(argv[3], "rb");
    if (!f) return 2;
    std::vector<double> in;
    double v;
    while (fread(&v, 8, 1, f) == 1) in.push_back(v);
    fclose(f);
    std::vector<double> out;
    if (!strcmp(fn, "atan2")) {
      for (size_t i = 0; i + 1 < in.size(); i += 2) out.push_back(crm_atan2(in[i], in[i + 1]));
    } else {
      for (double x : in)
        out.push_back(!strcmp(fn, "cos") ? crm_cos(x) : !strcmp(fn, "sin") ? crm_sin(x) : crm_atan(x));
    }
    f = fopen(argv[4], "wb");
    if (!f || fwrite(out.data(), 8, out.size(), f) != out.size()) return 2;
    fclose(f);
    return 0;
  }
  if (argc != 4) return 2;
  const char* fn = argv[1];
  const long n = atol(argv[2]);
  g_s = strtoull(argv[3], nullptr, 10);
  long wrong = 0, amb = 0, libc_wrong = 0, differ = 0;
  for (long i = 0; i < n; i++) {
    double ours, libc, cr, arg = 0;
    __float128 q;
    if (!strcmp(fn, "cos") || !strcmp(fn, "sin")) {
      const double x = arg_trig();
      arg = x;
      const bool c = fn[0] == 'c';
      ours = c ? crm_cos(x) : crm_sin(x);
      libc = c ? cos(x) : sin(x);
      q = ref_trig(x, c);
    } else if (!strcmp(fn, "atan")) {
      const double x = arg_atan();
      ours = crm_atan(x);
      libc = atan(x);
      q = atanq(static_cast<__float128>(x));
    } else {
      double y = (uni() * 2 - 1) * 3, x = (uni() * 2 - 1) * 3;
      if (next_u64() % 4 == 0) y = ldexp(y, -static_cast<int>(next_u64() % 60));
      if (next_u64() % 4 == 0) x = ldexp(x, -static_cast<int>(next_u64() % 60));
      ours = crm_atan2(y, x);
      libc = atan2(y, x);
      q = atan2q(static_cast<__float128>(y), static_cast<__float128>(x));
    }
    if (!cr_of(q, &cr)) {
      amb++;
      continue;
    }
    if (ours != cr) {
      if (wrong < 5) fprintf(stderr, "%s(%a): ours %a cr %a libc %a\n", fn, arg, ours, cr, libc);
      wrong++;
    }
    libc_wrong += libc != cr;
    differ += ours != libc;
  }
  printf("%s n %ld wrong %ld ambiguous %ld libc_wrong %ld differ %ld stage2 %ld undecided %ld\n", fn, n,
         wrong, amb, libc_wrong, differ, g_stage2, g_undecided);
  return wrong ? 1 : 0;
}
