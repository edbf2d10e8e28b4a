// wx_libm.h -- FP64 exp, pow, log and sin for the device, written so that
// their results are the host C library's (glibc 2.35, x86-64, FMA code path)
// bit for bit wherever that is known to be possible:
//
//   wexp, wpow  the method of Szabolcs Nagy (ARM optimized-routines, glibc's
//               exp and pow since 2.28): 128-entry tables (generated from
//               their defining formulas by gen_libm_tables.py), the same
//               polynomials, the same order of operations and the same fused
//               multiply-adds as the library's FMA build.  Equal to the host
//               library on every argument tested.
//   wlog        pow's extended-precision logarithm (about 15 bits beyond
//               double) rounded to double: correctly rounded except on about
//               3 arguments in 10^5.  The host library's log is not correctly
//               rounded on about 2 arguments in 10^4; there the two differ by
//               one ulp.
//   wsin        the structure of the C library's sin (IBM Accurate
//               Mathematical Library, glibc >= 2.28): table of sin/cos at
//               i/128 (generated), the same series, reduction constants and
//               fused multiply-adds; equal to the host library on every
//               argument tested below |x| = 105414350 (beyond: ocml's sin).
//
// The reference computes its rates and losses with the Fortran intrinsics
// exp, log, sin and the ** operator (src/coolingh.f90:41-165, JET/jet.f90:55,
// src/coolingdmc.f90:36), which resolve to these library functions; the
// neutral-fraction update of coolingh amplifies a one-ulp difference of a
// rate about 2000-fold.
//
// The header compiles as HIP device+host code and as plain C++ (the CPU test
// that compares it with the host library on random arguments).  It must be
// compiled without floating-point contraction (-ffp-contract=off): every
// fused multiply-add below is written out.
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WXM_FN __host__ __device__ __forceinline__
#define WXM_TABLE(type, name, n) static constexpr type name[n]
#else
#include <cmath>
#define WXM_FN static inline
#define WXM_TABLE(type, name, n) static constexpr type name[n]
#endif

namespace wx {
namespace libm {

#include "wx_libm_tables.inc"

WXM_FN uint64_t asu(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}
WXM_FN double asd(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
}
WXM_FN double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
WXM_FN uint32_t top12(double x) { return (uint32_t)(asu(x) >> 52); }

// ---------------------------------------------------------------------------
// exp and pow share the exp kernel: exp(x) = 2^(k/128) exp(r),
// x = k ln2/128 + r, |r| <= ln2/256, 2^(k/128) = H (1 + T) from kExpTab.
constexpr double kInvLn2N = 0x1.71547652b82fep0 * 128;
constexpr double kNegLn2hiN = -0x1.62e42fefa0000p-8;
constexpr double kNegLn2loN = -0x1.cf79abc9e3b3ap-47;
constexpr double kShift = 0x1.8p52;
constexpr double kC2 = 0x1.ffffffffffdbdp-2, kC3 = 0x1.555555555543cp-3, kC4 = 0x1.55555cf172b91p-5,
                 kC5 = 0x1.1111167a4d017p-7;
constexpr uint32_t kSignBias = 0x800 << 7;

// Results whose scale 2^(k/128) over- or underflows a double's exponent.
WXM_FN double exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  if ((ki & 0x80000000) == 0) {
    // k > 0: the exponent of scale overflowed by at most 460
    sbits -= 1009ull << 52;
    const double scale = asd(sbits);
    return 0x1p1009 * fmad(scale, tmp, scale);
  }
  // k < 0: care in the subnormal range (the product is not fused here)
  sbits += 1022ull << 52;
  const double scale = asd(sbits);
  const double st = scale * tmp;
  double y = scale + st;
  if (y < 1.0 && y > -1.0) {
    const double one = y < 0.0 ? -1.0 : 1.0;
    double lo = scale - y + st;
    const double hi = one + y;
    lo = one - hi + y + lo;
    y = (hi + lo) - one;
    if (y == 0.0) y = asd(sbits & 0x8000000000000000ull);
  }
  return 0x1p-1022 * y;
}

// exp(x + xtail) with the sign bias of pow's negative bases.
WXM_FN double exp_kernel(double x, double xtail, uint32_t sign_bias, bool with_tail) {
  uint32_t abstop = top12(x) & 0x7ff;
  if (abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54)) {
    if (abstop - top12(0x1p-54) >= 0x80000000u) {
      // tiny x: exp(x) rounds to 1 (+x gives the rounding of exp itself)
      if (with_tail) return sign_bias ? -1.0 : 1.0;
      return 1.0 + x;
    }
    if (abstop >= top12(1024.0)) {
      if (!with_tail) {
        if (asu(x) == 0xfff0000000000000ull) return 0.0;
        if (abstop >= top12(__builtin_inf())) return 1.0 + x;
      }
      if (asu(x) >> 63) return sign_bias ? -0.0 : 0.0;
      return sign_bias ? -__builtin_inf() : __builtin_inf();
    }
    abstop = 0;  // large |x|: scale handled by exp_special
  }
  double kd = fmad(kInvLn2N, x, kShift);
  const uint64_t ki = asu(kd);
  kd -= kShift;
  double r = fmad(kd, kNegLn2loN, fmad(kd, kNegLn2hiN, x));
  if (with_tail) r += xtail;
  const uint64_t idx = 2 * (ki % 128);
  const uint64_t top = (ki + sign_bias) << 45;
  const double tail = asd(kExpTab[idx]);
  const uint64_t sbits = kExpTab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fmad(r2 * r2, fmad(r, kC5, kC4), fmad(r2, fmad(r, kC3, kC2), tail + r));
  if (abstop == 0) return exp_special(tmp, sbits, ki);
  const double scale = asd(sbits);
  return fmad(scale, tmp, scale);
}

// The main path of exp_kernel alone, with no branch: its result for every
// argument that takes that path (2^-54 <= |x| < 512); `special` is set for
// the others, whose result exp_kernel gives.  (Straight-line code lets the
// compiler interleave independent calls; the table index is in range for
// any argument.)
WXM_FN double exp_main(double x, double xtail, uint32_t sign_bias, bool with_tail, bool &special) {
  const uint32_t abstop = top12(x) & 0x7ff;
  special = special || abstop - top12(0x1p-54) >= top12(512.0) - top12(0x1p-54);
  double kd = fmad(kInvLn2N, x, kShift);
  const uint64_t ki = asu(kd);
  kd -= kShift;
  double r = fmad(kd, kNegLn2loN, fmad(kd, kNegLn2hiN, x));
  if (with_tail) r += xtail;
  const uint64_t idx = 2 * (ki % 128);
  const uint64_t top = (ki + sign_bias) << 45;
  const double tail = asd(kExpTab[idx]);
  const uint64_t sbits = kExpTab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fmad(r2 * r2, fmad(r, kC5, kC4), fmad(r2, fmad(r, kC3, kC2), tail + r));
  const double scale = asd(sbits);
  return fmad(scale, tmp, scale);
}

WXM_FN double wexp(double x) {
  bool special = false;
  double y = exp_main(x, 0.0, 0, false, special);
  if (special) y = exp_kernel(x, 0.0, 0, false);
  return y;
}

// pow's logarithm: log(x) = hi + tail with about 15 bits beyond double.
// x = 2^k z, z in [OFF, 2 OFF); z/c - 1 is exact since 1/c has few bits.
constexpr uint64_t kPowOff = 0x3fe6955500000000ull;
constexpr double kLn2hi = 0x1.62e42fefa3800p-1, kLn2lo = 0x1.ef35793c76730p-45;
constexpr double kA0 = -0x1p-1, kA1 = 0x1.555555555556p-2 * -2, kA2 = -0x1.0000000000006p-2 * -2,
                 kA3 = 0x1.999999959554ep-3 * 4, kA4 = -0x1.555555529a47ap-3 * 4,
                 kA5 = 0x1.2495b9b4845e9p-3 * -8, kA6 = -0x1.0002b8b263fc3p-3 * -8;

WXM_FN double pow_log(uint64_t ix, double *tail) {
  const uint64_t tmp = ix - kPowOff;
  const int i = (int)((tmp >> 45) % 128);
  const int k = (int)((int64_t)tmp >> 52);
  const uint64_t iz = ix - (tmp & 0xfffull << 52);
  const double z = asd(iz);
  const double kd = (double)k;
  const double invc = asd(kPowLogTab[3 * i]);
  const double logc = asd(kPowLogTab[3 * i + 1]);
  const double logctail = asd(kPowLogTab[3 * i + 2]);
  const double r = fmad(z, invc, -1.0);
  const double t1 = fmad(kd, kLn2hi, logc);
  const double t2 = t1 + r;
  const double lo1 = fmad(kd, kLn2lo, logctail);
  const double lo2 = t1 - t2 + r;
  const double ar = kA0 * r;
  const double ar2 = r * ar;
  const double ar3 = r * ar2;
  const double hi = t2 + ar2;
  const double lo3 = fmad(ar, r, -ar2);
  const double lo4 = t2 - hi + ar2;
  // the library's p = ar3 * (...) has one use, the last addition of lo, and
  // is fused into it (one argument in 10^7 or fewer rounds differently)
  const double lo = fmad(ar3, fmad(ar2, fmad(ar2, fmad(r, kA6, kA5), fmad(r, kA4, kA3)), fmad(r, kA2, kA1)),
                         lo1 + lo2 + lo3 + lo4);
  const double y = hi + lo;
  *tail = hi - y + lo;
  return y;
}

// 0: not an integer, 1: odd integer, 2: even integer
WXM_FN int pow_checkint(uint64_t iy) {
  const int e = (int)(iy >> 52 & 0x7ff);
  if (e < 0x3ff) return 0;
  if (e > 0x3ff + 52) return 2;
  if (iy & ((1ull << (0x3ff + 52 - e)) - 1)) return 0;
  if (iy & (1ull << (0x3ff + 52 - e))) return 1;
  return 2;
}
WXM_FN bool pow_zeroinfnan(uint64_t i) { return 2 * i - 1 >= 2 * 0x7ff0000000000000ull - 1; }

WXM_FN double wpow_full(double x, double y) {
  uint32_t sign_bias = 0;
  uint64_t ix = asu(x);
  const uint64_t iy = asu(y);
  uint32_t topx = top12(x);
  const uint32_t topy = top12(y);
  if (topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
    // x < 2^-1022, inf or nan; or |y| < 2^-65, |y| >= 2^63 or nan
    if (pow_zeroinfnan(iy)) {
      if (2 * iy == 0) return 1.0;
      if (ix == asu(1.0)) return 1.0;
      if (2 * ix > 2 * 0x7ff0000000000000ull || 2 * iy > 2 * 0x7ff0000000000000ull) return x + y;
      if (2 * ix == 2 * asu(1.0)) return 1.0;
      if ((2 * ix < 2 * asu(1.0)) == !(iy >> 63)) return 0.0;
      return y * y;
    }
    if (pow_zeroinfnan(ix)) {
      double x2 = x * x;
      if ((ix >> 63) && pow_checkint(iy) == 1) x2 = -x2;
      return (iy >> 63) ? 1.0 / x2 : x2;
    }
    if (ix >> 63) {
      const int yint = pow_checkint(iy);
      if (yint == 0) return __builtin_nan("");
      if (yint == 1) sign_bias = kSignBias;
      ix &= 0x7fffffffffffffffull;
      topx &= 0x7ff;
    }
    if ((topy & 0x7ff) - 0x3be >= 0x43e - 0x3be) {
      if (ix == asu(1.0)) return 1.0;
      if ((topy & 0x7ff) < 0x3be) return ix > asu(1.0) ? 1.0 + y : 1.0 - y;
      return ((ix > asu(1.0)) == (topy < 0x800)) ? __builtin_inf() : 0.0;
    }
    if (topx == 0) {
      // subnormal x: normalise so that the exponent becomes negative
      ix = asu(x * 0x1p52);
      ix &= 0x7fffffffffffffffull;
      ix -= 52ull << 52;
    }
  }
  double lo;
  const double hi = pow_log(ix, &lo);
  const double ehi = y * hi;
  const double elo = fmad(y, lo, fmad(y, hi, -ehi));
  return exp_kernel(ehi, elo, sign_bias, true);
}

// pow's main path (x a positive normal number, 2^-65 <= |y| < 2^63, and exp's
// main path), with no branch; `special` is set for the other arguments.
WXM_FN double pow_main(double x, double y, bool &special) {
  const uint32_t topx = top12(x), topy = top12(y);
  special = special || topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be;
  double lo;
  const double hi = pow_log(asu(x), &lo);
  const double ehi = y * hi;
  const double elo = fmad(y, lo, fmad(y, hi, -ehi));
  return exp_main(ehi, elo, 0, true, special);
}

WXM_FN double wpow(double x, double y) {
  bool special = false;
  double r = pow_main(x, y, special);
  if (special) r = wpow_full(x, y);
  return r;
}

// log: pow's extended-precision logarithm rounded to double.
WXM_FN double wlog(double x) {
  uint64_t ix = asu(x);
  const uint32_t top = (uint32_t)(ix >> 48);
  if (top - 0x0010 >= 0x7ff0 - 0x0010) {
    // x < 2^-1022 or inf or nan
    if (ix * 2 == 0) return -__builtin_inf();
    if (ix == 0x7ff0000000000000ull) return x;
    if ((top & 0x8000) || (top & 0x7ff0) == 0x7ff0) return __builtin_nan("");
    ix = asu(x * 0x1p52);
    ix -= 52ull << 52;
  }
  if (ix == asu(1.0)) return 0.0;
  double tail;
  return pow_log(ix, &tail);
}

// log(x) and pow(x, y) of one argument x, with one evaluation of pow's
// logarithm for both (what pow computes first is what log rounds): *lg is
// wlog(x) and the result wpow(x, y), bit for bit; the special cases of either
// (x zero, subnormal, negative, infinite or NaN, or y outside pow's main
// range) take the two functions themselves.
WXM_FN double wlog_pow(double x, double y, double *lg) {
  const uint64_t ix = asu(x);
  const uint32_t topx = top12(x), topy = top12(y);
  bool special = topx - 0x001 >= 0x7ff - 0x001 || (topy & 0x7ff) - 0x3be >= 0x43e - 0x3be;
  double lo;
  const double hi = pow_log(ix, &lo);
  const double ehi = y * hi;
  const double elo = fmad(y, lo, fmad(y, hi, -ehi));
  double r = exp_main(ehi, elo, 0, true, special);
  *lg = ix == asu(1.0) ? 0.0 : hi;
  if (special) {
    *lg = wlog(x);
    r = wpow_full(x, y);
  }
  return r;
}

// ---------------------------------------------------------------------------
// sin: the structure of the C library's sin (glibc >= 2.28, the IBM Accurate
// Mathematical Library's s_sin.c after the removal of its slow paths): a
// table of sin and cos at i/128 (kSinCosTab) combined with short series of
// the remainder, the argument reduced by a four-part pi/2 below |x| <
// 105414350, a Taylor form near zero; the fused multiply-adds are those of
// the library's FMA build.
constexpr double kSinBig = 0x1.8p45;                  // 52776558133248: x + big rounds x to 1/128
constexpr double kHp0 = 0x1.921fb54442d18p0, kHp1 = 0x1.1a62633145c07p-54;   // pi/2 = hp0 + hp1
constexpr double kMp1 = 0x1.921fb58000000p0, kMp2 = -0x1.dde973c000000p-27;  // pi/2 in four parts
constexpr double kPp3 = -0x1.cb3b398000000p-55, kPp4 = -0x1.d747f23e32ed7p-83;
constexpr double kHpInv = 0x1.45f306dc9c883p-1, kToInt = 0x1.8p52;
constexpr double kSn3 = -1.66666666666664880952546298448555E-01, kSn5 = 8.33333214285722277379541354343671E-03,
                 kCs2 = 4.99999999999999999999950396842453E-01, kCs4 = -4.16666666666664434524222570944589E-02,
                 kCs6 = 1.38888874007937613028114285595617E-03;
constexpr double kS1 = -0x1.5555555555555p-3, kS2 = 0x1.1111111110ECEp-7, kS3 = -0.00019841269834414642,
                 kS4 = 2.755729806860771e-06, kS5 = -2.5022014848318398e-08;

// (a + da) - (a + da)^3/3! + ... near zero (|a| < 0.126)
WXM_FN double sin_taylor(double xx, double a, double da) {
  const double p = fmad(fmad(fmad(fmad(kS5, xx, kS4), xx, kS3), xx, kS2), xx, kS1);
  const double t = fmad(fmad(p, a, -(0.5 * da)), xx, da);
  return a + t;
}

// cos(x + dx) from the table entry nearest x
WXM_FN double sin_do_cos(double x, double dx) {
  if (x < 0) dx = -dx;
  const double ux = kSinBig + __builtin_fabs(x);
  x = __builtin_fabs(x) - (ux - kSinBig) + dx;
  const double xx = x * x;
  const double s = fmad(x * xx, fmad(xx, kSn5, kSn3), x);
  const double c = xx * fmad(xx, fmad(xx, kCs6, kCs4), kCs2);
  const int k = (int)((uint32_t)asu(ux) << 2);
  const double sn = asd(kSinCosTab[k]), ssn = asd(kSinCosTab[k + 1]);
  const double cs = asd(kSinCosTab[k + 2]), ccs = asd(kSinCosTab[k + 3]);
  const double cor = fmad(-sn, s, fmad(-cs, c, fmad(-s, ssn, ccs)));
  return cs + cor;
}

// sin(x + dx) from the table entry nearest x (Taylor form near zero)
WXM_FN double sin_do_sin(double x, double dx) {
  const double xold = x;
  if (__builtin_fabs(x) < 0.126) return sin_taylor(x * x, x, dx);
  if (x <= 0) dx = -dx;
  const double ux = kSinBig + __builtin_fabs(x);
  x = __builtin_fabs(x) - (ux - kSinBig);
  const double xx = x * x;
  const double s = x + fmad(x * xx, fmad(xx, kSn5, kSn3), dx);
  const double c = fmad(x, dx, xx * fmad(xx, fmad(xx, kCs6, kCs4), kCs2));
  const int k = (int)((uint32_t)asu(ux) << 2);
  const double sn = asd(kSinCosTab[k]), ssn = asd(kSinCosTab[k + 1]);
  const double cs = asd(kSinCosTab[k + 2]), ccs = asd(kSinCosTab[k + 3]);
  const double cor = fmad(cs, s, fmad(-sn, c, fmad(s, ccs, ssn)));
  return __builtin_copysign(sn + cor, xold);
}

WXM_FN double wsin(double x) {
  const uint32_t k = (uint32_t)(asu(x) >> 32) & 0x7fffffffu;
  if (k < 0x3e500000u) return x;                        // |x| < 2^-26
  if (k < 0x3feb6000u) return sin_do_sin(x, 0.0);      // |x| < 0.855469
  if (k < 0x400368fdu) {                                // |x| < 2.426265
    const double t = kHp0 - __builtin_fabs(x);
    return __builtin_copysign(sin_do_cos(t, kHp1), x);
  }
  if (k < 0x419921fbu) {                                // |x| < 105414350
    const double t = fmad(x, kHpInv, kToInt);
    const double xn = t - kToInt;
    const int n = (int)(asu(t) & 3);
    const double y = fmad(-xn, kMp2, fmad(-xn, kMp1, x));
    const double t2 = fmad(-xn, kPp3, y);
    double db = fmad(-xn, kPp3, y - t2);
    const double b = fmad(-xn, kPp4, t2);
    db = db + fmad(-xn, kPp4, t2 - b);
    const double r = (n & 1) ? sin_do_cos(b, db) : sin_do_sin(b, db);
    return (n & 2) ? -r : r;
  }
  // beyond (not reached by the reference's jet speed): the device's own sin
#if defined(__HIP_DEVICE_COMPILE__)
  return ::sin(x);
#else
  return std::sin(x);
#endif
}

}  // namespace libm
}  // namespace wx
