// Compares the device math functions of wx_libm.h, compiled for the host,
// with the host C library on seeded random arguments, and (when built with
// -DWX_QUAD, libquadmath) log with correctly rounded values and sin with the
// count of the library's own results that are not correctly rounded.
//   libm_check N      prints one line: counts of differing results
#include "../../walicxe2d_amd/csrc/wx_libm.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#ifdef WX_QUAD
#include <quadmath.h>
#endif

using namespace wx::libm;

static uint64_t rng = 0x9e3779b97f4a7c15ULL;
static inline uint64_t next() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}
static inline double u01() { return (double)(next() >> 11) * 0x1p-53; }
static inline bool same(double a, double b) { return asu(a) == asu(b) || (a != a && b != b); }

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 1000000;
  long nexp = 0, npow = 0, nlog = 0, nsin = 0, nlog_cr = 0, nsin_cr = 0, ntot_log = 0, nlogpow = 0;
  // wlog_pow(x, y) against wlog(x) and wpow(x, y), bit for bit
  auto logpow = [&](double x, double y) {
    double lg;
    const double pw = wlog_pow(x, y, &lg);
    if (!same(lg, wlog(x)) || !same(pw, wpow(x, y))) {
      ++nlogpow;
      printf("log_pow(%a, %a): %a %a vs %a %a\n", x, y, lg, pw, wlog(x), wpow(x, y));
    }
  };
  // special arguments first
  const double sp[] = {0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 10.0, 1e-310, -1e-310, 0x1p-1074, 1e308, -1e308,
                       709.78, 709.79, -745.13, -745.14, -708.4, 1e-20, -1e-20, __builtin_inf(), -__builtin_inf(),
                       __builtin_nan(""), 3.0, -3.0, 0.79, -0.33333, 1.63, 0.63};
  const int ns = sizeof sp / sizeof sp[0];
  for (int i = 0; i < ns; ++i) {
    if (!same(exp(sp[i]), wexp(sp[i]))) ++nexp;
    for (int j = 0; j < ns; ++j) logpow(sp[i], sp[j]);
    for (int j = 0; j < ns; ++j)
      if (!same(pow(sp[i], sp[j]), wpow(sp[i], sp[j]))) {
        ++npow;
        printf("pow(%a, %a): %a vs %a\n", sp[i], sp[j], pow(sp[i], sp[j]), wpow(sp[i], sp[j]));
      }
  }
  // arguments on which a restatement that leaves one product unfused and the
  // library round apart (found in the cooling jet's full run); read through
  // volatile so that the compiler cannot fold the library call
  static const double rare_pow[][2] = {{0x1.0714b6b6dad7bp+4, 0.79}};
  for (const auto &a : rare_pow) {
    volatile double vx = a[0], vy = a[1];
    if (!same(pow(vx, vy), wpow(vx, vy))) {
      ++npow;
      printf("pow(%a, %a): %a vs %a\n", (double)vx, (double)vy, pow(vx, vy), wpow(vx, vy));
    }
  }
  for (long i = 0; i < n; ++i) {
    // exp over its whole range and near zero
    double x = (i & 1) ? (u01() - 0.5) * 1460.0 : (u01() - 0.5) * 8.0;
    if (!same(exp(x), wexp(x))) ++nexp;
    // pow on the argument families of coolingh (src/coolingh.f90:47,74,132-146)
    // and on random positive bases
    double px, py;
    switch (i % 6) {
      case 0: px = 1.e4 / pow(10.0, u01() * 9.0); py = 0.79; break;           // alpha
      case 1: px = 157890. / pow(10.0, u01() * 9.0); py = -0.33333; break;   // betah
      case 2: px = 10.0; py = (u01() - 0.5) * 800.0; break;                  // 10.**x
      case 3: px = pow(10.0, u01() * 9.0); py = (next() & 1) ? 0.5 : 0.63; break;
      case 4: px = pow(10.0, u01() * 9.0) * 1.e-5; py = 1.63; break;
      default: px = asd(next() & 0x7fffffffffffffffULL); py = (u01() - 0.5) * 8.0; break;
    }
    if (!same(pow(px, py), wpow(px, py))) ++npow;
    logpow(px, py);
    // log on betah's arguments and on random positive doubles
    const double lx = (i % 3 == 0) ? 157890. / pow(10.0, u01() * 9.0) : asd(next() & 0x7fffffffffffffffULL);
    if (lx > 0.0 && lx < __builtin_inf()) {
      ++ntot_log;
      if (!same(log(lx), wlog(lx))) ++nlog;
#ifdef WX_QUAD
      if (!same((double)logq(lx), wlog(lx))) ++nlog_cr;
#endif
    }
    // sin on the jet's argument range 2 pi t / tau, t up to a few periods,
    // and on wider arguments below the library's 105414350 reduction limit
    const double sx = (i & 3) ? u01() * 100.0 : (u01() - 0.5) * 2.0e8;
    if (!same(sin(sx), wsin(sx))) ++nsin;
#ifdef WX_QUAD
    if (!same((double)sinq(sx), wsin(sx))) ++nsin_cr;
#endif
  }
  printf("n %ld exp %ld pow %ld log %ld logtotal %ld sin %ld log_cr %ld sin_cr %ld logpow %ld\n", n, nexp, npow, nlog,
         ntot_log, nsin, nlog_cr, nsin_cr, nlogpow);
  return 0;
}
