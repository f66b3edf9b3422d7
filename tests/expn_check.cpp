// Host check of qmc::exp_nonpos and qmc::exp_clamped against the C
// library's exp.  Line 1: exp_nonpos, the largest distance in ulps over
// uniform and log-uniform samples of [-745.2, 0] and a list of special
// arguments, then the values at -inf and below the underflow threshold.
// Line 2: exp_clamped, the same over [-745.2, 709.78] (both signs, uniform
// and log-uniform), then the values at +inf, just past the overflow
// threshold, at NaN and at -inf.
// Line 3: sincos_turns(u), the largest distance in ulps of sin and cos from
// an extended-precision reference (sinl and cosl of 2 pi (u - q/4) with q
// the nearest quarter turn, mapped to the quadrant by a table), and the
// largest absolute difference from the library's sin and cos of 2 pi u
// (whose own error, from rounding 2 pi u, is about 1e-16 absolute): over
// uniform u in [0, 1], the 24- and 32-bit grids of the Box-Muller uniforms
// and the quarter and eighth turns.
// Line 4: log_unit(x), the largest distance in ulps from the library's log
// over uniform x in (0, 1], the 24- and 32-bit uniform grids, log-uniform x
// down to the subnormals and special points, then log_unit(1).
#define QMC_EXPN_QUAL inline
#include "qmc_expn.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static int64_t ordered(double d) {
  int64_t i;
  std::memcpy(&i, &d, 8);
  return i < 0 ? INT64_MIN - i : i;
}

int main() {
  std::mt19937_64 gen(7);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  int64_t worst = 0;
  double worst_x = 0.0;
  long n = 0;
  double (*f)(double) = qmc::exp_nonpos;
  auto check = [&](double x) {
    int64_t d = ordered(f(x)) - ordered(std::exp(x));
    if (d < 0) d = -d;
    if (d > worst) {
      worst = d;
      worst_x = x;
    }
    ++n;
  };
  for (int i = 0; i < 1000000; ++i) {
    check(-745.2 * u(gen));
    check(-std::pow(10.0, -12.0 + 15.5 * u(gen)));
  }
  const double special[] = {0.0, -0.0, -1e-300, -0.34657359027997264, -0.3465735902799727, -1.0, -708.0,
                            -744.4, -745.1};
  for (double x : special) check(x);
  std::printf("%ld %lld %.17g %.17g %.17g\n", n, (long long)worst, worst_x, qmc::exp_nonpos(-INFINITY),
              qmc::exp_nonpos(-1e300) + qmc::exp_nonpos(-746.0));
  f = qmc::exp_clamped;
  worst = 0;
  worst_x = 0.0;
  n = 0;
  for (int i = 0; i < 1000000; ++i) {
    check(-745.2 + 1454.98 * u(gen));
    const double m = std::pow(10.0, -12.0 + 14.85 * u(gen));
    check(i & 1 ? m : -m);
  }
  const double special2[] = {0.0, -0.0, 1e-300, 0.34657359027997264, 1.0, 708.0, 709.0, 709.78, 709.782712893384,
                             -744.4, -745.1};
  for (double x : special2) check(x);
  std::printf("%ld %lld %.17g %.17g %.17g %.17g %.17g\n", n, (long long)worst, worst_x, qmc::exp_clamped(INFINITY),
              qmc::exp_clamped(709.79), qmc::exp_clamped(NAN), qmc::exp_clamped(-INFINITY));
  // sincos_turns
  worst = 0;
  worst_x = 0.0;
  n = 0;
  double worst_abs = 0.0;
  auto check_sc = [&](double v) {
    double sn, cs;
    qmc::sincos_turns(v, sn, cs);
    const long double q = rintl(4.0L * v), fr = (long double)v - q / 4.0L;
    const long double xr = 6.283185307179586476925286766559L * fr, sr = sinl(xr), cr = cosl(xr);
    const long double st[4] = {sr, cr, -sr, -cr}, ct[4] = {cr, -sr, -cr, sr};
    const int k = (int)q % 4;
    for (int j = 0; j < 2; ++j) {
      const double got = j ? cs : sn, want = (double)(j ? ct[k] : st[k]);
      int64_t d = ordered(got) - ordered(want);
      if (d < 0) d = -d;
      if (d > worst) {
        worst = d;
        worst_x = v;
      }
      const double lib = j ? std::cos(6.283185307179586 * v) : std::sin(6.283185307179586 * v);
      worst_abs = std::fmax(worst_abs, std::fabs(got - lib));
    }
    ++n;
  };
  for (int i = 0; i < 1000000; ++i) {
    check_sc(u(gen));
    check_sc(((double)(gen() >> 40) + 0.5) * 5.9604644775390625e-08);  // 24-bit grid
    check_sc(((double)(gen() >> 32) + 0.5) * 2.3283064365386963e-10);  // 32-bit grid
  }
  for (int j = 0; j <= 8; ++j) check_sc(j / 8.0);
  std::printf("%ld %lld %.17g %.3g\n", n, (long long)worst, worst_x, worst_abs);
  // log_unit
  f = qmc::log_unit;
  worst = 0;
  worst_x = 0.0;
  n = 0;
  auto check_log = [&](double x) {
    int64_t d = ordered(qmc::log_unit(x)) - ordered(std::log(x));
    if (d < 0) d = -d;
    if (d > worst) {
      worst = d;
      worst_x = x;
    }
    ++n;
  };
  for (int i = 0; i < 1000000; ++i) {
    check_log(u(gen));
    check_log(((double)(gen() >> 40) + 0.5) * 5.9604644775390625e-08);  // 24-bit grid
    check_log(((double)(gen() >> 32) + 0.5) * 2.3283064365386963e-10);  // 32-bit grid
    check_log(std::pow(2.0, -1074.0 * u(gen)));                          // log-uniform down to subnormals
  }
  const double special3[] = {1.0, 0.5, 0.70710678118654746, 0.70710678118654757, 1.0 - 0x1p-53, 0x1p-1074,
                             0x1p-1022, 0.1, 2.9802322387695312e-08};
  for (double x : special3) check_log(x);
  std::printf("%ld %lld %.17g %.17g\n", n, (long long)worst, worst_x, qmc::log_unit(1.0));
  return 0;
}
