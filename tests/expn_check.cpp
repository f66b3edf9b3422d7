// Host check of qmc::exp_nonpos against the C library's exp: prints the
// largest distance in ulps over uniform and log-uniform samples of
// [-745.2, 0] and a list of special arguments, then the values at -inf
// and below the underflow threshold.
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
  auto check = [&](double x) {
    int64_t d = ordered(qmc::exp_nonpos(x)) - ordered(std::exp(x));
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
  return 0;
}
