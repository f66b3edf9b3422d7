// qmc_expn.hpp -- e^x in FP64 for the arguments of the Gaussian and Slater
// primitives, x = -a r^2 or -zeta r: finite and not positive.
//
// Cody-Waite argument reduction (W. J. Cody and W. Waite, Software Manual
// for the Elementary Functions, 1980): n = rint(x / ln 2) and
// r = x - n ln 2 with ln 2 split into a double and the double nearest to
// the remainder, both products taken by fused multiply-add, so |r| stays
// within ln(2)/2 to about 1e-19.  e^r is the Taylor series to degree 13
// in Horner form (the truncation is below 6e-18 relative on the interval)
// and 2^n comes from ldexp, which also takes the far tail (x < -745) to
// zero; arguments below -2000 (and -inf) are raised to -2000 first, so n
// stays a small integer.  The result is within 1 ulp of e^x
// (tests/test_expn.py checks it against the C library on the host over the
// whole range).
//
// The includer defines QMC_EXPN_QUAL (the function qualifiers, e.g.
// "__host__ __device__ __forceinline__") before including this header.
#pragma once

#include <cmath>

namespace qmc {

QMC_EXPN_QUAL double exp_nonpos(double x) {
  x = fmax(x, -2000.0);
  const double n = rint(x * 0x1.71547652b82fep+0);  // x / ln 2
  double r = fma(n, -0x1.62e42fefa39efp-1, x);       // ln 2, rounded
  r = fma(n, -0x1.abc9e3b39803fp-56, r);             // ln 2 minus that
  double p = 0x1.6124613a86d09p-33;                  // 1/13!
  p = fma(p, r, 0x1.1eed8eff8d898p-29);              // 1/12!
  p = fma(p, r, 0x1.ae64567f544e4p-26);              // 1/11!
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);              // 1/10!
  p = fma(p, r, 0x1.71de3a556c734p-19);              // 1/9!
  p = fma(p, r, 0x1.a01a01a01a01ap-16);              // 1/8!
  p = fma(p, r, 0x1.a01a01a01a01ap-13);              // 1/7!
  p = fma(p, r, 0x1.6c16c16c16c17p-10);              // 1/6!
  p = fma(p, r, 0x1.1111111111111p-7);               // 1/5!
  p = fma(p, r, 0x1.5555555555555p-5);               // 1/4!
  p = fma(p, r, 0x1.5555555555555p-3);               // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)n);
}

}  // namespace qmc
