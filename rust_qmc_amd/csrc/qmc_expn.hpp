// qmc_expn.hpp -- e^x in FP64 without the library's range checks:
// exp_nonpos for the arguments of the Gaussian and Slater primitives,
// x = -a r^2 or -zeta r (finite and not positive), and exp_clamped for the
// Metropolis acceptance e^(2 du), whose argument has either sign.
//
// Cody-Waite argument reduction (W. J. Cody and W. Waite, Software Manual
// for the Elementary Functions, 1980): n = rint(x / ln 2) and
// r = x - n ln 2 with ln 2 split into a double and the double nearest to
// the remainder, both products taken by fused multiply-add, so |r| stays
// within ln(2)/2 to about 1e-19.  e^r is the Taylor series to degree 13
// in Horner form (the truncation is below 6e-18 relative on the interval)
// and 2^n comes from ldexp, which also takes the far tail (x < -745) to
// zero and x > 709.78 to infinity; arguments are held within [-2000, 2000]
// first (exp_nonpos raises those below -2000 and -inf to -2000;
// exp_clamped also lowers those above 2000 and +inf to 2000, and takes NaN
// to e^-2000 = 0), so n stays a small integer.  The result is within 1 ulp
// of e^x (tests/test_expn.py checks both against the C library on the host
// over the whole range).
//
// sincos_turns gives sin(2 pi u) and cos(2 pi u) of a Box-Muller angle u in
// [0, 1] (turns): the reduction to the nearest quarter turn q/4 is exact
// (f = u - q/4 in [-1/8, 1/8] by a fused multiply-add of a number with at
// most 53 - 2 significant bits), x = 2 pi f in [-pi/4, pi/4] is one rounded
// product, and sin x = x + x^3 P(x^2), cos x = 1 - x^2/2 + x^4 Q(x^2) are the
// Taylor series through x^19 and x^18 (truncation below 1e-19 relative),
// mapped to the quadrant by the symmetries of sin and cos.  Unlike
// sincos(2 pi u) of the library, no error of the product 2 pi u reaches the
// result (tests/test_expn.py checks it against an extended-precision
// reference on the host).
//
// log_unit gives ln x for x in (0, 1] (the Box-Muller radii's ln u): x =
// m 2^e with m in [sqrt(1/2), sqrt(2)) (frexp), f = m - 1 exact, then
// ln(1 + f) = 2 atanh(s) with s = f / (2 + f), evaluated as
// f - (f^2/2 - s (f^2/2 + R(s^2))) where R(z) = sum_k 2 z^k / (2k + 1) is
// the atanh series to k = 10 (truncation below 1e-19 relative), and
// e ln 2 added with ln 2 split into a 32-bit head (e ln2_hi exact) and a
// tail (the decomposition of fdlibm's log, Sun Microsystems 1993, with the
// series in place of its minimax polynomial).  Within 1 ulp of the library's log
// on the host (tests/test_expn.py).
//
// The includer defines QMC_EXPN_QUAL (the function qualifiers, e.g.
// "__host__ __device__ __forceinline__") before including this header.
#pragma once

#include <cmath>

namespace qmc {

// e^x for x in [-2000, 2000]
QMC_EXPN_QUAL double exp_core(double x) {
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

QMC_EXPN_QUAL double exp_nonpos(double x) { return exp_core(fmax(x, -2000.0)); }

QMC_EXPN_QUAL double exp_clamped(double x) { return exp_core(fmin(fmax(x, -2000.0), 2000.0)); }

QMC_EXPN_QUAL double log_unit(double x) {
  int e;
  double m = frexp(x, &e);  // x = m 2^e, m in [1/2, 1)
  if (m < 0x1.6a09e667f3bcdp-1) {  // sqrt(1/2)
    m *= 2.0;
    --e;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f), z = s * s;
  double r = 0x1.8618618618618p-4;      // 2/21
  r = fma(r, z, 0x1.af286bca1af28p-4);  // 2/19
  r = fma(r, z, 0x1.e1e1e1e1e1e1ep-4);  // 2/17
  r = fma(r, z, 0x1.1111111111111p-3);  // 2/15
  r = fma(r, z, 0x1.3b13b13b13b14p-3);  // 2/13
  r = fma(r, z, 0x1.745d1745d1746p-3);  // 2/11
  r = fma(r, z, 0x1.c71c71c71c71cp-3);  // 2/9
  r = fma(r, z, 0x1.2492492492492p-2);  // 2/7
  r = fma(r, z, 0x1.999999999999ap-2);  // 2/5
  r = fma(r, z, 0x1.5555555555555p-1);  // 2/3
  r *= z;
  const double hf = 0.5 * f * f, de = (double)e;
  // e ln2_hi + (f - (hf - (s (hf + r) + e ln2_lo)))
  return fma(de, 0x1.62e42ff000000p-1, f - (hf - fma(s, hf + r, de * -0x1.718432a1b0e26p-35)));
}

QMC_EXPN_QUAL void sincos_turns(double u, double& sn, double& cs) {
  const double q = rint(4.0 * u);
  const double f = fma(q, -0.25, u);
  // x + xl = 2 pi f to about 2^-106 relative: 2 pi as two doubles, the
  // product's rounding error recovered by a fused multiply-add
  const double x = f * 0x1.921fb54442d18p+2;
  const double xl = fma(f, 0x1.1a62633145c07p-52, fma(f, 0x1.921fb54442d18p+2, -x));
  const double x2 = x * x;
  double p = -0x1.2f49b46814157p-57;  // -1/19!
  p = fma(p, x2, 0x1.952c77030ad4ap-49);   // 1/17!
  p = fma(p, x2, -0x1.ae7f3e733b81fp-41);  // -1/15!
  p = fma(p, x2, 0x1.6124613a86d09p-33);   // 1/13!
  p = fma(p, x2, -0x1.ae64567f544e4p-26);  // -1/11!
  p = fma(p, x2, 0x1.71de3a556c734p-19);   // 1/9!
  p = fma(p, x2, -0x1.a01a01a01a01ap-13);  // -1/7!
  p = fma(p, x2, 0x1.1111111111111p-7);    // 1/5!
  p = fma(p, x2, -0x1.5555555555555p-3);   // -1/3!
  const double sx0 = fma(x * x2, p, x);
  double c = 0x1.6827863b97d97p-53;       // 1/18!
  c = fma(c, x2, -0x1.ae7f3e733b81fp-45);  // -1/16!
  c = fma(c, x2, 0x1.93974a8c07c9dp-37);   // 1/14!
  c = fma(c, x2, -0x1.1eed8eff8d898p-29);  // -1/12!
  c = fma(c, x2, 0x1.27e4fb7789f5cp-22);   // 1/10!
  c = fma(c, x2, -0x1.a01a01a01a01ap-16);  // -1/8!
  c = fma(c, x2, 0x1.6c16c16c16c17p-10);   // 1/6!
  c = fma(c, x2, -0x1.5555555555555p-5);   // -1/4!
  // 1 - x^2/2 as t + e: h = x^2/2 and t = 1 - h rounded, e = (1 - t) - h
  // exact (1 >= h), x2l the rounding error of x^2
  const double h = 0.5 * x2, t = 1.0 - h, e = (1.0 - t) - h;
  const double x2l = fma(x, x, -x2);
  const double cx0 = t + fma(x2 * x2, -c, fma(x2l, -0.5, e));
  // sin(x + xl) = sin x + xl cos x, cos(x + xl) = cos x - xl sin x
  const double sx = fma(xl, cx0, sx0), cx = fma(-xl, sx0, cx0);
  // 2 pi u = q pi/2 + x: quadrant q mod 4 swaps and negates
  const int k = (int)q & 3;
  const double s0 = k & 1 ? cx : sx, c0 = k & 1 ? sx : cx;
  sn = k & 2 ? -s0 : s0;
  cs = (k + 1) & 2 ? -c0 : c0;
}

}  // namespace qmc
