// rhy_math.hpp -- x**y for the exponents of the ideal-gas Riemann solver.
//
// The Riemann solver raises positive pressure and density ratios to
// gamma-derived constant exponents ((g-1)/2g, -(g+1)/2g, 1/g, 2/(g-1),
// 2g/(g-1)), often two exponents of the same base in a row.  pow_pos()
// evaluates x**y as exp(y * log(x)) with log and exp carried in double-double
// arithmetic (error of y*log(x) ~2^-66 relative), so the result is the
// correctly rounded value except when the exact value lies within ~2^-60
// relative of a rounding boundary; pow_pos2() shares the logarithm between
// two exponents.
//
// Method (textbook, e.g. Muller, "Elementary Functions", ch. 6):
//   log: x = 2^e m, m in [1/sqrt2, sqrt2); log m = 2 atanh(s), s = (m-1)/(m+1),
//        2s in double-double, the odd series tail 2 s^3 (1/3 + s^2/5 + ...)
//        in double (it is below 1% of the result).
//   exp: t = k ln2 + r with |r| <= ln2/2 (Cody-Waite split of ln2), then
//        e^r = 1 + r + r^2 (1/2 + r/6 + ... + r^13/15!) with the low part of
//        t folded into the linear term, and a final scale by 2^k.
// Only x > 0 finite and |y log x| < 700 reach the fast path; everything
// else (x <= 0, NaN, inf, overflow/underflow range) goes to the library pow.
// Explicit fma() calls are part of the method (error-free products); they
// are not contraction and are unaffected by -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace rhy {

#define RHY_HD __host__ __device__ __forceinline__

struct dd {
  double hi, lo;
};

RHY_HD dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  return {s, e};
}

RHY_HD dd fast_two_sum(double a, double b) {  // |a| >= |b|
  double s = a + b;
  return {s, b - (s - a)};
}

RHY_HD dd two_prod(double a, double b) {
  double p = a * b;
  return {p, fma(a, b, -p)};
}

// log(x) for x > 0 finite (normal or subnormal), as a double-double.
RHY_HD dd log_dd(double x) {
  // ln2 split: LN2_HI has 32 trailing zero bits so e * LN2_HI is exact
  const double LN2_HI = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
  const double LN2_LO = 1.90821492927058770002e-10;  // ln2 - LN2_HI
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  // s = (m - 1) / (m + 1) in double-double; m - 1 is exact (Sterbenz)
  const double f = m - 1.0;
  const dd d = two_sum(m, 1.0);
  const double s_hi = f / d.hi;
  double r = fma(-s_hi, d.hi, f);
  r = fma(-s_hi, d.lo, r);
  const double s_lo = r / d.hi;
  // tail: 2 s^3 (1/3 + z/5 + z^2/7 + ...), z = s^2 <= 0.0295
  dd z = two_prod(s_hi, s_hi);
  z.lo += 2.0 * s_hi * s_lo;
  double q = 2.0 / 25.0;
  q = fma(q, z.hi, 2.0 / 23.0);
  q = fma(q, z.hi, 2.0 / 21.0);
  q = fma(q, z.hi, 2.0 / 19.0);
  q = fma(q, z.hi, 2.0 / 17.0);
  q = fma(q, z.hi, 2.0 / 15.0);
  q = fma(q, z.hi, 2.0 / 13.0);
  q = fma(q, z.hi, 2.0 / 11.0);
  q = fma(q, z.hi, 2.0 / 9.0);
  q = fma(q, z.hi, 2.0 / 7.0);
  q = fma(q, z.hi, 2.0 / 5.0);
  // 2/3 carried in double-double: the product q*z is < 1.2% of 2/3
  const double TWO_THIRDS_HI = 6.66666666666666629659e-01;
  const double TWO_THIRDS_LO = 3.70074341541718826e-17;
  dd c = two_sum(TWO_THIRDS_HI, q * z.hi);
  c.lo += TWO_THIRDS_LO;
  // s^3 and tail = s^3 * c in double-double
  dd s3 = two_prod(s_hi, z.hi);
  s3.lo += s_hi * z.lo + s_lo * z.hi;
  const dd t1 = two_prod(s3.hi, c.hi);
  const double tail_hi = t1.hi;
  const double tail_lo = t1.lo + s3.hi * c.lo + s3.lo * c.hi;
  // log m = 2 s + tail
  dd acc = fast_two_sum(2.0 * s_hi, tail_hi);
  acc.lo += 2.0 * s_lo + tail_lo;
  // + e * ln2
  const double de = (double)e;
  dd el = fast_two_sum(de * LN2_HI, de * LN2_LO);
  dd sum = two_sum(el.hi, acc.hi);
  sum.lo += el.lo + acc.lo;
  return fast_two_sum(sum.hi, sum.lo);
}

// exp(t_hi + t_lo) for |t_hi| < 700.
RHY_HD double exp_dd(double t_hi, double t_lo) {
  const double INV_LN2 = 1.44269504088896338700e+00;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  const double kd = rint(t_hi * INV_LN2);
  const int k = (int)kd;
  // r = t - k ln2: t_hi - k*LN2_HI is exact (|k| < 2^11, Sterbenz)
  const double a = t_hi - kd * LN2_HI;
  const dd p = two_prod(kd, LN2_LO);
  dd r = two_sum(a, -p.hi);
  r.lo += -p.lo + t_lo;
  const dd rn = fast_two_sum(r.hi, r.lo);
  const double x = rn.hi;
  // e^x = 1 + x + x^2/2 + x^3/6 + x^4 q(x), |x| <= 0.3466; the terms up
  // to x^3 in double-double, x^4 q (< 6.2e-4) in double
  double q = 1.0 / 20922789888000.0;  // 1/16!
  q = fma(q, x, 1.0 / 1307674368000.0);
  q = fma(q, x, 1.0 / 87178291200.0);
  q = fma(q, x, 1.0 / 6227020800.0);
  q = fma(q, x, 1.0 / 479001600.0);
  q = fma(q, x, 1.0 / 39916800.0);
  q = fma(q, x, 1.0 / 3628800.0);
  q = fma(q, x, 1.0 / 362880.0);
  q = fma(q, x, 1.0 / 40320.0);
  q = fma(q, x, 1.0 / 5040.0);
  q = fma(q, x, 1.0 / 720.0);
  q = fma(q, x, 1.0 / 120.0);
  q = fma(q, x, 1.0 / 24.0);
  const double SIXTH_HI = 1.66666666666666657415e-01;
  const double SIXTH_LO = 9.25185853854297066e-18;
  const dd x2 = two_prod(x, x);
  dd x3 = two_prod(x2.hi, x);
  x3.lo += x2.lo * x;
  dd t3 = two_prod(x3.hi, SIXTH_HI);
  t3.lo += x3.hi * SIXTH_LO + x3.lo * SIXTH_HI;
  const double t4 = x2.hi * x2.hi * q;
  // s = x + x^2/2 + x^3/6 + x^4 q;  e^(x + lo) = e^x (1 + lo)
  dd s = fast_two_sum(x, 0.5 * x2.hi);
  dd s2 = two_sum(s.hi, t3.hi);
  s2.lo += s.lo + 0.5 * x2.lo + t3.lo + t4;
  dd one = fast_two_sum(1.0, s2.hi);
  one.lo += s2.lo;
  double res = one.hi + (one.lo + one.hi * rn.lo);
  return ldexp(res, k);
}

// x**y for x > 0; falls back to pow() outside the fast path.
RHY_HD double pow_pos(double x, double y) {
  if (!(x > 0.0) || !(x < INFINITY)) return pow(x, y);
  const dd l = log_dd(x);
  const dd t = two_prod(y, l.hi);
  const double t_lo = t.lo + y * l.lo;
  if (!(fabs(t.hi) < 700.0)) return pow(x, y);
  return exp_dd(t.hi, t_lo);
}

// x**y1 and x**y2 sharing log(x).
RHY_HD void pow_pos2(double x, double y1, double y2, double& r1, double& r2) {
  if (!(x > 0.0) || !(x < INFINITY)) {
    r1 = pow(x, y1);
    r2 = pow(x, y2);
    return;
  }
  const dd l = log_dd(x);
  const dd t1 = two_prod(y1, l.hi);
  const dd t2 = two_prod(y2, l.hi);
  r1 = fabs(t1.hi) < 700.0 ? exp_dd(t1.hi, t1.lo + y1 * l.lo) : pow(x, y1);
  r2 = fabs(t2.hi) < 700.0 ? exp_dd(t2.hi, t2.lo + y2 * l.lo) : pow(x, y2);
}

}  // namespace rhy
