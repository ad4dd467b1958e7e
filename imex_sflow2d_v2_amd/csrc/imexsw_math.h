// imexsw_math.h -- elementary functions with a fixed operation sequence.
//
// The settling law (Re^0.687, hindrance^4.65), the entrainment law (Ri^2.4)
// and the source time law (t^1.5, t^0.5, cos) need pow/cos; the complex
// square root of the Jacobian needs hypot.  Vendor math libraries differ in
// the last bits between CPU and GPU, so these are implemented here with
// plain IEEE operations (no FMA contraction) to make the device result
// reproducible bit for bit by any IEEE double implementation that performs
// the same sequence.  Methods: Cody & Waite (1980) argument reduction with
// ln2 split into a 32-bit head and a tail; log via log(m) = 2 atanh(s),
// s = (m-1)/(m+1), m in [sqrt(1/2), sqrt(2)); exp via a degree-14 Taylor
// polynomial on |r| <= ln2/2; cos via reduction by pi/2 (three-part split)
// and Taylor polynomials on |r| <= pi/4.  Accuracy is a few ulp over the
// ranges used (|y log x| < 40).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imexsw {

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;

// The special cases are selected after the fact rather than branched
// around, so that the straight-line code of independent calls can overlap.
__device__ __forceinline__ double det_log(double x) {
  // x > 0, finite
  const int64_t bits0 = __double_as_longlong(x);
  const bool sub = ((bits0 >> 52) & 0x7ff) == 0;  // subnormal: scale by 2^54
  const double xs = sub ? x * 18014398509481984.0 : x;
  const int64_t bits = __double_as_longlong(xs);
  int e = (int)((bits >> 52) & 0x7ff) - (sub ? 54 : 0) - 1022;  // x = m * 2^e, m in [0.5, 1)
  double m = __longlong_as_double((bits & 0x000fffffffffffffLL) | 0x3fe0000000000000LL);
  const bool lo = m < 0.70710678118654752440;
  m = lo ? m * 2.0 : m;
  e -= lo ? 1 : 0;
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  // 1/3 + s2/5 + ... + s2^10/23 by Estrin's scheme (depth 7, not 20)
  const double x2 = s2 * s2, x4 = x2 * x2, x8 = x4 * x4;
  const double a0 = 1.0 / 3.0 + 1.0 / 5.0 * s2, a1 = 1.0 / 7.0 + 1.0 / 9.0 * s2;
  const double a2 = 1.0 / 11.0 + 1.0 / 13.0 * s2, a3 = 1.0 / 15.0 + 1.0 / 17.0 * s2;
  const double a4 = 1.0 / 19.0 + 1.0 / 21.0 * s2;
  const double b0 = a0 + a1 * x2, b1 = a2 + a3 * x2, b2 = a4 + 1.0 / 23.0 * x2;
  const double p = (b0 + b1 * x4) + b2 * x8;
  const double logm = 2.0 * s + 2.0 * s * (s2 * p);
  const double de = (double)e;
  return (de * kLn2Hi + logm) + de * kLn2Lo;
}

__device__ __forceinline__ double det_exp(double y) {
  const bool hi = y > 709.0, lo = y < -745.0;
  const double yc = (hi || lo) ? 0.0 : y;
  const double n = rint(yc * kInvLn2);
  const double r = (yc - n * kLn2Hi) - n * kLn2Lo;
  // sum of r^k / k!, k = 0..14, by Estrin's scheme (depth 8, not 28)
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double a0 = 1.0 + r, a1 = 0.5 + 1.0 / 6.0 * r;
  const double a2 = 1.0 / 24.0 + 1.0 / 120.0 * r, a3 = 1.0 / 720.0 + 1.0 / 5040.0 * r;
  const double a4 = 1.0 / 40320.0 + 1.0 / 362880.0 * r;
  const double a5 = 1.0 / 3628800.0 + 1.0 / 39916800.0 * r;
  const double a6 = 1.0 / 479001600.0 + 1.0 / 6227020800.0 * r;
  const double b0 = a0 + a1 * r2, b1 = a2 + a3 * r2, b2 = a4 + a5 * r2;
  const double b3 = a6 + 1.0 / 87178291200.0 * r2;
  const double p = (b0 + b1 * r4) + (b2 + b3 * r4) * r8;
  // scale by 2^n in two steps to stay in range near the limits
  const int ni = (int)n;
  const int n1 = ni / 2, n2 = ni - n1;
  const double s1 = __longlong_as_double((int64_t)(1023 + n1) << 52);
  const double s2 = __longlong_as_double((int64_t)(1023 + n2) << 52);
  const double v = (p * s1) * s2;
  return hi ? __longlong_as_double(0x7ff0000000000000LL) : lo ? 0.0 : v;
}

// x**y for x >= 0 (the reference only raises non-negative quantities)
__device__ __forceinline__ double det_pow(double x, double y) {
  const bool z = x == 0.0, one = (x == 1.0 || y == 0.0), nan = (x != x || y != y);
  const double v = det_exp(y * det_log((z || one || nan) ? 2.0 : x));
  return z ? ((y > 0.0) ? 0.0 : (y == 0.0 ? 1.0 : __longlong_as_double(0x7ff0000000000000LL)))
           : one ? 1.0 : nan ? x + y : v;
}

__device__ __forceinline__ double det_cos(double x) {
  // |x| <= ~1e3; pi/2 = P1 + P2 + P3
  const double P1 = 1.57079632673412561417e+00, P2 = 6.07710050650619224932e-11,
               P3 = 2.02226624879595063154e-21;
  const double k = rint(x * 6.36619772367581382433e-01);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  const double r2 = r * r;
  double c = -1.0 / 87178291200.0;
  c = c * r2 + 1.0 / 479001600.0;
  c = c * r2 - 1.0 / 3628800.0;
  c = c * r2 + 1.0 / 40320.0;
  c = c * r2 - 1.0 / 720.0;
  c = c * r2 + 1.0 / 24.0;
  c = c * r2 - 0.5;
  c = c * r2 + 1.0;
  double s = -1.0 / 1307674368000.0;
  s = s * r2 + 1.0 / 6227020800.0;
  s = s * r2 - 1.0 / 39916800.0;
  s = s * r2 + 1.0 / 362880.0;
  s = s * r2 - 1.0 / 5040.0;
  s = s * r2 + 1.0 / 120.0;
  s = s * r2 - 1.0 / 6.0;
  s = r + r * (r2 * s);
  const int q = ((int)k) & 3;
  return q == 0 ? c : q == 1 ? -s : q == 2 ? -c : s;
}

__device__ __forceinline__ double det_hypot(double x, double y) {
  double a = fabs(x), b = fabs(y);
  if (a < b) {
    double t = a;
    a = b;
    b = t;
  }
  if (a == 0.0) return 0.0;
  const double r = b / a;
  return a * sqrt(1.0 + r * r);
}

}  // namespace imexsw
