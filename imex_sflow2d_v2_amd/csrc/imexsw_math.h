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

__device__ __forceinline__ double det_log(double x) {
  // x > 0, finite
  int64_t bits = __double_as_longlong(x);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) {  // subnormal: scale by 2^54
    x = x * 18014398509481984.0;
    bits = __double_as_longlong(x);
    e = (int)((bits >> 52) & 0x7ff) - 54;
  }
  e -= 1022;  // x = m * 2^e, m in [0.5, 1)
  double m = __longlong_as_double((bits & 0x000fffffffffffffLL) | 0x3fe0000000000000LL);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  const double logm = 2.0 * s + 2.0 * s * (s2 * p);
  const double de = (double)e;
  return (de * kLn2Hi + logm) + de * kLn2Lo;
}

__device__ __forceinline__ double det_exp(double y) {
  if (y > 709.0) return __longlong_as_double(0x7ff0000000000000LL);
  if (y < -745.0) return 0.0;
  const double n = rint(y * kInvLn2);
  const double r = (y - n * kLn2Hi) - n * kLn2Lo;
  double p = 1.0 / 87178291200.0;  // 1/14!
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  // scale by 2^n in two steps to stay in range near the limits
  const int ni = (int)n;
  const int n1 = ni / 2, n2 = ni - n1;
  const double s1 = __longlong_as_double((int64_t)(1023 + n1) << 52);
  const double s2 = __longlong_as_double((int64_t)(1023 + n2) << 52);
  return (p * s1) * s2;
}

// x**y for x >= 0 (the reference only raises non-negative quantities)
__device__ __forceinline__ double det_pow(double x, double y) {
  if (x == 0.0) return (y > 0.0) ? 0.0 : (y == 0.0 ? 1.0 : __longlong_as_double(0x7ff0000000000000LL));
  if (x == 1.0 || y == 0.0) return 1.0;
  if (x != x || y != y) return x + y;
  return det_exp(y * det_log(x));
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
