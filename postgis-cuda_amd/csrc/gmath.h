// Portable FP64 elementary functions for the geodesic paths: sin, cos, tan, atan, atan2.
//
// The same source compiles for the host (g++, used by the CPU restatement the engine is tested against) and for
// gfx950 (hipcc); with -ffp-contract=off on both sides every operation below is one IEEE-754 rounding, so a value
// computed on the GPU has exactly the bits of the same value computed on the host. The library math of either
// compiler would differ in the last ulp now and then, which would make threshold decisions (dwithin, erosion)
// and reported distances differ between the two.
//
// Methods (textbook):
//  * sin / cos: Cody–Waite reduction by pi/2 split into three parts (the first two carry 33 significant bits, so
//    k * part is exact for |k| < 2^20), then the Taylor series of sin and cos on |r| <= pi/4 evaluated by Horner's
//    rule in r^2 (terms to r^19 / r^20: truncation below 1e-19 relative). Valid for |x| < 2^20; NaN beyond.
//  * atan: reflection atan(x) = pi/2 - atan(1/x) for |x| > 1, then the addition formula around c = k/8
//    (atan(x) = atan(c) + atan((x - c) / (1 + x c)), |reduced| <= 1/16), Taylor series to t^17; atan(k/8) as a
//    two-double (hi + lo) constant.
//  * atan2 from atan of the smaller-over-larger ratio with the quadrant fixed by the signs.
// Accuracy: sin, cos, atan, atan2 within 2 ulp of the exact function, tan (a quotient) within 3, over the domains
// used here (pinned against mpmath by the tests).
// The constants are the correctly rounded values of pi/2 (split), pi, and atan(k/8), computed with mpmath.
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GSJ_PORT __host__ __device__ inline
#else
#define GSJ_PORT inline
#endif

namespace gsj {
namespace gm {

constexpr double PIO2_1 = 1.5707963267341256;       // pi/2 to 33 bits
constexpr double PIO2_2 = 6.077100506303966e-11;    // next 33 bits
constexpr double PIO2_3 = 2.0222662487959506e-21;   // remainder, rounded
constexpr double TWO_OVER_PI = 0.6366197723675814;
constexpr double PI_HI = 3.141592653589793, PI_LO = 1.2246467991473532e-16;     // pi = hi + lo
constexpr double PIO2_HI = 1.5707963267948966, PIO2_LO = 6.123233995736766e-17;  // pi/2 = hi + lo

GSJ_PORT bool is_nan(double x) { return x != x; }

// sin and cos of x (one reduction for both)
GSJ_PORT void sincos(double x, double& s, double& c) {
  if (is_nan(x) || !(x < 1048576.0 && x > -1048576.0)) { s = c = NAN; return; }
  const double k = rint(x * TWO_OVER_PI);
  const double r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
  const double z = r * r;
  const double ps = -1.0 / 6.0 +
                    z * (1.0 / 120.0 +
                         z * (-1.0 / 5040.0 +
                              z * (1.0 / 362880.0 +
                                   z * (-1.0 / 39916800.0 +
                                        z * (1.0 / 6227020800.0 +
                                             z * (-1.0 / 1307674368000.0 + z * (1.0 / 355687428096000.0 + z * (-1.0 / 121645100408832000.0))))))));
  const double pc = 1.0 / 24.0 +
                    z * (-1.0 / 720.0 +
                         z * (1.0 / 40320.0 +
                              z * (-1.0 / 3628800.0 +
                                   z * (1.0 / 479001600.0 +
                                        z * (-1.0 / 87178291200.0 +
                                             z * (1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0 + z * (1.0 / 2432902008176640000.0))))))));
  const double sr = r + (r * z) * ps;
  const double cr = 1.0 - (0.5 * z - (z * z) * pc);
  const int q = (int)((int64_t)k & 3);
  if (q == 0) { s = sr; c = cr; }
  else if (q == 1) { s = cr; c = -sr; }
  else if (q == 2) { s = -sr; c = -cr; }
  else { s = -cr; c = sr; }
}
GSJ_PORT double sin(double x) { double s, c; sincos(x, s, c); return s; }
GSJ_PORT double cos(double x) { double s, c; sincos(x, s, c); return c; }
GSJ_PORT double tan(double x) { double s, c; sincos(x, s, c); return s / c; }

// atan(k/8) = hi + lo, k = 0..8
GSJ_PORT void atan_tab(int k, double& hi, double& lo) {
  switch (k) {
    case 1: hi = 0.12435499454676144; lo = -3.1253241424539383e-18; return;
    case 2: hi = 0.24497866312686414; lo = 1.0698755618734451e-17; return;
    case 3: hi = 0.35877067027057225; lo = -2.4623815582638635e-17; return;
    case 4: hi = 0.4636476090008061; lo = 2.2698777452961687e-17; return;
    case 5: hi = 0.5585993153435624; lo = -5.4556305485916264e-18; return;
    case 6: hi = 0.6435011087932844; lo = 1.5834785051444286e-17; return;
    case 7: hi = 0.7188299996216245; lo = -2.1478388444456983e-17; return;
    case 8: hi = 0.7853981633974483; lo = 3.061616997868383e-17; return;
    default: hi = 0.0; lo = 0.0; return;
  }
}

// atan(x) for x >= 0 or NaN
GSJ_PORT double atan_pos(double ax) {
  if (is_nan(ax)) return ax;
  const bool inv = ax > 1.0;
  if (inv) ax = 1.0 / ax;
  const int k = (int)rint(ax * 8.0);
  const double cc = k * 0.125;
  const double t = (ax - cc) / (1.0 + ax * cc);
  const double z = t * t;
  const double p = -1.0 / 3.0 +
                   z * (1.0 / 5.0 + z * (-1.0 / 7.0 + z * (1.0 / 9.0 + z * (-1.0 / 11.0 + z * (1.0 / 13.0 + z * (-1.0 / 15.0 + z * (1.0 / 17.0)))))));
  double hi, lo;
  atan_tab(k, hi, lo);
  const double r = hi + (lo + (t + (t * z) * p));
  return inv ? (PIO2_HI - r) + PIO2_LO : r;
}
GSJ_PORT double atan(double x) {
  if (is_nan(x)) return x;
  const double r = atan_pos(fabs(x));
  return x < 0 ? -r : (x > 0 ? r : x);
}
// atan2(y, x): the angle of (x, y) in (-pi, pi]; signed zeros as IEEE (atan2(+-0, -0) = +-pi); both infinite: NaN
GSJ_PORT double atan2(double y, double x) {
  if (is_nan(x) || is_nan(y)) return NAN;
  const double ax = fabs(x), ay = fabs(y);
  double a;
  if (ay <= ax) a = ax == 0 ? 0.0 : atan_pos(ay / ax);
  else a = (PIO2_HI - atan_pos(ax / ay)) + PIO2_LO;
  if (__builtin_signbit(x)) a = (PI_HI - a) + PI_LO;
  return __builtin_signbit(y) ? -a : a;
}

}  // namespace gm
}  // namespace gsj
