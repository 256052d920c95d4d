// nb_sincos.h -- sine and cosine of one argument for the Ewald phases.
//
// Cody & Waite argument reduction by pi/2, with pi/2 split into three
// doubles (hi + mid + lo, from pi to 50 digits by Machin's formula) and the
// remainder r = x - n pi/2 formed with fused multiply-adds, then the Taylor
// series of sin and cos on |r| <= pi/4 through r^17 and r^16 (truncation
// below 2e-18), evaluated by Horner's rule in r^2; the quadrant n mod 4
// selects and signs the pair (n taken as a 32-bit integer: |x| < 3e9).
// About 30 FP64 operations; absolute error a few 1e-16 for |x| up to ~1e5
// (the Ewald phases k.r stay below ~3 pi kmax).
// Usable from host C++ (for its tests) and from HIP device code.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define NB_HD __host__ __device__
#else
#define NB_HD
#endif

NB_HD static inline void nb_sincos(double x, double *sp, double *cp) {
  const double n = rint(x * 0.6366197723675814);  // 2/pi
  double r = fma(-n, 1.5707963267948966, x);
  r = fma(-n, 6.123233995736766e-17, r);
  r = fma(-n, -1.4973849048591698e-33, r);
  const double r2 = r * r;
  // sin r = r + r^3 (-1/3! + r^2 (1/5! - ...)),  cos r = 1 + r^2 (-1/2! + ...)
  double ps = 2.8114572543455206e-15;
  ps = fma(ps, r2, -7.647163731819816e-13);
  ps = fma(ps, r2, 1.6059043836821613e-10);
  ps = fma(ps, r2, -2.505210838544172e-08);
  ps = fma(ps, r2, 2.7557319223985893e-06);
  ps = fma(ps, r2, -0.0001984126984126984);
  ps = fma(ps, r2, 0.008333333333333333);
  ps = fma(ps, r2, -0.16666666666666666);
  double pc = 4.779477332387385e-14;
  pc = fma(pc, r2, -1.1470745597729725e-11);
  pc = fma(pc, r2, 2.08767569878681e-09);
  pc = fma(pc, r2, -2.755731922398589e-07);
  pc = fma(pc, r2, 2.48015873015873e-05);
  pc = fma(pc, r2, -0.001388888888888889);
  pc = fma(pc, r2, 0.041666666666666664);
  pc = fma(pc, r2, -0.5);
  const double s = fma(r * r2, ps, r);
  const double c = fma(r2, pc, 1.0);
  const int q = (int)n & 3;  // n is integral and |n| < 2^31 for |x| < 3e9
  const double a = (q & 1) ? c : s, b = (q & 1) ? s : c;
  *sp = (q & 2) ? -a : a;
  *cp = ((q + 1) & 2) ? -b : b;
}
