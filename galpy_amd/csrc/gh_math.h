// Elementary functions written for the force kernels, compiled identically
// for the device and the host (so the host can check them without a GPU).
#pragma once
#include <math.h>

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#endif

// Natural logarithm of y >= 1 (finite), accurate to about 1 ulp, in ~30
// instructions: the library log1p/log routines carry double-double code that
// costs ~3x more.
//
// Argument reduction (Cody-Waite): y = 2^e (1 + f) with 1 + f in
// [sqrt(1/2), sqrt(2)), so f = m - 1 is exact and |f| <= 0.4143.  Then
//   log(1 + f) = 2 atanh(s),  s = f / (2 + f),  |s| <= 0.1716,
//   2 atanh(s) = 2 s + s R(z),  R(z) = sum_{k>=1} 2 z^k / (2k + 1),  z = s^2,
// and 2 s = f - f s, so log(1 + f) = f - s (f - R(z)): the leading term f is
// exact and the correction is at most 0.17 |f|, so rounding in s and R is
// damped.  Nine terms of R leave a truncation error below 2^-56 relative.
// ln 2 is split as LN2_HI (42 significant bits, so e * LN2_HI is exact for
// |e| < 2^11) + LN2_LO.
__host__ __device__ inline double gh_log_ge1(double y) {
  const double LN2_HI = 0x1.62e42fefa3800p-1;
  const double LN2_LO = 0x1.ef35793c76730p-45;
  int e;
  double m = frexp(y, &e);  // y = m 2^e, m in [0.5, 1)
  if (m < 0x1.6a09e667f3bcdp-1) {  // sqrt(1/2)
    m = 2.0 * m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  // 2/(2k+1), k = 9 .. 1
  double R = 0x1.af286bca1af28p-4;
  R = fma(R, z, 0x1.e1e1e1e1e1e1ep-4);
  R = fma(R, z, 0x1.1111111111111p-3);
  R = fma(R, z, 0x1.3b13b13b13b14p-3);
  R = fma(R, z, 0x1.745d1745d1746p-3);
  R = fma(R, z, 0x1.c71c71c71c71cp-3);
  R = fma(R, z, 0x1.2492492492492p-2);
  R = fma(R, z, 0x1.999999999999ap-2);
  R = fma(R, z, 0x1.5555555555555p-1);
  R = R * z;
  const double ed = (double)e;
  const double lf = fma(-s, f - R, f);   // log(1 + f)
  return fma(ed, LN2_HI, fma(ed, LN2_LO, lf));
}

// sin and cos of one argument with a single argument reduction, entirely in
// registers (the library's sincos returns cos through a pointer and carries
// a large-argument reduction with a stack array, which costs scratch-memory
// traffic in every force evaluation of the rotating components).
//
// Reduction (Cody-Waite, with fused multiply-adds): k = nearest integer to
// x 2/pi, r = x - k pi/2 with pi/2 split into three doubles P1 + P2 + P3
// (each the rounding of what remains of pi/2), so that |r| <= pi/4 with an
// absolute error ~1e-16 for |k| < 2^30.  Then the Taylor polynomials
//   sin r = r + r^3 sum_{j=1..9} (-1)^j r^(2j-2) / (2j+1)!,
//   cos r = 1 - r^2/2 + r^4 sum_{j=2..9} (-1)^j r^(2j-4) / (2j)!,
// whose truncation error at |r| = pi/4 is below 1e-22, and the quadrant
// k mod 4 picks the signs and the swap.  Arguments with |k| >= 2^30, and
// non-finite ones, go to the library routines.
struct GhSinCos {
  double s, c;
};

#ifdef __HIP_DEVICE_COMPILE__
__device__ __noinline__ inline GhSinCos gh_sincos_library(double x) {
  return GhSinCos{sin(x), cos(x)};
}
#else
inline GhSinCos gh_sincos_library(double x) { return GhSinCos{sin(x), cos(x)}; }
#endif

__host__ __device__ inline GhSinCos gh_sincos(double x) {
  const double kd = rint(x * 0x1.45f306dc9c883p-1);  // x 2/pi
  if (!(fabs(kd) < 0x1p30)) return gh_sincos_library(x);
  double r = fma(-kd, 0x1.921fb54442d18p+0, x);
  r = fma(-kd, 0x1.1a62633145c07p-54, r);
  r = fma(-kd, -0x1.f1976b7ed8fbcp-110, r);
  const double z = r * r;
  double ps = -0x1.2f49b46814157p-57;  // -1/19!
  ps = fma(ps, z, 0x1.952c77030ad4ap-49);
  ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
  ps = fma(ps, z, 0x1.6124613a86d09p-33);
  ps = fma(ps, z, -0x1.ae64567f544e4p-26);
  ps = fma(ps, z, 0x1.71de3a556c734p-19);
  ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
  ps = fma(ps, z, 0x1.1111111111111p-7);
  ps = fma(ps, z, -0x1.5555555555555p-3);  // -1/3!
  const double sr = fma(r * z, ps, r);
  double pc = -0x1.6827863b97d97p-53;  // -1/18!
  pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
  pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
  pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
  pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
  pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
  pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
  pc = fma(pc, z, 0x1.5555555555555p-5);  // 1/4!
  const double cr = fma(z * z, pc, fma(-0.5, z, 1.0));
  const int q = (int)(long long)kd & 3;
  GhSinCos out;
  out.s = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
  out.c = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
  return out;
}
