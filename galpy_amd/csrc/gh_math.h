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
