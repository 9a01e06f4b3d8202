// IAS15 (wez_ias15.c), galpy's 15th-order Gauss-Radau integrator (Everhart
// 1985; Rein & Spiegel 2015), on the device: one orbit per lane, the
// rectangular positions x and velocities v, the force's polynomial
// coefficients B (7 per component), their divided-difference form G and the
// previous step's predicted B, E.  Only the start-of-step force F0 is kept:
// each substep's force enters G at once and is not needed again.  Tables are
// compile-time constants; every loop that indexes them or B/E/G is unrolled,
// so they stay immediates and registers.  The operation order of the
// reference's polynomial updates is kept (divisions by the integer
// denominators included), so results agree with it to rounding.
#pragma once
#include "rk_dev.h"

namespace gh_radau {
// Gauss-Radau spacings on [0, 1]: 0, then the 7 nodes of the 8-point Radau
// rule (wez_ias15.c:43-52)
constexpr double h[8] = {0.0,
                         0.05626256053692215,
                         0.18024069173689236,
                         0.35262471711316964,
                         0.54715362633055538,
                         0.73421017721541053,
                         0.88532094683909577,
                         0.97752061356128750};
// rr[k][j] = 1 / (h[k] - h[j]) for j < k (wez_ias15.c:113-148)
constexpr double rr[8][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {17.773808914078, 0, 0, 0, 0, 0, 0},
    {5.548136718537217, 8.065938648381888, 0, 0, 0, 0, 0},
    {2.835876078644439, 3.3742499769626355, 5.801001559264062, 0, 0, 0, 0},
    {1.8276402675175978, 2.0371118353585844, 2.725442211808226, 5.140624105810932, 0, 0, 0},
    {1.3620078160624696, 1.4750402175604116, 1.8051535801402514, 2.620644926387035,
     5.3459768998711095, 0, 0},
    {1.1295338753367898, 1.2061876660584456, 1.418278263734739, 1.8772424961868102,
     2.957116017290456, 6.617662013702422, 0},
    {1.0229963298234868, 1.0854721939386425, 1.2542646222818779, 1.6002665494908161,
     2.3235983002196945, 4.109975778344558, 10.846026190236847}};
// G -> B: a change dG of G[k-1] adds cc[k-1][j] dG to B[j], j < k-1, and dG
// to B[k-1] (wez_ias15.c:54-81)
constexpr double cc[7][7] = {
    {1, 0, 0, 0, 0, 0, 0},
    {-0.05626256053692215, 1, 0, 0, 0, 0, 0},
    {0.01014080283006363, -0.23650325227381452, 1, 0, 0, 0, 0},
    {-0.003575897729251617, 0.09353769525946207, -0.5891279693869842, 1, 0, 0, 0},
    {0.001956565409947221, -0.05475538688906869, 0.4158812000823069, -1.1362815957175396, 1, 0,
     0},
    {-0.0014365302363708915, 0.042158527721268706, -0.3600995965020568, 1.250150711840691,
     -1.87049177293295, 1, 0},
    {0.0012717903090268678, -0.03876035791590677, 0.360962243452846, -1.466884208400427,
     2.906136259308429, -2.7558127197720457, 1}};
// B -> G: G[j] = sum over k >= j of dd[k][j] B[k] (wez_ias15.c:83-111)
constexpr double dd[7][7] = {
    {1, 0, 0, 0, 0, 0, 0},
    {0.05626256053692215, 1, 0, 0, 0, 0, 0},
    {0.0031654757181708297, 0.23650325227381452, 1, 0, 0, 0, 0},
    {0.00017809776922174343, 0.04579298550602792, 0.5891279693869842, 1, 0, 0, 0},
    {1.002023652232913e-05, 0.008431857153525702, 0.25353406905456927, 1.1362815957175396, 1, 0,
     0},
    {5.637641639318209e-07, 0.001529784002500466, 0.09783423653244401, 0.8752546646840911,
     1.87049177293295, 1, 0},
    {3.171881540176138e-08, 0.0002762930909826477, 0.03602855398373646, 0.5767330002770787,
     2.2485887607691595, 2.7558127197720457, 1}};
constexpr double kDeltaB = 1e-16;    // predictor-corrector convergence
constexpr double kPrecision = 1e-9;  // step-size precision parameter
constexpr int kMaxPasses = 12;       // predictor-corrector passes per step
}  // namespace gh_radau

// a^(1/7): scale a into [1e-7, 1e2] by factors of 1e7, then 20 Newton steps
// from 1 (wez_ias15.c:587-603); zero, subnormal, inf and NaN are not scaled.
__device__ __forceinline__ bool gh_isnormal(double a) {
  const double m = fabs(a);
  return m >= 0x1p-1022 && m <= 0x1.fffffffffffffp1023;
}
__device__ __forceinline__ double gh_seventhroot(double a) {
  double scale = 1.0;
  while (a < 1e-7 && gh_isnormal(a)) {
    scale *= 0.1;
    a *= 1e7;
  }
  while (a > 1e2 && gh_isnormal(a)) {
    scale *= 10.0;
    a *= 1e-7;
  }
  double x = 1.0;
#pragma unroll 1
  for (int k = 0; k < 20; k++) {
    const double x6 = x * x * x * x * x * x;
    x += (a / x6 - x) / 7.0;
  }
  return x * scale;
}

// Position and velocity at fraction hn of a step of length T
// (wez_ias15.c:441-476).
template <int D>
__device__ __forceinline__ void gh_ias15_position(double* y, const double* y1, const double* v,
                                                  double hn, double T, const double* F0,
                                                  const double (*B)[D]) {
#pragma unroll
  for (int i = 0; i < D; i++) {
    const double poly =
        F0[i] / 2.0 +
        hn * (B[0][i] / 6.0 +
              hn * (B[1][i] / 12.0 +
                    hn * (B[2][i] / 20.0 +
                          hn * (B[3][i] / 30.0 +
                                hn * (B[4][i] / 42.0 +
                                      hn * (B[5][i] / 56.0 + hn * (B[6][i] / 72.0)))))));
    y[i] = y1[i] + (v[i] * hn * T) + ((hn * T) * (hn * T)) * poly;
  }
}
template <int D>
__device__ __forceinline__ void gh_ias15_velocity(double* v, const double* v1, double hn, double T,
                                                  const double* F0, const double (*B)[D]) {
#pragma unroll
  for (int i = 0; i < D; i++) {
    const double poly =
        F0[i] + hn * (B[0][i] / 2.0 +
                      hn * (B[1][i] / 3.0 +
                            hn * (B[2][i] / 4.0 +
                                  hn * (B[3][i] / 5.0 +
                                        hn * (B[4][i] / 6.0 +
                                              hn * (B[5][i] / 7.0 + hn * (B[6][i] / 8.0)))))));
    v[i] = v1[i] + (hn * T * poly);
  }
}

// The force at the positions x (D = 3 or 2 components)
template <class F, class V, int D>
__device__ __forceinline__ void gh_ias15_force(const F& T, double t, const double* x, double* a) {
  gh_rk_put(a, gh_accel(T, t, gh_vec_at<V>(x)));
}

// One IAS15 step attempt of length dt_temp from (x, v) at force time
// to_temp: the predictor-corrector loop (wez_ias15.c:312-352) and the step
// the next attempt would take (wez_ias15.c:354-370).  B and G are updated in
// place; F0 receives the start-of-step force.
template <class F, class V, int D>
__device__ __forceinline__ double gh_ias15_attempt(const F& T, const double* x, const double* v,
                                                   double dt_temp, double to_temp, double* F0,
                                                   double (*B)[D], double (*G)[D]) {
  using namespace gh_radau;
  gh_ias15_force<F, V, D>(T, to_temp, x, F0);
#pragma unroll
  for (int j = 0; j < 7; j++) {
#pragma unroll
    for (int i = 0; i < D; i++) {
      double g = dd[j][j] * B[j][i];
#pragma unroll
      for (int k = j + 1; k < 7; k++) g = g + dd[k][j] * B[k][i];
      G[j][i] = g;
    }
  }
  double ierr = kDeltaB + 1.0;
#pragma unroll 1
  for (int pass = 0; pass < kMaxPasses && !(ierr < kDeltaB); pass++) {
    double max_dg = 0.0, max_a = 0.0;
#pragma unroll
    for (int k = 1; k < 8; k++) {
      double xs[D], a[D];
      gh_ias15_position<D>(xs, x, v, h[k], dt_temp, F0, B);
      gh_ias15_force<F, V, D>(T, to_temp, xs, a);
#pragma unroll
      for (int i = 0; i < D; i++) {
        double g = (a[i] - F0[i]) * rr[k][0];
#pragma unroll
        for (int j = 1; j < k; j++) g = (g - G[j - 1][i]) * rr[k][j];
        const double dg = g - G[k - 1][i];
        G[k - 1][i] = g;
#pragma unroll
        for (int j = 0; j < k - 1; j++) B[j][i] += dg * cc[k - 1][j];
        B[k - 1][i] += dg;
        if (k == 7) {
          if (fabs(dg) > max_dg) max_dg = fabs(dg);
          if (fabs(F0[i]) > max_a) max_a = fabs(F0[i]);
        }
      }
    }
    ierr = max_dg / max_a;
  }
  double max_b6 = 0.0, max_a = 0.0;
#pragma unroll
  for (int i = 0; i < D; i++) {
    if (fabs(B[6][i]) > max_b6) max_b6 = fabs(B[6][i]);
    if (fabs(F0[i]) > max_a) max_a = fabs(F0[i]);
  }
  return dt_temp * gh_seventhroot(kPrecision / (max_b6 / max_a));
}

// Advance (x, v) over the accepted step and predict the next step's B from
// the binomial expansion at step ratio 1 (wez_ias15.c:407-410, 531-552).
template <int D>
__device__ __forceinline__ void gh_ias15_advance(double* x, double* v, double dt_temp,
                                                 const double* F0, double (*B)[D],
                                                 double (*E)[D]) {
  gh_ias15_position<D>(x, x, v, 1.0, dt_temp, F0, B);
  gh_ias15_velocity<D>(v, v, 1.0, dt_temp, F0, B);
#pragma unroll
  for (int i = 0; i < D; i++) {
    double en[7];
    en[0] = (B[0][i] + 2.0 * B[1][i] + 3.0 * B[2][i] + 4.0 * B[3][i] + 5.0 * B[4][i] +
             6.0 * B[5][i] + 7.0 * B[6][i]);
    en[1] = (B[1][i] + 3.0 * B[2][i] + 6.0 * B[3][i] + 10.0 * B[4][i] + 15.0 * B[5][i] +
             21.0 * B[6][i]);
    en[2] = (B[2][i] + 4.0 * B[3][i] + 10.0 * B[4][i] + 20.0 * B[5][i] + 35.0 * B[6][i]);
    en[3] = (B[3][i] + 5.0 * B[4][i] + 15.0 * B[5][i] + 35.0 * B[6][i]);
    en[4] = (B[4][i] + 6.0 * B[5][i] + 21.0 * B[6][i]);
    en[5] = (B[5][i] + 7.0 * B[6][i]);
    en[6] = B[6][i];
#pragma unroll
    for (int k = 0; k < 7; k++) {
      const double bd = B[k][i] - E[k][i];
      E[k][i] = en[k];
      B[k][i] = en[k] + bd;
    }
  }
}
