// Runge-Kutta integrators of galpy (bovy_rk.c) on the device: RK4 and RK6
// with galpy's fixed steps and step estimators, one orbit per lane, on the
// rectangular state y = [q, p] (N = 2 D doubles; the variational systems of
// integrate_dxdv in galpy's row layout) with the derivative
// evalRectDeriv / evalPlanarRectDeriv (integrateFullOrbit.c:1097-1105,
// integratePlanarOrbit.c:795-803): dq/dt = p, dp/dt = the force.  The stage
// arithmetic keeps the reference's order of operations (e.g. yn1 += dt a / 6
// is yn1 + (dt a) / 6), so results agree with it to rounding.
#pragma once
#include "integrators_dev.h"

__device__ __forceinline__ void gh_rk_put(double* a, GhVec3 v) {
  a[0] = v.x;
  a[1] = v.y;
  a[2] = v.z;
}
__device__ __forceinline__ void gh_rk_put(double* a, GhVec2 v) {
  a[0] = v.x;
  a[1] = v.y;
}

// a = dy/dt at (t, y), in the row layout of GhPhase<V>: (q, p) for plain
// orbits, galpy's (q, p | dq, dp) for the variational systems
// (evalRectDeriv_dxdv, integrateFullOrbit.c:1496-1600: the deviation's
// position derivative is its velocity, its velocity's is K dq)
template <class F, class V>
__device__ __forceinline__ void gh_rk_deriv(const F& T, double t, const double* y, double* a) {
  GhPhase<V>::set(a, GhPhase<V>::p(y), gh_accel_v(T, t, GhPhase<V>::q(y), GhPhase<V>::p(y)));
}

// bovy_rk4_onestep, bovy_rk.c:142-169 (yn1 equals yn on entry)
template <class F, class V>
__device__ __forceinline__ void gh_rk4_onestep(const F& T, const double* yn, double* yn1, double tn,
                                               double dt) {
  constexpr int N = GhPhase<V>::N;
  double ynk[N], a[N];
  gh_rk_deriv<F, V>(T, tn, yn, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += dt * a[i] / 6.0;
    ynk[i] = yn[i] + dt * a[i] / 2.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 2.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += dt * a[i] / 3.0;
    ynk[i] = yn[i] + dt * a[i] / 2.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 2.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += dt * a[i] / 3.0;
    ynk[i] = yn[i] + dt * a[i];
  }
  gh_rk_deriv<F, V>(T, tn + dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) yn1[i] += dt * a[i] / 6.0;
}

// bovy_rk6_onestep, bovy_rk.c:270-326 (seven evaluations; k6 reuses k5's
// storage as the reference does)
template <class F, class V>
__device__ __forceinline__ void gh_rk6_onestep(const F& T, const double* yn, double* yn1, double tn,
                                               double dt) {
  constexpr int N = GhPhase<V>::N;
  double ynk[N], a[N], k1[N], k2[N], k3[N], k4[N], k5[N];
  gh_rk_deriv<F, V>(T, tn, yn, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += 11.0 * dt * a[i] / 120.0;
    k1[i] = dt * a[i];
    ynk[i] = yn[i] + k1[i] / 3.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 3.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    k2[i] = dt * a[i];
    ynk[i] = yn[i] + 2.0 * k2[i] / 3.0;
  }
  gh_rk_deriv<F, V>(T, tn + 2.0 * dt / 3.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += 81.0 * dt * a[i] / 120.0;
    k3[i] = dt * a[i];
    ynk[i] = yn[i] + (k1[i] + 4.0 * k2[i] - k3[i]) / 12.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 3.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] += 81.0 * dt * a[i] / 120.0;
    k4[i] = dt * a[i];
    ynk[i] = yn[i] + (-k1[i] + 18.0 * k2[i] - 3.0 * k3[i] - 6.0 * k4[i]) / 16.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 2.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] -= 32.0 * dt * a[i] / 120.0;
    k5[i] = dt * a[i];
    ynk[i] = yn[i] + (9.0 * k2[i] - 3.0 * k3[i] - 6.0 * k4[i] + 4.0 * k5[i]) / 8.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt / 2.0, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    yn1[i] -= 32.0 * dt * a[i] / 120.0;
    k5[i] = dt * a[i];
    ynk[i] = yn[i] + (9.0 * k1[i] - 36.0 * k2[i] + 63.0 * k3[i] + 72.0 * k4[i] - 64.0 * k5[i]) / 44.0;
  }
  gh_rk_deriv<F, V>(T, tn + dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) yn1[i] += 11.0 * dt * a[i] / 120.0;
}

template <int ORDER, class F, class V>
__device__ __forceinline__ void gh_rk_onestep(const F& T, const double* yn, double* yn1, double tn,
                                              double dt) {
  if constexpr (ORDER == 4)
    gh_rk4_onestep<F, V>(T, yn, yn1, tn, dt);
  else
    gh_rk6_onestep<F, V>(T, yn, yn1, tn, dt);
}

// rk4_estimate_step / rk6_estimate_step, bovy_rk.c:328-500: one step of dt
// against two of dt/2, scaled by the state's largest log|y|; dt is divided by
// ceil(err^(1/5)) (RK4) or ceil(err^(1/7)) (RK6) while that exceeds 1 and the
// total reduction stays below 10000.  (The reduction is not a power of two,
// so the reference's ndt = (long)(init_dt / dt) can round down and leave part
// of each output interval unintegrated; the same divisions reproduce it.)
template <int ORDER, class F, class V>
__device__ __forceinline__ double gh_rk_estimate_step(const F& T, const double* yo, double dt,
                                                      double to, double rtol, double atol) {
  constexpr int N = GhPhase<V>::N;
  double err = 2.0;
  const double init_dt = dt;
  double max_val = log(fabs(yo[0]));
#pragma unroll
  for (int i = 1; i < N; i++) max_val = fmax(max_val, log(fabs(yo[i])));
  const double c = fmax(atol, rtol + max_val);
  const double s = log(exp(atol - c) + exp(rtol + max_val - c)) + c;
  const double expo = 1.0 / (ORDER == 4 ? 5.0 : 7.0);
  while (err > 1.0) {
    double y1[N], y21[N], y2[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
      y1[i] = yo[i];
      y21[i] = yo[i];
    }
    gh_rk_onestep<ORDER, F, V>(T, yo, y1, to, dt);
    gh_rk_onestep<ORDER, F, V>(T, yo, y21, to, dt / 2.0);
#pragma unroll
    for (int i = 0; i < N; i++) y2[i] = y21[i];
    gh_rk_onestep<ORDER, F, V>(T, y21, y2, to + dt / 2.0, dt / 2.0);
    err = 0.0;
#pragma unroll
    for (int i = 0; i < N; i++) err += exp(2.0 * log(fabs(y1[i] - y2[i])) - 2.0 * s);
    err = sqrt(err / (double)N);
    const double f = ceil(pow(err, expo));
    if (f > 1.0 && init_dt / dt * f < GH_MAX_DT_REDUCE)
      dt /= f;
    else
      break;
  }
  return dt;
}

// bovy_dopr54_actualstep, bovy_rk.c:617-760: one Dormand-Prince 5(4) step of
// dt from (to, yo) with the first-same-as-last slope a1; the correction
// 0.85 err^-0.2 rounded to a power of two within [2^-3, 2^3]; the step is
// taken (yo, a1 and to advanced) if that power is >= 0 or `accept` is
// forced.  Returns the next step dt 2^power.
template <class F, class V>
__device__ __forceinline__ double gh_dopr54_actualstep(const F& T, double* yo, double dt, double& to,
                                                       double rtol, double atol, double* a1,
                                                       bool accept, bool& accepted) {
  constexpr int N = GhPhase<V>::N;
  constexpr double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8. / 9.;
  constexpr double a21 = 0.2, a31 = 3. / 40., a41 = 44. / 45., a51 = 19372. / 6561,
                   a61 = 9017. / 3168., a71 = 35. / 384., a32 = 9. / 40., a42 = -56. / 15.,
                   a52 = -25360. / 2187., a62 = -355. / 33., a43 = 32. / 9., a53 = 64448. / 6561.,
                   a63 = 46732. / 5247., a73 = 500. / 1113., a54 = -212. / 729., a64 = 49. / 176.,
                   a74 = 125. / 192., a65 = -5103. / 18656., a75 = -2187. / 6784., a76 = 11. / 84.;
  constexpr double b1 = 35. / 384., b3 = 500. / 1113., b4 = 125. / 192., b5 = -2187. / 6784.,
                   b6 = 11. / 84.;
  constexpr double be1 = b1 - 5179. / 57600., be3 = b3 - 7571. / 16695., be4 = b4 - 393. / 640.,
                   be5 = b5 + 92097. / 339200., be6 = b6 - 187. / 2100., be7 = -1. / 40.;
  double yn1[N], yerr[N], ynk[N], a[N], k1[N], k2[N], k3[N], k4[N], k5[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    k1[i] = dt * a1[i];
    yn1[i] = yo[i] + b1 * k1[i];
    yerr[i] = be1 * k1[i];
    ynk[i] = yo[i] + a21 * k1[i];
  }
  gh_rk_deriv<F, V>(T, to + c2 * dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    k2[i] = dt * a[i];
    ynk[i] = yo[i] + a31 * k1[i] + a32 * k2[i];
  }
  gh_rk_deriv<F, V>(T, to + c3 * dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    k3[i] = dt * a[i];
    yn1[i] += b3 * k3[i];
    yerr[i] += be3 * k3[i];
    ynk[i] = yo[i] + a41 * k1[i] + a42 * k2[i] + a43 * k3[i];
  }
  gh_rk_deriv<F, V>(T, to + c4 * dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    k4[i] = dt * a[i];
    yn1[i] += b4 * k4[i];
    yerr[i] += be4 * k4[i];
    ynk[i] = yo[i] + a51 * k1[i] + a52 * k2[i] + a53 * k3[i] + a54 * k4[i];
  }
  gh_rk_deriv<F, V>(T, to + c5 * dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    k5[i] = dt * a[i];
    yn1[i] += b5 * k5[i];
    yerr[i] += be5 * k5[i];
    ynk[i] = yo[i] + a61 * k1[i] + a62 * k2[i] + a63 * k3[i] + a64 * k4[i] + a65 * k5[i];
  }
  gh_rk_deriv<F, V>(T, to + dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double k6 = dt * a[i];
    yn1[i] += b6 * k6;
    yerr[i] += be6 * k6;
    ynk[i] = yo[i] + a71 * k1[i] + a73 * k3[i] + a74 * k4[i] + a75 * k5[i] + a76 * k6;
  }
  gh_rk_deriv<F, V>(T, to + dt, ynk, a);
#pragma unroll
  for (int i = 0; i < N; i++) yerr[i] += be7 * dt * a[i];
  double max_val = log(fabs(yo[0]));
#pragma unroll
  for (int i = 1; i < N; i++) max_val = fmax(max_val, log(fabs(yo[i])));
  const double c = fmax(atol, rtol + max_val);
  const double s = log(exp(atol - c) + exp(rtol + max_val - c)) + c;
  double err = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) err += exp(2.0 * log(fabs(yerr[i])) - 2.0 * s);
  err = sqrt(err / (double)N);
  const double corr = 0.85 * pow(err, -.2);
  double powertwo = round(log(corr) / log(2.0));
  powertwo = fmin(fmax(powertwo, -3.0), 3.0);
  accepted = (powertwo >= 0.0) || accept;
  if (accepted) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      a1[i] = a[i];
      yo[i] = yn1[i];
    }
    to += dt;
  }
  return dt * pow(2.0, powertwo);
}
