// CPU oracle: integrators -- TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Restates galpy's leapfrog / symplec4 (+ step estimators) from
// galpy/util/bovy_symplecticode.c and DOP853 from galpy/util/leung_dop853.c,
// keeping the reference's arithmetic sequence so results agree to rounding.
// Templated on the scalar type R (double, or the op-counting CountedD) and on
// the spatial dimension D: 3 for integrateFullOrbit, 2 for
// integratePlanarOrbit (the reference passes dim = 3 / 2 to the symplectic
// steppers and 6 / 4 to dop853, integratePlanarOrbit.c:788-825).
#pragma once
#include "oracle.hpp"

namespace orc {

constexpr double kDtSentinel = -9999.99;      // integrateFullOrbit.py:666-667
constexpr double kMaxDtReduce = 10000.;       // bovy_symplecticode.c:35

// symplec4 coefficients, bovy_symplecticode.c:96-99 (Kinoshita et al.)
constexpr double S4C1 = 0.6756035959798289, S4C2 = -0.1756035959798288,
                 S4C3 = -0.1756035959798288, S4C4 = 0.6756035959798289;
constexpr double S4D1 = 1.3512071919596578, S4D2 = -1.7024143839193153,
                 S4D3 = 1.3512071919596578;

struct IntStats {
  long long steps;     // orbit-steps: ndt*(nt-1) (symplectic) / accepted (DOP853)
  long long rejected;  // DOP853 rejected steps
  long long rhs;       // force / RHS evaluations
  double dt;           // step actually used (symplectic)
};

// evalRectForce / evalRectDeriv (full orbits, D = 3) or evalPlanarRectForce /
// evalPlanarRectDeriv (planar orbits, D = 2).  D = 6 and D = 4 are the
// variational systems of integrate_dxdv (orbit + one deviation, 12 and 8
// first-order components), whose right-hand sides are evalRectDeriv_dxdv and
// evalPlanarRectDeriv_dxdv; only DOP853 integrates those through deriv().
template <class R, int D = 3> struct ForceCtx {
  int npot;
  const Component* P;
  long long* nrhs;
  void force(R t, const R* q, R* a) const {
    (*nrhs)++;
    if constexpr (D == 2) planar_rect_force(npot, P, t, q, a);
    else rect_force(npot, P, t, q, a);
  }
  void deriv(R t, const R* q, R* a) const {
    (*nrhs)++;
    if constexpr (D == 2) planar_rect_deriv(npot, P, t, q, a);
    else if constexpr (D == 6) rect_deriv_dxdv(npot, P, t, q, a);
    else if constexpr (D == 4) planar_rect_deriv_dxdv(npot, P, t, q, a);
    else rect_deriv(npot, P, t, q, a);
  }
};

// bovy_symplecticode.c:63-72
template <int D, class R> inline void leapq(const R* q, const R* p, R dt, R* qn) {
  for (int i = 0; i < D; i++) qn[i] = q[i] + dt * p[i];
}
template <int D, class R> inline void leapp(const R* p, R dt, const R* a, R* pn) {
  for (int i = 0; i < D; i++) pn[i] = p[i] + dt * a[i];
}

// Shared scale set-up of the step estimators, bovy_symplecticode.c:606-623
template <int D, class R> void estimate_scales(const R* qo, const R* po, double rtol, double atol,
                                               R& sq, R& sp) {
  R mq = log(fabs(qo[0]));
  for (int i = 1; i < D; i++) if (log(fabs(qo[i])) > mq) mq = log(fabs(qo[i]));
  R mp = log(fabs(po[0]));
  for (int i = 1; i < D; i++) if (log(fabs(po[i])) > mp) mp = log(fabs(po[i]));
  R c = rmax(R(atol), R(rtol) + mq);
  sq = log(exp(R(atol) - c) + exp(R(rtol) + mq - c)) + c;
  c = rmax(R(atol), R(rtol) + mp);
  sp = log(exp(R(atol) - c) + exp(R(rtol) + mp - c)) + c;
}

// error norm, bovy_symplecticode.c:641-646 / 795-800
template <int D, class R> R estimate_err(const R* q11, const R* q12, const R* p11, const R* p12, R sq, R sp) {
  R err = 0.;
  for (int i = 0; i < D; i++) {
    err += exp(R(2.) * log(fabs(q11[i] - q12[i])) - R(2.) * sq);
    err += exp(R(2.) * log(fabs(p11[i] - p12[i])) - R(2.) * sp);
  }
  return sqrt(err / R(2.) / R((double)D));
}

// leapfrog_estimate_step, bovy_symplecticode.c:588-663
template <int D, class R>
R leapfrog_estimate_step(const ForceCtx<R, D>& F, const R* qo, const R* po, R dt, R to,
                         double rtol, double atol) {
  R err = 2.;
  R init_dt = dt;
  R q11[3], q12[3], p11[3], p12[3], qtmp[3], ptmp[3], a[3];
  R sq, sp;
  estimate_scales<D>(qo, po, rtol, atol, sq, sp);
  dt *= 2.;
  while (err > R(1.) && init_dt / dt < R(kMaxDtReduce)) {
    dt /= 2.;
    leapq<D>(qo, po, dt / R(2.), q12);
    F.force(to + dt / R(2.), q12, a);
    leapp<D>(po, dt, a, p11);
    leapq<D>(q12, p11, dt / R(2.), q11);
    leapq<D>(qo, po, dt / R(4.), q12);
    F.force(to + dt / R(4.), q12, a);
    leapp<D>(po, dt / R(2.), a, ptmp);
    leapq<D>(q12, ptmp, dt / R(2.), qtmp);
    F.force(to + R(3.) * dt / R(4.), qtmp, a);
    leapp<D>(ptmp, dt / R(2.), a, p12);
    leapq<D>(qtmp, p12, dt / R(4.), q12);
    err = estimate_err<D>(q11, q12, p11, p12, sq, sp);
  }
  return dt;
}

// symplec4_estimate_step, bovy_symplecticode.c:665-808
template <int D, class R>
R symplec4_estimate_step(const ForceCtx<R, D>& F, const R* qo, const R* po, R dt, R to,
                         double rtol, double atol) {
  const R c1 = S4C1, c2 = S4C2, c3 = S4C3, c4 = S4C4, d1 = S4D1, d2 = S4D2, d3 = S4D3;
  R err = 2.;
  R init_dt = dt;
  R q11[3], q12[3], p11[3], p12[3], qtmp[3], ptmp[3], a[3];
  R sq, sp;
  estimate_scales<D>(qo, po, rtol, atol, sq, sp);
  dt *= 2.;
  while (err > R(1.) && init_dt / dt < R(kMaxDtReduce)) {
    dt /= 2.;
    // one step of dt
    leapq<D>(qo, po, c1 * dt, q12);
    to += c1 * dt;
    F.force(to, q12, a);
    leapp<D>(po, d1 * dt, a, p12);
    leapq<D>(q12, p12, c2 * dt, qtmp);
    to += c2 * dt;
    F.force(to, qtmp, a);
    leapp<D>(p12, d2 * dt, a, ptmp);
    leapq<D>(qtmp, ptmp, c3 * dt, q12);
    to += c3 * dt;
    F.force(to, q12, a);
    leapp<D>(ptmp, d3 * dt, a, p11);
    leapq<D>(q12, p11, c4 * dt, q11);
    to += c4 * dt;
    to -= dt;
    // two steps of dt/2
    leapq<D>(qo, po, c1 * dt / R(2.), q12);
    to += c1 * dt / R(2.);
    F.force(to, q12, a);
    leapp<D>(po, d1 * dt / R(2.), a, p12);
    leapq<D>(q12, p12, c2 * dt / R(2.), qtmp);
    to += c2 * dt / R(2.);
    F.force(to, qtmp, a);
    leapp<D>(p12, d2 * dt / R(2.), a, ptmp);
    leapq<D>(qtmp, ptmp, c3 * dt / R(2.), q12);
    to += c3 * dt / R(2.);
    F.force(to, q12, a);
    leapp<D>(ptmp, d3 * dt / R(2.), a, p12);
    leapq<D>(q12, p12, (c1 + c4) * dt / R(2.), qtmp);
    to += (c1 + c4) * dt / R(2.);
    F.force(to, qtmp, a);
    leapp<D>(p12, d1 * dt / R(2.), a, ptmp);
    leapq<D>(qtmp, ptmp, c2 * dt / R(2.), q12);
    to += c2 * dt / R(2.);
    F.force(to, q12, a);
    leapp<D>(ptmp, d2 * dt / R(2.), a, p12);
    leapq<D>(q12, p12, c3 * dt / R(2.), qtmp);
    to += c3 * dt / R(2.);
    F.force(to, qtmp, a);
    leapp<D>(p12, d3 * dt / R(2.), a, ptmp);
    leapq<D>(qtmp, ptmp, c4 * dt / R(2.), q12);
    to += c4 * dt / R(2.);
    for (int i = 0; i < D; i++) p12[i] = ptmp[i];
    err = estimate_err<D>(q11, q12, p11, p12, sq, sp);
    to -= dt;
  }
  return dt;
}

// symplec6 coefficients (Yoshida 1990), bovy_symplecticode.c:100-107
constexpr double S6C[8] = {0.392256805238780, 0.510043411918458, -0.471053385409758,
                           0.687531682525198e-1, 0.687531682525198e-1, -0.471053385409758,
                           0.510043411918458, 0.392256805238780};
constexpr double S6D[7] = {0.784513610477560, 0.235573213359357, -0.117767998417887e1,
                           0.131518632068391e1, -0.117767998417887e1, 0.235573213359357,
                           0.784513610477560};

// symplec6_estimate_step, bovy_symplecticode.c:810-1000: the symplec4
// estimator with the seven-kick splitting -- one step of dt against two of
// dt/2 (the middle drifts merged as (c1 + c8) dt/2), the same scales and
// norm, halving from 2 init_dt.
template <int D, class R>
R symplec6_estimate_step(const ForceCtx<R, D>& F, const R* qo, const R* po, R dt, R to,
                         double rtol, double atol) {
  R err = 2.;
  R init_dt = dt;
  R q[3], p[3], q11[3], p11[3], a[3];
  R sq, sp;
  estimate_scales<D>(qo, po, rtol, atol, sq, sp);
  dt *= 2.;
  while (err > R(1.) && init_dt / dt < R(kMaxDtReduce)) {
    dt /= 2.;
    // one step of dt
    leapq<D>(qo, po, R(S6C[0]) * dt, q);
    to += R(S6C[0]) * dt;
    for (int i = 0; i < D; i++) p[i] = po[i];
    for (int k = 0; k < 7; k++) {
      F.force(to, q, a);
      leapp<D>(p, R(S6D[k]) * dt, a, p);
      leapq<D>(q, p, R(S6C[k + 1]) * dt, q);
      to += R(S6C[k + 1]) * dt;
    }
    for (int i = 0; i < D; i++) {
      q11[i] = q[i];
      p11[i] = p[i];
    }
    to -= dt;
    // two steps of dt/2
    leapq<D>(qo, po, R(S6C[0]) * dt / R(2.), q);
    to += R(S6C[0]) * dt / R(2.);
    for (int i = 0; i < D; i++) p[i] = po[i];
    for (int half = 0; half < 2; half++) {
      for (int k = 0; k < 7; k++) {
        F.force(to, q, a);
        leapp<D>(p, R(S6D[k]) * dt / R(2.), a, p);
        const R cn = (k == 6 && half == 0) ? R(S6C[0] + S6C[7]) : R(S6C[k + 1]);
        leapq<D>(q, p, cn * dt / R(2.), q);
        to += cn * dt / R(2.);
      }
    }
    err = estimate_err<D>(q11, q, p11, p, sq, sp);
    to -= dt;
  }
  return dt;
}

// save_qp, bovy_symplecticode.c:74-78
template <int D, class R> inline void save_qp(const R* q, const R* p, double* out) {
  for (int i = 0; i < D; i++) { out[i] = dval(q[i]); out[D + i] = dval(p[i]); }
}

// leapfrog, bovy_symplecticode.c:132-224.  yo = [q, p] (2D values);
// result = nt rows of 2D; returns the step count in st.
template <int D, class R>
void leapfrog(const ForceCtx<R, D>& F, const double* yo, int nt, double dt_in, const double* t,
              double rtol, double atol, double* result, int* err, IntStats* st) {
  R qo[3], po[3], q12[3], p12[3], a[3];
  for (int i = 0; i < D; i++) { qo[i] = yo[i]; po[i] = yo[D + i]; }
  save_qp<D>(qo, po, result);
  result += 2 * D;
  *err = 0;
  R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in;
  if (dt_in == kDtSentinel) dt = leapfrog_estimate_step<D>(F, qo, po, init_dt, R(t[0]), rtol, atol);
  long ndt = (long)(dval(init_dt / dt));
  R to = t[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    leapq<D>(qo, po, dt / R(2.), q12);
    for (long jj = 0; jj < ndt - 1; jj++) {
      F.force(to + dt / R(2.), q12, a);
      leapp<D>(po, dt, a, p12);
      leapq<D>(q12, p12, dt, qo);
      to = to + dt;
      for (int k = 0; k < D; k++) { q12[k] = qo[k]; po[k] = p12[k]; }
    }
    F.force(to + dt / R(2.), q12, a);
    leapp<D>(po, dt, a, po);
    leapq<D>(q12, po, dt / R(2.), qo);
    to = to + dt;
    save_qp<D>(qo, po, result);
    result += 2 * D;
  }
  if (st) { st->steps = (long long)ndt * (nt - 1); st->dt = dval(dt); }
}

// symplec4, bovy_symplecticode.c:250-377
template <int D, class R>
void symplec4(const ForceCtx<R, D>& F, const double* yo, int nt, double dt_in, const double* t,
              double rtol, double atol, double* result, int* err, IntStats* st) {
  const R c1 = S4C1, c2 = S4C2, c3 = S4C3, c4 = S4C4, d1 = S4D1, d2 = S4D2, d3 = S4D3;
  R qo[3], po[3], q12[3], p12[3], a[3];
  for (int i = 0; i < D; i++) { qo[i] = yo[i]; po[i] = yo[D + i]; }
  save_qp<D>(qo, po, result);
  result += 2 * D;
  *err = 0;
  R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in;
  if (dt_in == kDtSentinel) dt = symplec4_estimate_step<D>(F, qo, po, init_dt, R(t[0]), rtol, atol);
  long ndt = (long)(dval(init_dt / dt));
  R to = t[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    leapq<D>(qo, po, c1 * dt, q12);
    to += c1 * dt;
    for (long jj = 0; jj < ndt - 1; jj++) {
      F.force(to, q12, a);
      leapp<D>(po, d1 * dt, a, p12);
      leapq<D>(q12, p12, c2 * dt, qo);
      to += c2 * dt;
      F.force(to, qo, a);
      leapp<D>(p12, d2 * dt, a, po);
      leapq<D>(qo, po, c3 * dt, q12);
      to += c3 * dt;
      F.force(to, q12, a);
      leapp<D>(po, d3 * dt, a, p12);
      leapq<D>(q12, p12, (c4 + c1) * dt, qo);
      to += (c4 + c1) * dt;
      for (int k = 0; k < D; k++) { q12[k] = qo[k]; po[k] = p12[k]; }
    }
    F.force(to, q12, a);
    leapp<D>(po, d1 * dt, a, p12);
    leapq<D>(q12, p12, c2 * dt, qo);
    to += c2 * dt;
    F.force(to, qo, a);
    leapp<D>(p12, d2 * dt, a, po);
    leapq<D>(qo, po, c3 * dt, q12);
    to += c3 * dt;
    F.force(to, q12, a);
    leapp<D>(po, d3 * dt, a, p12);
    leapq<D>(q12, p12, c4 * dt, qo);
    to += c4 * dt;
    for (int k = 0; k < D; k++) po[k] = p12[k];
    save_qp<D>(qo, po, result);
    result += 2 * D;
  }
  if (st) { st->steps = (long long)ndt * (nt - 1); st->dt = dval(dt); }
}

// symplec6, bovy_symplecticode.c:403-586: the symplec4 loop with seven
// kicks (the last drift of a step and the first of the next merged as
// (c8 + c1) dt).
template <int D, class R>
void symplec6(const ForceCtx<R, D>& F, const double* yo, int nt, double dt_in, const double* t,
              double rtol, double atol, double* result, int* err, IntStats* st) {
  R q[3], p[3], a[3];
  for (int i = 0; i < D; i++) { q[i] = yo[i]; p[i] = yo[D + i]; }
  save_qp<D>(q, p, result);
  result += 2 * D;
  *err = 0;
  R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in;
  if (dt_in == kDtSentinel) dt = symplec6_estimate_step<D>(F, q, p, init_dt, R(t[0]), rtol, atol);
  long ndt = (long)(dval(init_dt / dt));
  R to = t[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    leapq<D>(q, p, R(S6C[0]) * dt, q);
    to += R(S6C[0]) * dt;
    for (long jj = 0; jj < ndt; jj++) {
      const bool last = jj == ndt - 1;
      for (int k = 0; k < 7; k++) {
        F.force(to, q, a);
        leapp<D>(p, R(S6D[k]) * dt, a, p);
        const R cn = (k == 6 && !last) ? R(S6C[7] + S6C[0]) : R(S6C[k + 1]);
        leapq<D>(q, p, cn * dt, q);
        to += cn * dt;
      }
    }
    save_qp<D>(q, p, result);
    result += 2 * D;
  }
  if (st) { st->steps = (long long)ndt * (nt - 1); st->dt = dval(dt); }
}

// ---------------------------------------------------------------------------
// Runge-Kutta steppers of bovy_rk.c on the rectangular state y = [q, p]
// (2D values) with the derivative evalRectDeriv (integrateFullOrbit.c:1091-1105).
// ---------------------------------------------------------------------------
// bovy_rk4_onestep, bovy_rk.c:142-169: yn1 (equal to yn on entry) gains the
// weighted slopes; ynk holds the stage states.
template <int D, class R>
void rk4_onestep(const ForceCtx<R, D>& F, const R* yn, R* yn1, R tn, R dt) {
  constexpr int N = 2 * D;
  R ynk[N], a[N];
  F.deriv(tn, yn, a);
  for (int i = 0; i < N; i++) yn1[i] += dt * a[i] / R(6.);
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + dt * a[i] / R(2.);
  F.deriv(tn + dt / R(2.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += dt * a[i] / R(3.);
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + dt * a[i] / R(2.);
  F.deriv(tn + dt / R(2.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += dt * a[i] / R(3.);
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + dt * a[i];
  F.deriv(tn + dt, ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += dt * a[i] / R(6.);
}

// bovy_rk6_onestep, bovy_rk.c:270-326 (seven evaluations)
template <int D, class R>
void rk6_onestep(const ForceCtx<R, D>& F, const R* yn, R* yn1, R tn, R dt) {
  constexpr int N = 2 * D;
  R ynk[N], a[N], k1[N], k2[N], k3[N], k4[N], k5[N];
  F.deriv(tn, yn, a);
  for (int i = 0; i < N; i++) yn1[i] += R(11.) * dt * a[i] / R(120.);
  for (int i = 0; i < N; i++) k1[i] = dt * a[i];
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + k1[i] / R(3.);
  F.deriv(tn + dt / R(3.), ynk, a);
  for (int i = 0; i < N; i++) k2[i] = dt * a[i];
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + R(2.) * k2[i] / R(3.);
  F.deriv(tn + R(2.) * dt / R(3.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += R(81.) * dt * a[i] / R(120.);
  for (int i = 0; i < N; i++) k3[i] = dt * a[i];
  for (int i = 0; i < N; i++) ynk[i] = yn[i] + (k1[i] + R(4.) * k2[i] - k3[i]) / R(12.);
  F.deriv(tn + dt / R(3.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += R(81.) * dt * a[i] / R(120.);
  for (int i = 0; i < N; i++) k4[i] = dt * a[i];
  for (int i = 0; i < N; i++)
    ynk[i] = yn[i] + (-k1[i] + R(18.) * k2[i] - R(3.) * k3[i] - R(6.) * k4[i]) / R(16.);
  F.deriv(tn + dt / R(2.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] -= R(32.) * dt * a[i] / R(120.);
  for (int i = 0; i < N; i++) k5[i] = dt * a[i];
  for (int i = 0; i < N; i++)
    ynk[i] = yn[i] + (R(9.) * k2[i] - R(3.) * k3[i] - R(6.) * k4[i] + R(4.) * k5[i]) / R(8.);
  F.deriv(tn + dt / R(2.), ynk, a);
  for (int i = 0; i < N; i++) yn1[i] -= R(32.) * dt * a[i] / R(120.);
  for (int i = 0; i < N; i++) k5[i] = dt * a[i];  // k6 in k5's place, as the reference
  for (int i = 0; i < N; i++)
    ynk[i] = yn[i] + (R(9.) * k1[i] - R(36.) * k2[i] + R(63.) * k3[i] + R(72.) * k4[i] -
                      R(64.) * k5[i]) / R(44.);
  F.deriv(tn + dt, ynk, a);
  for (int i = 0; i < N; i++) yn1[i] += R(11.) * dt * a[i] / R(120.);
}

// rk4_estimate_step / rk6_estimate_step, bovy_rk.c:328-500: one step of dt
// against two of dt/2 in the scaled norm of the state's largest log|y|;
// dt is divided by ceil(err^(1/5)) (RK4) or ceil(err^(1/7)) (RK6) until that
// is 1 or the reduction would pass 10000.
template <int ORDER, int D, class R>
R rk_estimate_step(const ForceCtx<R, D>& F, const R* yo, R dt, R to, double rtol, double atol) {
  constexpr int N = 2 * D;
  R err = 2.;
  R init_dt = dt;
  R max_val = log(fabs(yo[0]));
  for (int i = 1; i < N; i++)
    if (log(fabs(yo[i])) > max_val) max_val = log(fabs(yo[i]));
  R c = rmax(R(atol), R(rtol) + max_val);
  R s = log(exp(R(atol) - c) + exp(R(rtol) + max_val - c)) + c;
  const double expo = 1. / (ORDER == 4 ? 5. : 7.);
  while (err > R(1.)) {
    R y1[N], y21[N], y2[N];
    for (int i = 0; i < N; i++) { y1[i] = yo[i]; y21[i] = yo[i]; }
    if (ORDER == 4) {
      rk4_onestep<D>(F, yo, y1, to, dt);
      rk4_onestep<D>(F, yo, y21, to, dt / R(2.));
      for (int i = 0; i < N; i++) y2[i] = y21[i];
      rk4_onestep<D>(F, y21, y2, to + dt / R(2.), dt / R(2.));
    } else {
      rk6_onestep<D>(F, yo, y1, to, dt);
      rk6_onestep<D>(F, yo, y21, to, dt / R(2.));
      for (int i = 0; i < N; i++) y2[i] = y21[i];
      rk6_onestep<D>(F, y21, y2, to + dt / R(2.), dt / R(2.));
    }
    err = 0.;
    for (int i = 0; i < N; i++) err += exp(R(2.) * log(fabs(y1[i] - y2[i])) - R(2.) * s);
    err = sqrt(err / R(N));
    const double f = std::ceil(std::pow(dval(err), expo));
    if (f > 1. && dval(init_dt / dt) * f < kMaxDtReduce)
      dt /= R(f);
    else
      break;
  }
  return dt;
}

// bovy_rk4 / bovy_rk6, bovy_rk.c:66-140, 171-268: ndt = (long)(init_dt/dt)
// steps per output interval.
template <int ORDER, int D, class R>
void rk_fixed(const ForceCtx<R, D>& F, const double* yo, int nt, double dt_in, const double* t,
              double rtol, double atol, double* result, int* err, IntStats* st) {
  constexpr int N = 2 * D;
  R yn[N], yn1[N];
  for (int i = 0; i < N; i++) { yn[i] = yo[i]; yn1[i] = yo[i]; result[i] = yo[i]; }
  result += N;
  *err = 0;
  R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in;
  if (dt_in == kDtSentinel) dt = rk_estimate_step<ORDER, D>(F, yn, init_dt, R(t[0]), rtol, atol);
  long ndt = (long)(dval(init_dt / dt));
  R to = t[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    for (long jj = 0; jj < ndt; jj++) {
      if (ORDER == 4) rk4_onestep<D>(F, yn, yn1, to, dt);
      else rk6_onestep<D>(F, yn, yn1, to, dt);
      to += dt;
      for (int k = 0; k < N; k++) yn[k] = yn1[k];
    }
    for (int k = 0; k < N; k++) result[k] = dval(yn1[k]);
    result += N;
  }
  if (st) { st->steps = (long long)ndt * (nt - 1); st->dt = dval(dt); }
}

// bovy_dopr54_actualstep, bovy_rk.c:617-760: one Dormand-Prince 5(4) step
// of dt from (to, yo) with the first-same-as-last slope a1; the correction
// 0.85 err^-0.2 rounded to a power of two in [2^-3, 2^3]; accepted (yo, a1,
// to advanced) if that power is >= 0 or `accept` is forced.  Returns the next
// step, dt 2^power.
template <int D, class R>
R dopr54_actualstep(const ForceCtx<R, D>& F, R* yo, R dt, R& to, double rtol, double atol, R* a1,
                    bool accept, bool& accepted) {
  constexpr int N = 2 * D;
  const double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8. / 9.;
  const double a21 = 0.2, a31 = 3. / 40., a41 = 44. / 45., a51 = 19372. / 6561, a61 = 9017. / 3168.,
               a71 = 35. / 384., a32 = 9. / 40., a42 = -56. / 15., a52 = -25360. / 2187.,
               a62 = -355. / 33., a43 = 32. / 9., a53 = 64448. / 6561., a63 = 46732. / 5247.,
               a73 = 500. / 1113., a54 = -212. / 729., a64 = 49. / 176., a74 = 125. / 192.,
               a65 = -5103. / 18656., a75 = -2187. / 6784., a76 = 11. / 84.;
  const double b1 = 35. / 384., b3 = 500. / 1113., b4 = 125. / 192., b5 = -2187. / 6784.,
               b6 = 11. / 84.;
  const double be1 = b1 - 5179. / 57600., be3 = b3 - 7571. / 16695., be4 = b4 - 393. / 640.,
               be5 = b5 + 92097. / 339200., be6 = b6 - 187. / 2100., be7 = -1. / 40.;
  R yn1[N], yerr[N], ynk[N], a[N], k1[N], k2[N], k3[N], k4[N], k5[N], k6[N];
  for (int i = 0; i < N; i++) yn1[i] = yo[i];
  for (int i = 0; i < N; i++) a[i] = a1[i];
  for (int i = 0; i < N; i++) {
    k1[i] = dt * a[i];
    yn1[i] += R(b1) * k1[i];
    yerr[i] = R(be1) * k1[i];
    ynk[i] = yo[i] + R(a21) * k1[i];
  }
  F.deriv(to + R(c2) * dt, ynk, a);
  for (int i = 0; i < N; i++) {
    k2[i] = dt * a[i];
    ynk[i] = yo[i] + R(a31) * k1[i] + R(a32) * k2[i];
  }
  F.deriv(to + R(c3) * dt, ynk, a);
  for (int i = 0; i < N; i++) {
    k3[i] = dt * a[i];
    yn1[i] += R(b3) * k3[i];
    yerr[i] += R(be3) * k3[i];
    ynk[i] = yo[i] + R(a41) * k1[i] + R(a42) * k2[i] + R(a43) * k3[i];
  }
  F.deriv(to + R(c4) * dt, ynk, a);
  for (int i = 0; i < N; i++) {
    k4[i] = dt * a[i];
    yn1[i] += R(b4) * k4[i];
    yerr[i] += R(be4) * k4[i];
    ynk[i] = yo[i] + R(a51) * k1[i] + R(a52) * k2[i] + R(a53) * k3[i] + R(a54) * k4[i];
  }
  F.deriv(to + R(c5) * dt, ynk, a);
  for (int i = 0; i < N; i++) {
    k5[i] = dt * a[i];
    yn1[i] += R(b5) * k5[i];
    yerr[i] += R(be5) * k5[i];
    ynk[i] = yo[i] + R(a61) * k1[i] + R(a62) * k2[i] + R(a63) * k3[i] + R(a64) * k4[i] +
             R(a65) * k5[i];
  }
  F.deriv(to + dt, ynk, a);
  for (int i = 0; i < N; i++) {
    k6[i] = dt * a[i];
    yn1[i] += R(b6) * k6[i];
    yerr[i] += R(be6) * k6[i];
    ynk[i] = yo[i] + R(a71) * k1[i] + R(a73) * k3[i] + R(a74) * k4[i] + R(a75) * k5[i] +
             R(a76) * k6[i];
  }
  F.deriv(to + dt, ynk, a);
  for (int i = 0; i < N; i++) yerr[i] += R(be7) * dt * a[i];
  R max_val = log(fabs(yo[0]));
  for (int i = 1; i < N; i++)
    if (log(fabs(yo[i])) > max_val) max_val = log(fabs(yo[i]));
  R c = rmax(R(atol), R(rtol) + max_val);
  R s = log(exp(R(atol) - c) + exp(R(rtol) + max_val - c)) + c;
  R err = 0.;
  for (int i = 0; i < N; i++) err += exp(R(2.) * log(fabs(yerr[i])) - R(2.) * s);
  err = sqrt(err / R(N));
  const double corr = 0.85 * std::pow(dval(err), -.2);
  double powertwo = std::round(std::log(corr) / std::log(2.));
  if (powertwo > 3.) powertwo = 3.;
  else if (powertwo < -3.) powertwo = -3.;
  accepted = (powertwo >= 0.) || accept;
  if (accepted) {
    for (int i = 0; i < N; i++) {
      a1[i] = a[i];
      yo[i] = yn1[i];
    }
    to += dt;
  }
  return dt * R(std::pow(2., powertwo));
}

// bovy_dopr54 / bovy_dopr54_onestep, bovy_rk.c:501-616: the first step from
// the RK4 estimator; each output interval of the grid's first spacing is
// stepped to its end (the last step clipped to land on it); a step reduced
// more than 10000-fold from the interval's first one (or NaN) is forced
// through at that floor and flags err as odd.
template <int D, class R>
void dopr54(const ForceCtx<R, D>& F, const double* yo_in, int nt, double dt_one_in,
            const double* t, double rtol, double atol, double* result, int* err, IntStats* st) {
  constexpr int N = 2 * D;
  R yn[N], a1[N];
  for (int i = 0; i < N; i++) { yn[i] = yo_in[i]; result[i] = yo_in[i]; }
  result += N;
  *err = 0;
  const R dt = R(t[1]) - R(t[0]);
  R dt_one = dt_one_in;
  if (dt_one_in == kDtSentinel) dt_one = rk_estimate_step<4, D>(F, yn, dt, R(t[0]), rtol, atol);
  R to = t[0];
  F.deriv(to, yn, a1);
  long long nacc = 0, nrej = 0;
  for (int ii = 0; ii < nt - 1; ii++) {
    const R init_dt_one = dt_one, init_to = to;
    while ((dt >= R(0.) && to < init_to + dt) || (dt < R(0.) && to > init_to + dt)) {
      bool accept = false;
      if (init_dt_one / dt_one > R(10000.) || dval(dt_one) != dval(dt_one)) {
        dt_one = init_dt_one / R(10000.);
        accept = true;
        if (*err % 2 == 0) *err += 1;
      }
      if (dt >= R(0.) && dt_one > init_to + dt - to) dt_one = init_to + dt - to;
      if (dt < R(0.) && dt_one < init_to + dt - to) dt_one = init_to + dt - to;
      bool accepted;
      dt_one = dopr54_actualstep<D>(F, yn, dt_one, to, rtol, atol, a1, accept, accepted);
      if (accepted) nacc++;
      else nrej++;
    }
    for (int k = 0; k < N; k++) result[k] = dval(yn[k]);
    result += N;
  }
  if (st) { st->steps = nacc; st->rejected = nrej; st->dt = dval(dt_one); }
}

// Variational (tangent-map) steppers, bovy_symplecticode.c:1053-1268:
// the same DKD sequence on the augmented state [q | dq_1..dq_nde],
// [p | dp_1..dp_nde]; the kick of each deviation column is h K dq_j
// (rect_force_dxdv, or planar_rect_force_dxdv for DB = 2).  The step is
// estimated from the base orbit only (:1081-1084), so the base trajectory is
// bit-identical to a plain run.  w = rectangular base state [q(DB), p(DB)]
// followed by nde deviation blocks [dq(DB), dp(DB)] (save_qp_dxdv's layout,
// :79-90); final = the final augmented state (nullable); traj (nullable) =
// all nt rows in the same layout.  t = output grid (only its spacing and
// count matter, as for the plain steppers).  Returns false if a component
// has no restated Hessian.
template <int DB, class R>
bool symplectic_dxdv_traj(int npot, const Component* P, int method, const double* w_in, int nde,
                          int nt, const double* t, double dt_in, double rtol, double atol,
                          double* final_out, double* traj) {
  const int nb = 1 + nde;
  constexpr int W = 2 * DB;  // doubles per block
  R q[3 * 7], p[3 * 7], a[3 * 7];
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < DB; k++) {
      q[DB * b + k] = w_in[W * b + k];
      p[DB * b + k] = w_in[W * b + DB + k];
    }
  auto save = [&](double* out) {
    for (int b = 0; b < nb; b++)
      for (int k = 0; k < DB; k++) {
        out[W * b + k] = dval(q[DB * b + k]);
        out[W * b + DB + k] = dval(p[DB * b + k]);
      }
  };
  if (traj) save(traj);
  long long nrhs = 0;
  ForceCtx<R, DB> F{npot, P, &nrhs};
  R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in;
  if (dt_in == kDtSentinel)
    dt = (method == 0)   ? leapfrog_estimate_step<DB>(F, q, p, init_dt, R(t[0]), rtol, atol)
         : (method == 4) ? symplec6_estimate_step<DB>(F, q, p, init_dt, R(t[0]), rtol, atol)
                         : symplec4_estimate_step<DB>(F, q, p, init_dt, R(t[0]), rtol, atol);
  long ndt = (long)(dval(init_dt / dt));
  R to = t[0];
  auto drift = [&](R h) {
    for (int i = 0; i < DB * nb; i++) q[i] = q[i] + h * p[i];
  };
  bool ok = true;
  auto kick = [&](R tk, R h) {
    if constexpr (DB == 3) ok &= rect_force_dxdv(npot, P, tk, q, a, nde);
    else ok &= planar_rect_force_dxdv(npot, P, tk, q, a, nde);
    for (int i = 0; i < DB * nb; i++) p[i] = p[i] + h * a[i];
  };
  const R c1 = S4C1, c2 = S4C2, c3 = S4C3, c4 = S4C4, d1 = S4D1, d2 = S4D2, d3 = S4D3;
  for (int ii = 0; ii < nt - 1; ii++) {
    if (method == 0) {
      drift(dt / R(2.));
      for (long jj = 0; jj < ndt - 1; jj++) {
        kick(to + dt / R(2.), dt);
        drift(dt);
        to = to + dt;
      }
      kick(to + dt / R(2.), dt);
      drift(dt / R(2.));
      to = to + dt;
    } else if (method == 4) {
      // symplec6_dxdv, bovy_symplecticode.c:1269-: symplec6's drift/kick
      // sequence on the augmented state (as symplec6 above)
      drift(R(S6C[0]) * dt);
      to += R(S6C[0]) * dt;
      for (long jj = 0; jj < ndt; jj++) {
        const bool last = jj == ndt - 1;
        for (int k = 0; k < 7; k++) {
          kick(to, R(S6D[k]) * dt);
          const R cn = (k == 6 && !last) ? R(S6C[7] + S6C[0]) : R(S6C[k + 1]);
          drift(cn * dt);
          to += cn * dt;
        }
      }
    } else {
      drift(c1 * dt);
      to += c1 * dt;
      for (long jj = 0; jj < ndt - 1; jj++) {
        kick(to, d1 * dt);
        drift(c2 * dt);
        to += c2 * dt;
        kick(to, d2 * dt);
        drift(c3 * dt);
        to += c3 * dt;
        kick(to, d3 * dt);
        drift((c4 + c1) * dt);
        to += (c4 + c1) * dt;
      }
      kick(to, d1 * dt);
      drift(c2 * dt);
      to += c2 * dt;
      kick(to, d2 * dt);
      drift(c3 * dt);
      to += c3 * dt;
      kick(to, d3 * dt);
      drift(c4 * dt);
      to += c4 * dt;
    }
    if (traj) save(traj + (long)(ii + 1) * W * nb);
  }
  if (final_out) save(final_out);
  return ok;
}

// the final augmented state of a full orbit (DB = 3)
template <class R>
bool symplectic_dxdv(int npot, const Component* P, int method, const double* w_in, int nde, int nt,
                     const double* t, double dt_in, double rtol, double atol, double* out) {
  return symplectic_dxdv_traj<3, R>(npot, P, method, w_in, nde, nt, t, dt_in, rtol, atol, out,
                                    nullptr);
}

// ---------------------------------------------------------------------------
// DOP853 tableau (Hairer, Norsett & Wanner, "Solving ODEs I", dop853.f);
// the same published constants galpy uses at leung_dop853.c:41-205.
// ---------------------------------------------------------------------------
namespace dp {
constexpr double c2 = 0.526001519587677318785587544488e-1, c3 = 0.789002279381515978178381316732e-1,
  c4 = 0.118350341907227396726757197510, c5 = 0.281649658092772603273242802490,
  c6 = 0.333333333333333333333333333333, c7 = 0.25, c8 = 0.307692307692307692307692307692,
  c9 = 0.651282051282051282051282051282, c10 = 0.6, c11 = 0.857142857142857142857142857142,
  c14 = 0.1, c15 = 0.2, c16 = 0.777777777777777777777777777778;
constexpr double a21 = 5.26001519587677318785587544488e-2,
  a31 = 1.97250569845378994544595329183e-2, a32 = 5.91751709536136983633785987549e-2,
  a41 = 2.95875854768068491816892993775e-2, a43 = 8.87627564304205475450678981324e-2,
  a51 = 2.41365134159266685502369798665e-1, a53 = -8.84549479328286085344864962717e-1,
  a54 = 9.24834003261792003115737966543e-1,
  a61 = 3.7037037037037037037037037037e-2, a64 = 1.70828608729473871279604482173e-1,
  a65 = 1.25467687566822425016691814123e-1,
  a71 = 3.7109375e-2, a74 = 1.70252211019544039314978060272e-1,
  a75 = 6.02165389804559606850219397283e-2, a76 = -1.7578125e-2,
  a81 = 3.70920001185047927108779319836e-2, a84 = 1.70383925712239993810214054705e-1,
  a85 = 1.07262030446373284651809199168e-1, a86 = -1.53194377486244017527936158236e-2,
  a87 = 8.27378916381402288758473766002e-3,
  a91 = 6.24110958716075717114429577812e-1, a94 = -3.36089262944694129406857109825e0,
  a95 = -8.68219346841726006818189891453e-1, a96 = 2.75920996994467083049415600797e1,
  a97 = 2.01540675504778934086186788979e1, a98 = -4.34898841810699588477366255144e1,
  a101 = 4.77662536438264365890433908527e-1, a104 = -2.48811461997166764192642586468e0,
  a105 = -5.90290826836842996371446475743e-1, a106 = 2.12300514481811942347288949897e1,
  a107 = 1.52792336328824235832596922938e1, a108 = -3.32882109689848629194453265587e1,
  a109 = -2.03312017085086261358222928593e-2,
  a111 = -9.3714243008598732571704021658e-1, a114 = 5.18637242884406370830023853209e0,
  a115 = 1.09143734899672957818500254654e0, a116 = -8.14978701074692612513997267357e0,
  a117 = -1.85200656599969598641566180701e1, a118 = 2.27394870993505042818970056734e1,
  a119 = 2.49360555267965238987089396762e0, a1110 = -3.0467644718982195003823669022e0,
  a121 = 2.27331014751653820792359768449e0, a124 = -1.05344954667372501984066689879e1,
  a125 = -2.00087205822486249909675718444e0, a126 = -1.79589318631187989172765950534e1,
  a127 = 2.79488845294199600508499808837e1, a128 = -2.85899827713502369474065508674e0,
  a129 = -8.87285693353062954433549289258e0, a1210 = 1.23605671757943030647266201528e1,
  a1211 = 6.43392746015763530355970484046e-1,
  a141 = 5.61675022830479523392909219681e-2, a147 = 2.53500210216624811088794765333e-1,
  a148 = -2.46239037470802489917441475441e-1, a149 = -1.24191423263816360469010140626e-1,
  a1410 = 1.5329179827876569731206322685e-1, a1411 = 8.20105229563468988491666602057e-3,
  a1412 = 7.56789766054569976138603589584e-3, a1413 = -8.298e-3,
  a151 = 3.18346481635021405060768473261e-2, a156 = 2.83009096723667755288322961402e-2,
  a157 = 5.35419883074385676223797384372e-2, a158 = -5.49237485713909884646569340306e-2,
  a1511 = -1.08347328697249322858509316994e-4, a1512 = 3.82571090835658412954920192323e-4,
  a1513 = -3.40465008687404560802977114492e-4, a1514 = 1.41312443674632500278074618366e-1,
  a161 = -4.28896301583791923408573538692e-1, a166 = -4.69762141536116384314449447206e0,
  a167 = 7.68342119606259904184240953878e0, a168 = 4.06898981839711007970213554331e0,
  a169 = 3.56727187455281109270669543021e-1, a1613 = -1.39902416515901462129418009734e-3,
  a1614 = 2.9475147891527723389556272149e0, a1615 = -9.15095847217987001081870187138e0;
constexpr double b1 = 5.42937341165687622380535766363e-2, b6 = 4.45031289275240888144113950566,
  b7 = 1.89151789931450038304281599044, b8 = -5.8012039600105847814672114227,
  b9 = 3.1116436695781989440891606237e-1, b10 = -1.52160949662516078556178806805e-1,
  b11 = 2.01365400804030348374776537501e-1, b12 = 4.47106157277725905176885569043e-2;
constexpr double bhh1 = 0.244094488188976377952755905512, bhh2 = 0.733846688281611857341361741547,
  bhh3 = 0.220588235294117647058823529412e-1;
constexpr double er1 = 0.1312004499419488073250102996e-1, er6 = -0.1225156446376204440720569753e+1,
  er7 = -0.4957589496572501915214079952, er8 = 0.1664377182454986536961530415e+1,
  er9 = -0.3503288487499736816886487290, er10 = 0.3341791187130174790297318841,
  er11 = 0.8192320648511571246570742613e-1, er12 = -0.2235530786388629525884427845e-1;
constexpr double d41 = -0.84289382761090128651353491142e+1, d46 = 0.56671495351937776962531783590,
  d47 = -0.30689499459498916912797304727e+1, d48 = 0.23846676565120698287728149680e+1,
  d49 = 0.21170345824450282767155149946e+1, d410 = -0.87139158377797299206789907490,
  d411 = 0.22404374302607882758541771650e+1, d412 = 0.63157877876946881815570249290,
  d413 = -0.88990336451333310820698117400e-1, d414 = 0.18148505520854727256656404962e+2,
  d415 = -0.91946323924783554000451984436e+1, d416 = -0.44360363875948939664310572000e+1;
constexpr double d51 = 0.10427508642579134603413151009e+2, d56 = 0.24228349177525818288430175319e+3,
  d57 = 0.16520045171727028198505394887e+3, d58 = -0.37454675472269020279518312152e+3,
  d59 = -0.22113666853125306036270938578e+2, d510 = 0.77334326684722638389603898808e+1,
  d511 = -0.30674084731089398182061213626e+2, d512 = -0.93321305264302278729567221706e+1,
  d513 = 0.15697238121770843886131091075e+2, d514 = -0.31139403219565177677282850411e+2,
  d515 = -0.93529243588444783865713862664e+1, d516 = 0.35816841486394083752465898540e+2;
constexpr double d61 = 0.19985053242002433820987653617e+2, d66 = -0.38703730874935176555105901742e+3,
  d67 = -0.18917813819516756882830838328e+3, d68 = 0.52780815920542364900561016686e+3,
  d69 = -0.11573902539959630126141871134e+2, d610 = 0.68812326946963000169666922661e+1,
  d611 = -0.10006050966910838403183860980e+1, d612 = 0.77771377980534432092869265740,
  d613 = -0.27782057523535084065932004339e+1, d614 = -0.60196695231264120758267380846e+2,
  d615 = 0.84320405506677161018159903784e+2, d616 = 0.11992291136182789328035130030e+2;
constexpr double d71 = -0.25693933462703749003312586129e+2, d76 = -0.15418974869023643374053993627e+3,
  d77 = -0.23152937917604549567536039109e+3, d78 = 0.35763911791061412378285349910e+3,
  d79 = 0.93405324183624310003907691704e+2, d710 = -0.37458323136451633156875139351e+2,
  d711 = 0.10409964950896230045147246184e+3, d712 = 0.29840293426660503123344363579e+2,
  d713 = -0.43533456590011143754432175058e+2, d714 = 0.96324553959188282948394950600e+2,
  d715 = -0.39177261675615439165231486172e+2, d716 = -0.14972683625798562581422125276e+3;
constexpr double uround = 2.3e-16;
}  // namespace dp

template <class R> inline R csign(R x, R y) { return (y > R(0.)) ? fabs(x) : -fabs(x); }

// dop853, leung_dop853.c:231-517.  y0 (2D values, rectangular) is advanced
// in place, as in the reference; result = nt rows of 2D at the requested
// times.
template <int D, class R>
void dop853(const ForceCtx<R, D>& F, double* y0d, int nt, const double* t, double rtol_log,
            double atol_log, double* result, int* err, IntStats* st) {
  using namespace dp;
  const int dim = 2 * D;
  R rtol = exp(R(rtol_log));
  R atol = exp(R(atol_log));
  const R safe = 0.9, beta = 0.0;
  R facold = 1.0e-4;
  const R expo1 = R(1.0 / 8.0) - beta * R(0.2);
  const R facc1 = R(1.0) / R(0.333), facc2 = R(1.0) / R(6.0);
  R hmax = R(t[nt - 1]) - R(t[0]);
  R pos_neg = csign(R(1.0), hmax);
  R y0[2 * D], yy1[2 * D], k1[2 * D], k2[2 * D], k3[2 * D], k4[2 * D], k5[2 * D], k6[2 * D],
      k7[2 * D], k8[2 * D], k9[2 * D], k10[2 * D];
  R rc1[2 * D], rc2[2 * D], rc3[2 * D], rc4[2 * D], rc5[2 * D], rc6[2 * D], rc7[2 * D], rc8[2 * D];
  long long nacc = 0, nrej = 0;
  for (int i = 0; i < dim; i++) { y0[i] = y0d[i]; result[i] = y0d[i]; }
  result += dim;
  *err = 0;
  F.deriv(R(t[0]), y0, k1);
  R dnf = 0.0, dny = 0.0, sk, sqr;
  for (int i = 0; i < dim; i++) {
    sk = atol + rtol * fabs(y0[i]);
    sqr = k1[i] / sk;
    dnf += sqr * sqr;
    sqr = y0[i] / sk;
    dny += sqr * sqr;
  }
  R h = csign(rmin(sqrt(dny / dnf) * R(0.01), fabs(hmax)), pos_neg);
  for (int i = 0; i < dim; i++) k3[i] = y0[i] + h * k1[i];
  F.deriv(R(t[0]) + h, k3, k2);
  R der2 = 0.0;
  for (int i = 0; i < dim; i++) {
    sk = atol + rtol * fabs(y0[i]);
    sqr = (k2[i] - k1[i]) / sk;
    der2 += sqr * sqr;
  }
  der2 = sqrt(der2) / h;
  R der12 = rmax(fabs(der2), sqrt(dnf));
  R h1 = pow(R(0.01) / der12, R(1.0 / 8.0));
  h = csign(rmin(R(100.0) * fabs(h), rmin(fabs(h1), fabs(hmax))), pos_neg);
  int reject = 0;
  R t_current = t[0], t_old = t[0], t_old_older = t[0];
  int done = 0;
  while (done < nt - 1) {
    h = pos_neg * rmax(fabs(h), R(1e3 * uround));
    for (int i = 0; i < dim; i++) yy1[i] = y0[i] + h * R(a21) * k1[i];
    F.deriv(t_current + R(c2) * h, yy1, k2);
    for (int i = 0; i < dim; i++) yy1[i] = y0[i] + h * (R(a31) * k1[i] + R(a32) * k2[i]);
    F.deriv(t_current + R(c3) * h, yy1, k3);
    for (int i = 0; i < dim; i++) yy1[i] = y0[i] + h * (R(a41) * k1[i] + R(a43) * k3[i]);
    F.deriv(t_current + R(c4) * h, yy1, k4);
    for (int i = 0; i < dim; i++) yy1[i] = y0[i] + h * (R(a51) * k1[i] + R(a53) * k3[i] + R(a54) * k4[i]);
    F.deriv(t_current + R(c5) * h, yy1, k5);
    for (int i = 0; i < dim; i++) yy1[i] = y0[i] + h * (R(a61) * k1[i] + R(a64) * k4[i] + R(a65) * k5[i]);
    F.deriv(t_current + R(c6) * h, yy1, k6);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a71) * k1[i] + R(a74) * k4[i] + R(a75) * k5[i] + R(a76) * k6[i]);
    F.deriv(t_current + R(c7) * h, yy1, k7);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a81) * k1[i] + R(a84) * k4[i] + R(a85) * k5[i] + R(a86) * k6[i] + R(a87) * k7[i]);
    F.deriv(t_current + R(c8) * h, yy1, k8);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a91) * k1[i] + R(a94) * k4[i] + R(a95) * k5[i] + R(a96) * k6[i] + R(a97) * k7[i] +
                            R(a98) * k8[i]);
    F.deriv(t_current + R(c9) * h, yy1, k9);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a101) * k1[i] + R(a104) * k4[i] + R(a105) * k5[i] + R(a106) * k6[i] +
                            R(a107) * k7[i] + R(a108) * k8[i] + R(a109) * k9[i]);
    F.deriv(t_current + R(c10) * h, yy1, k10);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a111) * k1[i] + R(a114) * k4[i] + R(a115) * k5[i] + R(a116) * k6[i] +
                            R(a117) * k7[i] + R(a118) * k8[i] + R(a119) * k9[i] + R(a1110) * k10[i]);
    F.deriv(t_current + R(c11) * h, yy1, k2);
    for (int i = 0; i < dim; i++)
      yy1[i] = y0[i] + h * (R(a121) * k1[i] + R(a124) * k4[i] + R(a125) * k5[i] + R(a126) * k6[i] +
                            R(a127) * k7[i] + R(a128) * k8[i] + R(a129) * k9[i] + R(a1210) * k10[i] +
                            R(a1211) * k2[i]);
    t_old_older = t_old;
    t_old = t_current;
    t_current = t_current + h;
    F.deriv(t_current, yy1, k3);
    for (int i = 0; i < dim; i++) {
      k4[i] = R(b1) * k1[i] + R(b6) * k6[i] + R(b7) * k7[i] + R(b8) * k8[i] + R(b9) * k9[i] + R(b10) * k10[i] +
              R(b11) * k2[i] + R(b12) * k3[i];
      k5[i] = y0[i] + h * k4[i];
    }
    R errv = 0.0, err2 = 0.0, erri;
    for (int i = 0; i < dim; i++) {
      sk = atol + rtol * rmax(fabs(y0[i]), fabs(k5[i]));
      erri = k4[i] - R(bhh1) * k1[i] - R(bhh2) * k9[i] - R(bhh3) * k3[i];
      sqr = erri / sk;
      err2 += sqr * sqr;
      erri = R(er1) * k1[i] + R(er6) * k6[i] + R(er7) * k7[i] + R(er8) * k8[i] + R(er9) * k9[i] +
             R(er10) * k10[i] + R(er11) * k2[i] + R(er12) * k3[i];
      sqr = erri / sk;
      errv += sqr * sqr;
    }
    R deno = errv + R(0.01) * err2;
    if (deno <= R(0.0)) deno = 1.0;
    errv = fabs(h) * errv * sqrt(R(1.0) / (deno * R((double)dim)));
    R fac11 = pow(errv, expo1);
    R fac = fac11 / pow(facold, beta);
    fac = rmax(facc2, rmin(facc1, fac / safe));
    R hnew = h / fac;
    if (errv <= R(1.0)) {
      nacc++;
      facold = rmax(errv, R(1.0e-4));
      F.deriv(t_current, k5, k4);
      // The dense-output coefficients and the three extra stages
      // (leung_dop853.c:430-463) feed only the interpolation at requested
      // times inside this step; the next step recomputes k2, k3 and k10 before
      // reading them.  They are formed only when a requested time falls in
      // this step, which leaves every output unchanged.
      const bool need_dense = pos_neg * R(t[done + 1]) < pos_neg * t_current;
      R ydiff, bspl;
      if (need_dense) {
      for (int i = 0; i < dim; i++) {
        rc1[i] = y0[i];
        ydiff = k5[i] - y0[i];
        rc2[i] = ydiff;
        bspl = h * k1[i] - ydiff;
        rc3[i] = bspl;
        rc4[i] = ydiff - h * k4[i] - bspl;
        rc5[i] = R(d41) * k1[i] + R(d46) * k6[i] + R(d47) * k7[i] + R(d48) * k8[i] + R(d49) * k9[i] +
                 R(d410) * k10[i] + R(d411) * k2[i] + R(d412) * k3[i];
        rc6[i] = R(d51) * k1[i] + R(d56) * k6[i] + R(d57) * k7[i] + R(d58) * k8[i] + R(d59) * k9[i] +
                 R(d510) * k10[i] + R(d511) * k2[i] + R(d512) * k3[i];
        rc7[i] = R(d61) * k1[i] + R(d66) * k6[i] + R(d67) * k7[i] + R(d68) * k8[i] + R(d69) * k9[i] +
                 R(d610) * k10[i] + R(d611) * k2[i] + R(d612) * k3[i];
        rc8[i] = R(d71) * k1[i] + R(d76) * k6[i] + R(d77) * k7[i] + R(d78) * k8[i] + R(d79) * k9[i] +
                 R(d710) * k10[i] + R(d711) * k2[i] + R(d712) * k3[i];
      }
      for (int i = 0; i < dim; i++)
        yy1[i] = y0[i] + h * (R(a141) * k1[i] + R(a147) * k7[i] + R(a148) * k8[i] + R(a149) * k9[i] +
                              R(a1410) * k10[i] + R(a1411) * k2[i] + R(a1412) * k3[i] + R(a1413) * k4[i]);
      F.deriv(t_old + R(c14) * h, yy1, k10);
      for (int i = 0; i < dim; i++)
        yy1[i] = y0[i] + h * (R(a151) * k1[i] + R(a156) * k6[i] + R(a157) * k7[i] + R(a158) * k8[i] +
                              R(a1511) * k2[i] + R(a1512) * k3[i] + R(a1513) * k4[i] + R(a1514) * k10[i]);
      F.deriv(t_old + R(c15) * h, yy1, k2);
      for (int i = 0; i < dim; i++)
        yy1[i] = y0[i] + h * (R(a161) * k1[i] + R(a166) * k6[i] + R(a167) * k7[i] + R(a168) * k8[i] +
                              R(a169) * k9[i] + R(a1613) * k4[i] + R(a1614) * k10[i] + R(a1615) * k2[i]);
      F.deriv(t_old + R(c16) * h, yy1, k3);
      for (int i = 0; i < dim; i++) {
        rc5[i] = h * (rc5[i] + R(d413) * k4[i] + R(d414) * k10[i] + R(d415) * k2[i] + R(d416) * k3[i]);
        rc6[i] = h * (rc6[i] + R(d513) * k4[i] + R(d514) * k10[i] + R(d515) * k2[i] + R(d516) * k3[i]);
        rc7[i] = h * (rc7[i] + R(d613) * k4[i] + R(d614) * k10[i] + R(d615) * k2[i] + R(d616) * k3[i]);
        rc8[i] = h * (rc8[i] + R(d713) * k4[i] + R(d714) * k10[i] + R(d715) * k2[i] + R(d716) * k3[i]);
      }
      }
      while (done < nt - 1 && pos_neg * R(t[done + 1]) < pos_neg * t_current) {
        R s = (R(t[done + 1]) - t_old) / h;
        R s1 = R(1.0) - s;
        for (int i = 0; i < dim; i++)
          result[i] = dval(rc1[i] + s * (rc2[i] + s1 * (rc3[i] + s * (rc4[i] + s1 * (rc5[i] + s * (rc6[i] +
                                                       s1 * (rc7[i] + s * rc8[i])))))));
        result += dim;
        done++;
      }
      for (int i = 0; i < dim; i++) {
        k1[i] = k4[i];
        y0[i] = k5[i];
      }
      hnew = (fabs(hnew) > fabs(hmax)) ? pos_neg * hmax : hnew;
      if (reject) hnew = pos_neg * rmin(fabs(hnew), fabs(h));
      reject = 0;
    } else {
      nrej++;
      hnew = h / rmin(facc1, fac11 / safe);
      reject = 1;
      t_current = t_old;
      t_old = t_old_older;
    }
    h = hnew;
  }
  for (int i = 0; i < dim; i++) y0d[i] = dval(y0[i]);
  if (st) { st->steps = nacc; st->rejected = nrej; st->dt = 0.; }
}

// ---------------------------------------------------------------------------
// IAS15 (wez_ias15.c:137-425), galpy's 15th-order Gauss-Radau integrator
// (Everhart 1985; Rein & Spiegel 2015) on the rectangular positions q and
// velocities p.  Per step: the force at the step start, then a
// predictor-corrector loop over the seven Radau substeps that refines the
// polynomial coefficients B of the force (through their divided-difference
// form G) until the change of G[6] relative to the largest start-of-step
// acceleration component falls below 1e-16, or 12 passes; the step is
// accepted when the next step dt (B[6] based, 1e-9 precision parameter,
// seventh root by Newton's method) is not smaller than the one taken.
// Outputs come from the step's interpolating polynomial at every grid time
// the accepted step passes.  Behaviour kept from the reference:
//   * every force in a step is evaluated at the step's END time to + dt;
//   * the integration stops once the accepted steps add up to |nt dt|
//     (dt the grid's first spacing, or the caller's dt), one dt past the
//     last output time; rows never passed stay unwritten;
//   * B is carried to the next step by the binomial predictor at step
//     ratio 1 and is not reset after a rejected step.
// ---------------------------------------------------------------------------
namespace radau {
// Gauss-Radau spacings on [0, 1] (h[0] = 0, then the 7 nodes of the 8-point
// Radau rule), wez_ias15.c:43-52
constexpr double h[8] = {0.0,
                         0.05626256053692215,
                         0.18024069173689236,
                         0.35262471711316964,
                         0.54715362633055538,
                         0.73421017721541053,
                         0.88532094683909577,
                         0.97752061356128750};
// rr[k][j] = 1 / (h[k] - h[j]), j < k (wez_ias15.c:113-148)
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
// G -> B conversion, cc[k][j] (c_{k,j} of wez_ias15.c:54-81): B[j] gains
// cc[k-1][j] times the change of G[k-1], B[k-1] the change itself
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
// B -> G conversion, dd[k][j] (d_{k,j} of wez_ias15.c:83-111): G[j] =
// sum over k >= j of dd[k][j] B[k]
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
constexpr double kDeltaB = 1e-16;    // predictor-corrector convergence (e_deltab)
constexpr double kPrecision = 1e-9;  // step-size precision parameter (e_b)
constexpr int kMaxPasses = 12;       // predictor-corrector passes per step
}  // namespace radau

// a^(1/7) by 20 Newton steps from 1 after scaling a into [1e-7, 1e2] by
// factors of 1e7 (wez_ias15.c:587-603); NaN/inf/zero are not scaled.
template <class R> R ias15_seventhroot(R a) {
  R scale = 1.;
  while (a < R(1e-7) && std::isnormal(dval(a))) {
    scale = scale * R(0.1);
    a = a * R(1e7);
  }
  while (a > R(1e2) && std::isnormal(dval(a))) {
    scale = scale * R(10.);
    a = a * R(1e-7);
  }
  R x = 1.;
  for (int k = 0; k < 20; k++) {
    R x6 = x * x * x * x * x * x;
    x += (a / x6 - x) / R(7.);
  }
  return x * scale;
}

// position at substep fraction hn of a step of length T (wez_ias15.c:459-476)
template <int D, class R>
void ias15_position(R* y, const R* y1, const R* v, R hn, R T, const R (*F0), const R (*B)[D]) {
  for (int i = 0; i < D; i++) {
    const R poly =
        F0[i] / R(2.) +
        hn * (B[0][i] / R(6.) +
              hn * (B[1][i] / R(12.) +
                    hn * (B[2][i] / R(20.) +
                          hn * (B[3][i] / R(30.) +
                                hn * (B[4][i] / R(42.) +
                                      hn * (B[5][i] / R(56.) + hn * (B[6][i] / R(72.))))))));
    y[i] = y1[i] + (v[i] * hn * T) + ((hn * T) * (hn * T)) * poly;
  }
}
// velocity at substep fraction hn (wez_ias15.c:441-456)
template <int D, class R>
void ias15_velocity(R* v, const R* v1, R hn, R T, const R* F0, const R (*B)[D]) {
  for (int i = 0; i < D; i++) {
    const R poly =
        F0[i] + hn * (B[0][i] / R(2.) +
                      hn * (B[1][i] / R(3.) +
                            hn * (B[2][i] / R(4.) +
                                  hn * (B[3][i] / R(5.) +
                                        hn * (B[4][i] / R(6.) +
                                              hn * (B[5][i] / R(7.) + hn * (B[6][i] / R(8.))))))));
    v[i] = v1[i] + (hn * T * poly);
  }
}

template <int D, class R>
void ias15(const ForceCtx<R, D>& F, const double* yo, int nt, double dt_in, const double* t,
           double, double, double* result, int* err, IntStats* st) {
  using namespace radau;
  R x[D], v[D], xs[D], vs[D], a[D], F0[D];
  R B[7][D], E[7][D], G[7][D];
  for (int i = 0; i < D; i++) {
    x[i] = yo[i];
    v[i] = yo[D + i];
  }
  for (int k = 0; k < 7; k++)
    for (int i = 0; i < D; i++) B[k][i] = E[k][i] = G[k][i] = 0.;
  for (int i = 0; i < 2 * D; i++) result[i] = yo[i];
  *err = 0;
  const R init_dt = R(t[1]) - R(t[0]);
  R dt = dt_in == kDtSentinel ? init_dt : R(dt_in);
  R to = t[0];
  R time_remaining = fabs(R(nt) * dt);
  int row = 1;
  long long nacc = 0, nrej = 0;
  while (time_remaining > R(0.)) {
    const R dt_temp = dt;
    const R to_temp = to + dt_temp;
    F.force(to_temp, x, a);
    for (int i = 0; i < D; i++) F0[i] = a[i];
    for (int j = 0; j < 7; j++)  // G from B (update_Gs_from_Bs, wez_ias15.c:514-529)
      for (int i = 0; i < D; i++) {
        R g = dd[j][j] * B[j][i];
        for (int k = j + 1; k < 7; k++) g = g + dd[k][j] * B[k][i];
        G[j][i] = g;
      }
    R ierr = R(kDeltaB + 1.);
    for (int pass = 0; pass < kMaxPasses && !(ierr < R(kDeltaB)); pass++) {
      R max_dg = 0., max_a = 0.;
      for (int k = 1; k < 8; k++) {
        ias15_position<D>(xs, x, v, R(h[k]), dt_temp, F0, B);
        F.force(to_temp, xs, a);
        for (int i = 0; i < D; i++) {
          // G[k-1] from the new force by divided differences (wez_ias15.c:478-496)
          R g = (a[i] - F0[i]) * R(rr[k][0]);
          for (int j = 1; j < k; j++) g = (g - G[j - 1][i]) * R(rr[k][j]);
          const R dg = g - G[k - 1][i];
          G[k - 1][i] = g;
          // B from the change of G[k-1] (wez_ias15.c:554-585)
          for (int j = 0; j < k - 1; j++) B[j][i] += dg * R(cc[k - 1][j]);
          B[k - 1][i] += dg;
          if (k == 7) {
            if (fabs(dg) > max_dg) max_dg = fabs(dg);
            if (fabs(F0[i]) > max_a) max_a = fabs(F0[i]);
          }
        }
      }
      ierr = max_dg / max_a;
    }
    R max_b6 = 0., max_a = 0.;
    for (int i = 0; i < D; i++) {
      if (fabs(B[6][i]) > max_b6) max_b6 = fabs(B[6][i]);
      if (fabs(F0[i]) > max_a) max_a = fabs(F0[i]);
    }
    const R dt_required = dt_temp * ias15_seventhroot(R(kPrecision) / (max_b6 / max_a));
    if (fabs(dt_temp) > fabs(dt_required)) {
      dt = dt_required;
      nrej++;
      continue;
    }
    nacc++;
    time_remaining -= fabs(dt);
    // outputs inside the step (wez_ias15.c:381-405)
    while (row < nt && (init_dt > R(0.) ? (to < R(t[row]) && R(t[row]) <= to_temp)
                                        : (to > R(t[row]) && R(t[row]) >= to_temp))) {
      const R hs = fabs(R(t[row]) - to) / fabs(to_temp - to);
      ias15_position<D>(xs, x, v, hs, dt_temp, F0, B);
      ias15_velocity<D>(vs, v, hs, dt_temp, F0, B);
      for (int i = 0; i < D; i++) {
        result[2 * D * row + i] = dval(xs[i]);
        result[2 * D * row + D + i] = dval(vs[i]);
      }
      row++;
    }
    ias15_position<D>(x, x, v, R(1.), dt_temp, F0, B);
    ias15_velocity<D>(v, v, R(1.), dt_temp, F0, B);
    // binomial predictor of the next step's B at step ratio 1 (wez_ias15.c:531-552)
    for (int i = 0; i < D; i++) {
      R bd[7], en[7];
      for (int k = 0; k < 7; k++) bd[k] = B[k][i] - E[k][i];
      en[0] = (B[0][i] + R(2.) * B[1][i] + R(3.) * B[2][i] + R(4.) * B[3][i] + R(5.) * B[4][i] +
               R(6.) * B[5][i] + R(7.) * B[6][i]);
      en[1] = (B[1][i] + R(3.) * B[2][i] + R(6.) * B[3][i] + R(10.) * B[4][i] + R(15.) * B[5][i] +
               R(21.) * B[6][i]);
      en[2] = (B[2][i] + R(4.) * B[3][i] + R(10.) * B[4][i] + R(20.) * B[5][i] + R(35.) * B[6][i]);
      en[3] = (B[3][i] + R(5.) * B[4][i] + R(15.) * B[5][i] + R(35.) * B[6][i]);
      en[4] = (B[4][i] + R(6.) * B[5][i] + R(21.) * B[6][i]);
      en[5] = (B[5][i] + R(7.) * B[6][i]);
      en[6] = B[6][i];
      for (int k = 0; k < 7; k++) {
        E[k][i] = en[k];
        B[k][i] = en[k] + bd[k];
      }
    }
    to = to_temp;
    dt = dt_required;
  }
  // rows never reached (a caller's dt shorter than the grid spacing): the
  // reference leaves them uninitialised; NaN here, as on the device
  for (; row < nt; row++)
    for (int i = 0; i < 2 * D; i++) result[2 * D * row + i] = std::nan("");
  if (st) { st->steps = nacc; st->rejected = nrej; st->dt = dval(dt); }
}

}  // namespace orc
