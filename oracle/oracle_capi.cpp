// CPU oracle C-ABI -- TEST INFRASTRUCTURE ONLY.
//
// `oracle_integrateFullOrbit` has the signature of galpy's
// `integrateFullOrbit` (galpy/orbit/orbit_c_ext/integrateFullOrbit.c:1046-1061)
// and the same OpenMP structure (:1133, schedule(dynamic,1) over orbits);
// `oracle_integratePlanarOrbit` is the same for galpy's `integratePlanarOrbit`
// (galpy/orbit/orbit_c_ext/integratePlanarOrbit.c:743-845).
// The remaining entry points expose the pieces (forces, potential values,
// incomplete gamma, coordinate transforms) for unit checks, plus an
// op-counting run that yields the algorithmic FLOP count per orbit-step.
#include <omp.h>
#include <cstdio>
#include <vector>
#include "integrators.hpp"

namespace orc {
thread_local OpCount g_ops = {0, 0, 0, 0, 0};
}

using namespace orc;

#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

// The caller's time functions for the following calls (galpy's pot_tfuncs;
// NULL / 0 for none).
ORC_EXPORT void oracle_set_tfuncs(double (**fns)(double), int n) {
  orc::g_tfuncs = fns;
  orc::g_ntfuncs = n;
}

namespace {

constexpr int kMaxPot = kOracleMaxComp;

// One orbit: integrateFullOrbit.c:1133-1141 (D = 3, rows of 6) or
// integratePlanarOrbit.c:826-835 (D = 2, rows of 4).
template <int D, class R>
void integrate_one(int npot, const Component* P, int odeint_type, double* yo, int nt,
                   const double* t, double dt, double rtol, double atol, double* result,
                   int* err, IntStats* st) {
  constexpr int N = 2 * D;
  long long nrhs = 0;
  ForceCtx<R, D> F{npot, P, &nrhs};
  IntStats local = {0, 0, 0, 0.};
  // in place, as the reference (integrateFullOrbit.c:1135, integratePlanarOrbit.c:828)
  if (D == 3) cyl_to_rect(yo);
  else polar_to_rect(yo);
  switch (odeint_type) {
    case 0: leapfrog<D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 3: symplec4<D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 1: rk_fixed<4, D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 2: rk_fixed<6, D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 4: symplec6<D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 5: dopr54<D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    case 6: dop853<D>(F, yo, nt, t, rtol, atol, result, err, &local); break;
    case 7: ias15<D>(F, yo, nt, dt, t, rtol, atol, result, err, &local); break;
    default: *err = -1; return;
  }
  for (int j = 0; j < nt; j++) {  // integrateFullOrbit.c:1139-1140, integratePlanarOrbit.c:832-833
    if (D == 3) rect_to_cyl(result + N * j);
    else rect_to_polar(result + N * j);
  }
  local.rhs = nrhs;
  if (st) *st = local;
}

template <int D>
int run(int nobj, double* yo, int nt, const double* t, int indiv_t, int npot,
        const int* pot_type, const double* pot_args, double dt, double rtol, double atol,
        double* result, int* err, int odeint_type, void (*cb)(void), long long* steps,
        long long* rejected, long long* nrhs, double* dt_used, int nthreads) {
  constexpr int N = 2 * D;
  if (npot > kMaxPot || npot < 0 || nt < 2) {
    for (int i = 0; i < nobj; i++) err[i] = -1;
    return -1;
  }
  Component P[kMaxPot];
  if (parse_potential(npot, pot_type, pot_args, P, D == 2) != 0) {
    for (int i = 0; i < nobj; i++) err[i] = -1;
    return -1;
  }
  int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
  if (nobj < nth) nth = nobj > 0 ? nobj : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
  for (int ii = 0; ii < nobj; ii++) {
    IntStats st = {0, 0, 0, 0.};
    integrate_one<D, double>(npot, P, odeint_type, yo + N * ii, nt, t + (long)nt * ii * indiv_t,
                             dt, rtol, atol, result + (long)N * nt * ii, err + ii, &st);
    if (steps) steps[ii] = st.steps;
    if (rejected) rejected[ii] = st.rejected;
    if (nrhs) nrhs[ii] = st.rhs;
    if (dt_used) dt_used[ii] = st.dt;
    if (cb) cb();
  }
  return 0;
}

}  // namespace

// Same signature and argument meaning as galpy's integrateFullOrbit.
// pot_tfuncs must be NULL (none of the supported potentials use them).
ORC_EXPORT void oracle_integrateFullOrbit(int nobj, double* yo, int nt, double* t, int indiv_t,
                                          int npot, int* pot_type, double* pot_args,
                                          void* pot_tfuncs, double dt, double rtol, double atol,
                                          double* result, int* err, int odeint_type,
                                          void (*cb)(void)) {
  (void)pot_tfuncs;
  run<3>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result, err,
         odeint_type, cb, nullptr, nullptr, nullptr, nullptr, 0);
}

// Same signature and argument meaning as galpy's integratePlanarOrbit
// (integratePlanarOrbit.c:743-758): rows of (R,vR,vT,phi), planar packing.
ORC_EXPORT void oracle_integratePlanarOrbit(int nobj, double* yo, int nt, double* t, int indiv_t,
                                            int npot, int* pot_type, double* pot_args,
                                            void* pot_tfuncs, double dt, double rtol, double atol,
                                            double* result, int* err, int odeint_type,
                                            void (*cb)(void)) {
  (void)pot_tfuncs;
  run<2>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result, err,
         odeint_type, cb, nullptr, nullptr, nullptr, nullptr, 0);
}

// As above plus per-orbit statistics and an explicit thread count (0 = all).
ORC_EXPORT int oracle_integrate_stats(int nobj, double* yo, int nt, double* t, int indiv_t,
                                      int npot, int* pot_type, double* pot_args, double dt,
                                      double rtol, double atol, double* result, int* err,
                                      int odeint_type, long long* steps, long long* rejected,
                                      long long* nrhs, double* dt_used, int nthreads) {
  return run<3>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result, err,
                odeint_type, nullptr, steps, rejected, nrhs, dt_used, nthreads);
}

// The planar twin of oracle_integrate_stats (rows of 4, planar packing).
ORC_EXPORT int oracle_integrate_planar_stats(int nobj, double* yo, int nt, double* t, int indiv_t,
                                             int npot, int* pot_type, double* pot_args, double dt,
                                             double rtol, double atol, double* result, int* err,
                                             int odeint_type, long long* steps, long long* rejected,
                                             long long* nrhs, double* dt_used, int nthreads) {
  return run<2>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result, err,
                odeint_type, nullptr, steps, rejected, nrhs, dt_used, nthreads);
}

namespace {
template <int D>
int count_ops(int nobj, double* yo, int nt, double* t, int indiv_t, int npot, int* pot_type,
              double* pot_args, double dt, double rtol, double atol, double* result, int* err,
              int odeint_type, double* out) {
  constexpr int N = 2 * D;
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P, D == 2) != 0) return -1;
  g_ops = OpCount{0, 0, 0, 0, 0};
  long long steps = 0, rej = 0, rhs = 0;
  for (int ii = 0; ii < nobj; ii++) {
    IntStats st = {0, 0, 0, 0.};
    // the coordinate transforms act on plain doubles and are not counted:
    // they are per output row, not per orbit-step
    integrate_one<D, CountedD>(npot, P, odeint_type, yo + N * ii, nt,
                               t + (long)nt * ii * indiv_t, dt, rtol, atol,
                               result + (long)N * nt * ii, err + ii, &st);
    steps += st.steps;
    rej += st.rejected;
    rhs += st.rhs;
  }
  out[0] = (double)g_ops.add;
  out[1] = (double)g_ops.mul;
  out[2] = (double)g_ops.div;
  out[3] = (double)g_ops.sqrt;
  out[4] = (double)g_ops.special;
  out[5] = (double)steps;
  out[6] = (double)rej;
  out[7] = (double)rhs;
  return 0;
}
}  // namespace

// Op-counting run (single thread).  out[0..7] = add, mul, div, sqrt, special,
// orbit-steps, rejected steps, force/RHS evaluations, summed over the orbits.
ORC_EXPORT int oracle_count_ops(int nobj, double* yo, int nt, double* t, int indiv_t, int npot,
                                int* pot_type, double* pot_args, double dt, double rtol,
                                double atol, double* result, int* err, int odeint_type,
                                double* out) {
  return count_ops<3>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result,
                      err, odeint_type, out);
}

// The same for planar orbits (rows of 4, planar packing).
ORC_EXPORT int oracle_count_ops_planar(int nobj, double* yo, int nt, double* t, int indiv_t,
                                       int npot, int* pot_type, double* pot_args, double dt,
                                       double rtol, double atol, double* result, int* err,
                                       int odeint_type, double* out) {
  return count_ops<2>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result,
                      err, odeint_type, out);
}

// Planar forces (calcPlanarRforce, calcPlanarphitorque) at n points, from
// the planar packing.
ORC_EXPORT int oracle_eval_planar_forces(int npot, int* pot_type, double* pot_args, int n,
                                         const double* R, const double* phi, const double* t,
                                         double* FR, double* Fphi) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P, true) != 0) return -1;
  for (int i = 0; i < n; i++) planar_forces<double>(npot, P, R[i], phi[i], t[i], FR[i], Fphi[i]);
  return 0;
}

// Planar rectangular accelerations (evalPlanarRectForce), q[n*2] -> a[n*2].
ORC_EXPORT int oracle_planar_rect_force(int npot, int* pot_type, double* pot_args, int n,
                                        const double* t, const double* q, double* a) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P, true) != 0) return -1;
  for (int i = 0; i < n; i++) planar_rect_force<double>(npot, P, t[i], q + 2 * i, a + 2 * i);
  return 0;
}

ORC_EXPORT void oracle_polar_to_rect(int n, double* v) {
  for (int i = 0; i < n; i++) polar_to_rect<double>(v + 4 * i);
}
ORC_EXPORT void oracle_rect_to_polar(int n, double* v) {
  for (int i = 0; i < n; i++) rect_to_polar<double>(v + 4 * i);
}

// Cylindrical forces (FR, Fz, Fphi) at n points.
ORC_EXPORT int oracle_eval_forces(int npot, int* pot_type, double* pot_args, int n,
                                  const double* R, const double* z, const double* phi,
                                  const double* t, double* FR, double* Fz, double* Fphi) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++) cyl_forces<double>(npot, P, R[i], z[i], phi[i], t[i], FR[i], Fz[i], Fphi[i]);
  return 0;
}

// Density (calcDensity, galpy_potentials.c:380-390) at n points.
ORC_EXPORT int oracle_eval_density(int npot, int* pot_type, double* pot_args, int n, const double* R,
                                   const double* z, const double* phi, double* out) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++) {
    out[i] = 0.;
    for (int k = 0; k < npot; k++) out[i] += comp_density(P[k], R[i], z[i], phi[i]);
  }
  return 0;
}

// The velocity-dependent forces (F_R, F_z, F_phi) at n phase-space points.
ORC_EXPORT int oracle_eval_diss_forces(int npot, int* pot_type, double* pot_args, int n, const double* R,
                                       const double* z, const double* phi, const double* t,
                                       const double* vR, const double* vT, const double* vz, double* FR,
                                       double* Fz, double* Fphi) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++) diss_forces(npot, P, R[i], z[i], phi[i], t[i], vR[i], vT[i], vz[i], FR[i], Fz[i], Fphi[i]);
  return 0;
}

// The cosine integral Ci at n points (x > 0).
ORC_EXPORT void oracle_cos_integral(int n, const double* x, double* out) {
  for (int i = 0; i < n; i++) out[i] = cos_integral(x[i]);
}

// Potential value at n points.
ORC_EXPORT int oracle_eval_potential(int npot, int* pot_type, double* pot_args, int n,
                                     const double* R, const double* z, const double* phi,
                                     const double* t, double* out) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++) out[i] = potential_value<double>(npot, P, R[i], z[i], phi[i], t[i]);
  return 0;
}

// Rectangular accelerations (evalRectForce) at n points q[n*3] -> a[n*3].
ORC_EXPORT int oracle_rect_force(int npot, int* pot_type, double* pot_args, int n,
                                 const double* t, const double* q, double* a) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++) rect_force<double>(npot, P, t[i], q + 3 * i, a + 3 * i);
  return 0;
}

ORC_EXPORT void oracle_lower_incgamma(double a, int n, const double* x, double* out) {
  double ga = std::tgamma(a);
  for (int i = 0; i < n; i++) out[i] = lower_incgamma<double>(a, x[i], ga);
}
ORC_EXPORT void oracle_upper_incgamma(double a, int n, const double* x, double* out) {
  double ga = std::tgamma(a);
  for (int i = 0; i < n; i++) out[i] = upper_incgamma<double>(a, x[i], ga);
}
ORC_EXPORT void oracle_cyl_to_rect(int n, double* v) {
  for (int i = 0; i < n; i++) cyl_to_rect<double>(v + 6 * i);
}
ORC_EXPORT void oracle_rect_to_cyl(int n, double* v) {
  for (int i = 0; i < n; i++) rect_to_cyl<double>(v + 6 * i);
}

// Tangent-map integration of n orbits (bovy_symplecticode.c:1053-1268 with
// integrateFullOrbit.c:1619-1681): w[n][(1+nde)*6] rectangular base state and
// deviation columns in, out[n][(1+nde)*6] final augmented state.  method 0
// (leapfrog) or 3 (symplec4); nde <= 6.  Returns -1 on an unsupported
// component or method.
ORC_EXPORT int oracle_symplectic_dxdv(int n, const double* w, int nde, int nt, const double* t,
                                      int npot, int* pot_type, double* pot_args, double dt,
                                      double rtol, double atol, int method, double* out) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  if ((method != 0 && method != 3) || nde < 0 || nde > 6) return -1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int i = 0; i < n; i++)
    bad += !symplectic_dxdv<double>(npot, P, method, w + (long)i * 6 * (1 + nde), nde, nt, t, dt,
                                    rtol, atol, out + (long)i * 6 * (1 + nde));
  return bad ? -1 : 0;
}

// Cylindrical second derivatives (R2deriv, z2deriv, Rzderiv) at n points.
ORC_EXPORT int oracle_eval_hessian(int npot, int* pot_type, double* pot_args, int n,
                                   const double* R, const double* z, double* R2, double* z2,
                                   double* Rz) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++)
    if (!cyl_hessian<double>(npot, P, R[i], z[i], R2[i], z2[i], Rz[i])) return -1;
  return 0;
}

// All six cylindrical second derivatives (calcR2deriv, calcz2deriv,
// calcRzderiv, calcphi2deriv, calcRphideriv, calczphideriv) at (R, z, phi, t).
ORC_EXPORT int oracle_eval_hessian_full(int npot, int* pot_type, double* pot_args, int n,
                                        const double* R, const double* z, const double* phi,
                                        const double* t, double* R2, double* z2, double* Rz,
                                        double* phi2, double* Rphi, double* zphi) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P) != 0) return -1;
  for (int i = 0; i < n; i++)
    if (!cyl_hessian_full<double>(npot, P, R[i], z[i], phi[i], t[i], R2[i], z2[i], Rz[i], phi2[i],
                                  Rphi[i], zphi[i]))
      return -1;
  return 0;
}

// calcPlanarR2deriv, calcPlanarphi2deriv, calcPlanarRphideriv at (R, phi, t)
// (planar packing).
ORC_EXPORT int oracle_eval_planar_hessian(int npot, int* pot_type, double* pot_args, int n,
                                          const double* R, const double* phi, const double* t,
                                          double* R2, double* phi2, double* Rphi) {
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P, true) != 0) return -1;
  for (int i = 0; i < n; i++)
    if (!planar_hessian<double>(npot, P, R[i], phi[i], t[i], R2[i], phi2[i], Rphi[i])) return -1;
  return 0;
}

// ---------------------------------------------------------------------------
// integrate_dxdv: the orbit plus one phase-space deviation, rectangular rows
// (q, p | dq, dp) of 12 (full) or 8 (planar) doubles; symplectic methods by
// the tangent-map steppers, DOP853 on the variational system.
// ---------------------------------------------------------------------------
namespace {
// integrateFullOrbit_dxdv, integrateFullOrbit.c:1234-1306 (DB = 3) and
// integratePlanarOrbit_dxdv, integratePlanarOrbit.c:940-1012 (DB = 2).
template <int DB, class R = double>
void dxdv_one(int npot, const Component* P, int odeint_type, double* yo, int nt, const double* t,
              double dt, double rtol, double atol, double* result, int* err,
              IntStats* st = nullptr) {
  *err = 0;
  bool ok = true;
  switch (odeint_type) {
    case 0:
    case 3:
    case 4:
      ok = symplectic_dxdv_traj<DB, R>(npot, P, odeint_type, yo, 1, nt, t, dt, rtol, atol,
                                       nullptr, result);
      if (st) {  // orbit-steps: the step estimate again, in plain (uncounted) doubles
        long long nrhs = 0;
        ForceCtx<double, DB> F{npot, P, &nrhs};
        double qo[3], po[3];
        for (int k = 0; k < DB; k++) { qo[k] = yo[k]; po[k] = yo[DB + k]; }
        const double init_dt = t[1] - t[0];
        double d = dt;
        if (dt == kDtSentinel)
          d = (odeint_type == 0)   ? leapfrog_estimate_step<DB>(F, qo, po, init_dt, t[0], rtol, atol)
              : (odeint_type == 4) ? symplec6_estimate_step<DB>(F, qo, po, init_dt, t[0], rtol, atol)
                                   : symplec4_estimate_step<DB>(F, qo, po, init_dt, t[0], rtol, atol);
        st->steps = (long long)(init_dt / d) * (nt - 1);
      }
      break;
    case 1:  // RK4, RK6, DOPR54 and DOP853 on the variational system
    case 2:  // (evalRectDeriv_dxdv, integrateFullOrbit.c:1272-1296)
    case 5:
    case 6: {
      long long nrhs = 0;
      ForceCtx<R, 2 * DB> F{npot, P, &nrhs};
      if (odeint_type == 1) rk_fixed<4, 2 * DB>(F, yo, nt, dt, t, rtol, atol, result, err, st);
      else if (odeint_type == 2) rk_fixed<6, 2 * DB>(F, yo, nt, dt, t, rtol, atol, result, err, st);
      else if (odeint_type == 5) dopr54<2 * DB>(F, yo, nt, dt, t, rtol, atol, result, err, st);
      else dop853<2 * DB>(F, yo, nt, t, rtol, atol, result, err, st);
      break;
    }
    default:
      *err = -1;
  }
  if (!ok) *err = -1;
}

template <int DB>
int run_dxdv(int nobj, double* yo, int nt, const double* t, int indiv_t, int npot,
             const int* pot_type, const double* pot_args, double dt, double rtol, double atol,
             double* result, int* err, int odeint_type, int nthreads,
             long long* steps = nullptr) {
  constexpr int N = 4 * DB;
  Component P[kMaxPot];
  if (npot > kMaxPot || npot < 0 || nt < 2 ||
      parse_potential(npot, pot_type, pot_args, P, DB == 2) != 0 || !has_hessian(npot, P)) {
    for (int i = 0; i < nobj; i++) err[i] = -1;
    return -1;
  }
  int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
  if (nobj < nth) nth = nobj > 0 ? nobj : 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
  for (int ii = 0; ii < nobj; ii++) {
    IntStats st = {0, 0, 0, 0.};
    dxdv_one<DB>(npot, P, odeint_type, yo + (long)N * ii, nt, t + (long)nt * ii * indiv_t, dt,
                 rtol, atol, result + (long)N * nt * ii, err + ii, steps ? &st : nullptr);
    if (steps) steps[ii] = st.steps;
  }
  return 0;
}
}  // namespace

// Same signature and argument meaning as galpy's integrateFullOrbit_dxdv
// (one orbit; yo[12], result[nt*12]).
ORC_EXPORT void oracle_integrateFullOrbit_dxdv(double* yo, int nt, double* t, int npot,
                                               int* pot_type, double* pot_args, void* pot_tfuncs,
                                               double dt, double rtol, double atol,
                                               double* result, int* err, int odeint_type) {
  (void)pot_tfuncs;
  run_dxdv<3>(1, yo, nt, t, 0, npot, pot_type, pot_args, dt, rtol, atol, result, err,
              odeint_type, 1);
}

// Same signature and argument meaning as galpy's integratePlanarOrbit_dxdv
// (one orbit; yo[8], result[nt*8]; planar packing).
ORC_EXPORT void oracle_integratePlanarOrbit_dxdv(double* yo, int nt, double* t, int npot,
                                                 int* pot_type, double* pot_args,
                                                 void* pot_tfuncs, double dt, double rtol,
                                                 double atol, double* result, int* err,
                                                 int odeint_type, void (*cb)(void)) {
  (void)pot_tfuncs;
  (void)cb;
  run_dxdv<2>(1, yo, nt, t, 0, npot, pot_type, pot_args, dt, rtol, atol, result, err,
              odeint_type, 1);
}

// Many orbits at once (dim 3: rows of 12, full packing; dim 2: rows of 8,
// planar packing), OpenMP over orbits; steps[i] (nullable) = orbit-steps.
ORC_EXPORT int oracle_integrate_dxdv(int nobj, int dim, double* yo, int nt, double* t, int indiv_t,
                                     int npot, int* pot_type, double* pot_args, double dt,
                                     double rtol, double atol, double* result, int* err,
                                     int odeint_type, int nthreads, long long* steps) {
  if (dim == 3)
    return run_dxdv<3>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result,
                       err, odeint_type, nthreads, steps);
  if (dim == 2)
    return run_dxdv<2>(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol, result,
                       err, odeint_type, nthreads, steps);
  return -1;
}

// Op-counting integrate_dxdv (single thread), out[0..5] = add, mul, div, sqrt,
// special, orbit-steps, summed over the orbits.  For the symplectic methods
// the count includes the plain step estimate that the device runs.
ORC_EXPORT int oracle_count_ops_dxdv(int nobj, int dim, double* yo, int nt, double* t, int npot,
                                     int* pot_type, double* pot_args, double dt, double rtol,
                                     double atol, double* result, int* err, int odeint_type,
                                     double* out) {
  if (dim != 2 && dim != 3) return -1;
  const bool planar = dim == 2;
  Component P[kMaxPot];
  if (npot > kMaxPot || parse_potential(npot, pot_type, pot_args, P, planar) != 0 ||
      !has_hessian(npot, P))
    return -1;
  const int N = 4 * dim;
  g_ops = OpCount{0, 0, 0, 0, 0};
  long long steps = 0;
  for (int ii = 0; ii < nobj; ii++) {
    IntStats st = {0, 0, 0, 0.};
    if (dim == 3)
      dxdv_one<3, CountedD>(npot, P, odeint_type, yo + (long)N * ii, nt, t, dt, rtol, atol,
                            result + (long)N * nt * ii, err + ii, &st);
    else
      dxdv_one<2, CountedD>(npot, P, odeint_type, yo + (long)N * ii, nt, t, dt, rtol, atol,
                            result + (long)N * nt * ii, err + ii, &st);
    steps += st.steps;
  }
  out[0] = (double)g_ops.add;
  out[1] = (double)g_ops.mul;
  out[2] = (double)g_ops.div;
  out[3] = (double)g_ops.sqrt;
  out[4] = (double)g_ops.special;
  out[5] = (double)steps;
  return 0;
}

// e^x E1(x) and 2F1(A, B; A+1; -x) at n points (their own pins against
// mpmath and, for 2F1, against the reference's xsf build)
ORC_EXPORT int oracle_e1_scaled(int n, const double* x, double* out) {
  for (int i = 0; i < n; i++) out[i] = e1_scaled(x[i]);
  return 0;
}
ORC_EXPORT int oracle_hyp2f1_ap1(int n, double A, double B, double amb, const double* x,
                                 double* out) {
  for (int i = 0; i < n; i++) out[i] = hyp2f1_ap1(A, B, amb, x[i]);
  return 0;
}
