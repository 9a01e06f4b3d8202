// Integrator kernels of galpy's eight C integrators (leapfrog, RK4, RK6,
// symplec4, symplec6, DOPR54, DOP853, IAS15) and their launch logic, as a
// template over the force composition Spec and the orbit kind V.  Each
// translation unit that includes this header instantiates launch_spec for
// its own (Spec, V) pairs, so the instances compile in parallel.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "gh_internal.h"
#include "dop853_dev.h"
#include "dxdv_dev.h"
#include "rk_dev.h"
#include "ias15_dev.h"

namespace gh {
namespace {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// leapfrog (bovy_symplecticode.c:132-224) and symplec4 (:250-377), for full
// (V = GhVec3) or planar (V = GhVec2) orbits; with V = GhVec6 / GhVec4 the
// variational steppers of integrate_dxdv (leapfrog_dxdv, symplec4_dxdv,
// bovy_symplecticode.c:1053-1268: the same drift/kick sequence on the orbit
// and one deviation, the step estimated from the orbit alone)
template <int METHOD, class Spec, class V>
__global__ void __launch_bounds__(kBlock)
gh_symplectic_kernel(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                     int indiv_t, double dt_in, double rtol, double atol,
                     const GhPotTable* __restrict__ T, double* __restrict__ result,
                     int* __restrict__ err, double* __restrict__ stats,
                     const volatile int* cancel) {
  const long orbit = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (orbit >= nobj) return;
  constexpr int N = GhPhase<V>::N;
  const GhForce<Spec> F{T};
  const double* tt = t + (indiv_t ? orbit * nt : 0);
  double w[N];
  {
    double v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = yo[orbit * N + k];
    GhPhase<V>::to_rect(v, w);
#pragma unroll
    for (int k = 0; k < N; k++) yo[orbit * N + k] = w[k];  // galpy converts yo in place
  }
  V q = GhPhase<V>::q(w), p = GhPhase<V>::p(w);
  double* out = result + orbit * (long)nt * N;
  GhPhase<V>::store(out, q, p);
  int e = 0;
  const double init_dt = tt[1] - tt[0];
  double dt = dt_in;
  if (dt_in == GH_DT_SENTINEL) {
    if (METHOD == 0)
      dt = gh_leapfrog_estimate_step(F, gh_base(q), gh_base(p), init_dt, tt[0], rtol, atol);
    else if (METHOD == 4)
      dt = gh_symplec6_estimate_step(F, gh_base(q), gh_base(p), init_dt, tt[0], rtol, atol);
    else
      dt = gh_symplec4_estimate_step(F, gh_base(q), gh_base(p), init_dt, tt[0], rtol, atol);
  }
  const long ndt = (long)(init_dt / dt);
  double to = tt[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    if (cancel != nullptr && (ii & 63) == 0 && *cancel) {
      e = -10;
      break;
    }
    if (METHOD == 0) {
      // leapfrog: half drift, ndt-1 fused kick+drift, kick + half drift
      V q12 = gh_drift(q, p, dt / 2.0);
      for (long jj = 0; jj < ndt - 1; jj++) {
        p = gh_drift(p, gh_accel(F, to + dt / 2.0, q12), dt);
        q12 = gh_drift(q12, p, dt);
        to = to + dt;
      }
      p = gh_drift(p, gh_accel(F, to + dt / 2.0, q12), dt);
      q = gh_drift(q12, p, dt / 2.0);
      to = to + dt;
    } else if (METHOD == 4) {
      // symplec6 (bovy_symplecticode.c:403-586): seven kicks per step, the
      // last drift of a step merged with the first of the next, (c8 + c1) dt
      q = gh_drift(q, p, gh_s6c[0] * dt);
      to += gh_s6c[0] * dt;
      const double cmerge = gh_s6c[7] + gh_s6c[0];
      for (long jj = 0; jj < ndt; jj++) {
        const bool last = jj == ndt - 1;
#pragma unroll 1
        for (int k = 0; k < 7; k++) {
          p = gh_drift(p, gh_accel(F, to, q), gh_s6d[k] * dt);
          const double cn = (k == 6) ? (last ? gh_s6c[7] : cmerge) : gh_s6c[k + 1];
          q = gh_drift(q, p, cn * dt);
          to += cn * dt;
        }
      }
    } else {
      const double c1 = GH_S4C1, c2 = GH_S4C2, c3 = GH_S4C3, c4 = GH_S4C4;
      const double d1 = GH_S4D1, d2 = GH_S4D2, d3 = GH_S4D3;
      q = gh_drift(q, p, c1 * dt);
      to += c1 * dt;
      for (long jj = 0; jj < ndt - 1; jj++) {
        p = gh_drift(p, gh_accel(F, to, q), d1 * dt);
        q = gh_drift(q, p, c2 * dt);
        to += c2 * dt;
        p = gh_drift(p, gh_accel(F, to, q), d2 * dt);
        q = gh_drift(q, p, c3 * dt);
        to += c3 * dt;
        p = gh_drift(p, gh_accel(F, to, q), d3 * dt);
        q = gh_drift(q, p, (c4 + c1) * dt);
        to += (c4 + c1) * dt;
      }
      p = gh_drift(p, gh_accel(F, to, q), d1 * dt);
      q = gh_drift(q, p, c2 * dt);
      to += c2 * dt;
      p = gh_drift(p, gh_accel(F, to, q), d2 * dt);
      q = gh_drift(q, p, c3 * dt);
      to += c3 * dt;
      p = gh_drift(p, gh_accel(F, to, q), d3 * dt);
      q = gh_drift(q, p, c4 * dt);
      to += c4 * dt;
    }
    GhPhase<V>::store(out + (long)N * (ii + 1), q, p);
  }
  err[orbit] = e;
  if (stats) {
    stats[orbit * 3 + 0] = (double)(ndt * (long)(nt - 1));
    stats[orbit * 3 + 1] = 0.0;
    stats[orbit * 3 + 2] = dt;
  }
}

// bovy_rk4 / bovy_rk6 (bovy_rk.c:66-140, 171-268), one orbit per lane:
// ndt = (long)(init_dt / dt) steps per output interval on the rectangular
// state, rows stored as the symplectic kernel stores them.
template <int ORDER, class Spec, class V>
__global__ void __launch_bounds__(kBlock)
gh_rk_kernel(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t, int indiv_t,
             double dt_in, double rtol, double atol, const GhPotTable* __restrict__ T,
             double* __restrict__ result, int* __restrict__ err, double* __restrict__ stats,
             const volatile int* cancel) {
  const long orbit = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (orbit >= nobj) return;
  constexpr int N = GhPhase<V>::N;
  const GhForce<Spec> F{T};
  const double* tt = t + (indiv_t ? orbit * nt : 0);
  double yn[N], yn1[N];
  {
    double v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = yo[orbit * N + k];
    GhPhase<V>::to_rect(v, yn);
#pragma unroll
    for (int k = 0; k < N; k++) {
      yo[orbit * N + k] = yn[k];  // galpy converts yo in place
      yn1[k] = yn[k];
    }
  }
  double* out = result + orbit * (long)nt * N;
  GhPhase<V>::store(out, GhPhase<V>::q(yn), GhPhase<V>::p(yn));
  int e = 0;
  const double init_dt = tt[1] - tt[0];
  double dt = dt_in;
  if (dt_in == GH_DT_SENTINEL)
    dt = gh_rk_estimate_step<ORDER, GhForce<Spec>, V>(F, yn, init_dt, tt[0], rtol, atol);
  const long ndt = (long)(init_dt / dt);
  double to = tt[0];
  for (int ii = 0; ii < nt - 1; ii++) {
    if (cancel != nullptr && (ii & 63) == 0 && *cancel) {
      e = -10;
      break;
    }
    for (long jj = 0; jj < ndt; jj++) {
      gh_rk_onestep<ORDER, GhForce<Spec>, V>(F, yn, yn1, to, dt);
      to += dt;
#pragma unroll
      for (int k = 0; k < N; k++) yn[k] = yn1[k];
    }
    GhPhase<V>::store(out + (long)N * (ii + 1), GhPhase<V>::q(yn1), GhPhase<V>::p(yn1));
  }
  err[orbit] = e;
  if (stats) {
    stats[orbit * 3 + 0] = (double)(ndt * (long)(nt - 1));
    stats[orbit * 3 + 1] = 0.0;
    stats[orbit * 3 + 2] = dt;
  }
}

// bovy_dopr54 / bovy_dopr54_onestep (bovy_rk.c:501-616), one orbit per lane:
// the first step from the RK4 estimator; each output interval (of the grid's
// first spacing, as the reference) is stepped to its end with the last step
// clipped onto it; a step reduced more than 10000-fold from the interval's
// first one (or NaN) is forced through at that floor and makes err odd.
// Every pass of the loop either advances t or shrinks the step towards that
// floor, so the loop ends; it is also capped at kDopr54MaxIter passes per
// output interval (err = -3 then), so that no input can keep a lane busy.
constexpr long kDopr54MaxIter = 1L << 22;

template <class Spec, class V>
__global__ void __launch_bounds__(kBlock)
gh_dopr54_kernel(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                 int indiv_t, double dt_in, double rtol, double atol,
                 const GhPotTable* __restrict__ T, double* __restrict__ result,
                 int* __restrict__ err, double* __restrict__ stats, const volatile int* cancel) {
  const long orbit = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (orbit >= nobj) return;
  constexpr int N = GhPhase<V>::N;
  const GhForce<Spec> F{T};
  const double* tt = t + (indiv_t ? orbit * nt : 0);
  double yn[N];
  {
    double v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = yo[orbit * N + k];
    GhPhase<V>::to_rect(v, yn);
#pragma unroll
    for (int k = 0; k < N; k++) yo[orbit * N + k] = yn[k];  // galpy converts yo in place
  }
  double* out = result + orbit * (long)nt * N;
  GhPhase<V>::store(out, GhPhase<V>::q(yn), GhPhase<V>::p(yn));
  const double dt = tt[1] - tt[0];
  double dt_one = dt_in;
  if (dt_in == GH_DT_SENTINEL)
    dt_one = gh_rk_estimate_step<4, GhForce<Spec>, V>(F, yn, dt, tt[0], rtol, atol);
  double to = tt[0];
  double a1[N];
  gh_rk_deriv<GhForce<Spec>, V>(F, to, yn, a1);
  int e = 0;
  long nacc = 0, nrej = 0;
  for (int ii = 0; ii < nt - 1 && e != -3; ii++) {
    if (cancel != nullptr && (ii & 63) == 0 && *cancel) {
      e = -10;
      break;
    }
    const double init_dt_one = dt_one, init_to = to;
    long iter = 0;
    while ((dt >= 0.0 && to < init_to + dt) || (dt < 0.0 && to > init_to + dt)) {
      if (++iter > kDopr54MaxIter) {
        e = -3;
        break;
      }
      bool accept = false;
      if (init_dt_one / dt_one > 10000.0 || dt_one != dt_one) {
        dt_one = init_dt_one / 10000.0;
        accept = true;
        if (e % 2 == 0) e += 1;
      }
      if (dt >= 0.0 && dt_one > init_to + dt - to) dt_one = init_to + dt - to;
      if (dt < 0.0 && dt_one < init_to + dt - to) dt_one = init_to + dt - to;
      bool accepted;
      dt_one = gh_dopr54_actualstep<GhForce<Spec>, V>(F, yn, dt_one, to, rtol, atol, a1, accept,
                                                       accepted);
      if (accepted)
        nacc++;
      else
        nrej++;
    }
    GhPhase<V>::store(out + (long)N * (ii + 1), GhPhase<V>::q(yn), GhPhase<V>::p(yn));
  }
  err[orbit] = e;
  if (stats) {
    stats[orbit * 3 + 0] = (double)nacc;
    stats[orbit * 3 + 1] = (double)nrej;
    stats[orbit * 3 + 2] = dt_one;
  }
}

// The same integration with persistent lanes: a lane that finishes its
// orbit claims the next one from a counter (zeroed before the launch; every
// lane claims its first orbit there too), so the lanes of a wavefront keep
// stepping until the queue is empty instead of idling behind the slowest
// orbit of the launch, and a workgroup that starts late finds the queue
// empty and leaves at once.
// Each pass of the loop is one pass of the reference's inner loop (at most
// one step), one output row or one interval's set-up; per orbit the steps
// and their order are those of gh_dopr54_kernel.  A lane leaves once no lane of its wavefront holds an orbit or
// a claim below nobj, which every lane reaches: each orbit ends after at
// most nt - 1 intervals of at most kDopr54MaxIter passes.  `order`
// (nullable) permutes the claims: claim k integrates orbit order[k].  A
// PILOT launch writes nothing but cost[o] = step attempts (yo untouched).
template <class Spec, class V, bool PILOT>
__global__ void __launch_bounds__(kBlock)
gh_dopr54_persistent(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                     int indiv_t, double dt_in, double rtol, double atol,
                     const GhPotTable* __restrict__ T, double* __restrict__ result,
                     int* __restrict__ err, double* __restrict__ stats, const volatile int* cancel,
                     unsigned long long* __restrict__ counter, const int* __restrict__ order,
                     unsigned* __restrict__ cost) {
  constexpr int N = GhPhase<V>::N;
  const GhForce<Spec> F{T};
  long next = (long)atomicAdd(counter, 1ULL);  // the orbit to start next
  long orbit = 0;
  double yn[N], a1[N];
  const double* tt = t;
  double* out = result;
  double dt = 0.0, dt_one = 0.0, to = 0.0, init_dt_one = 0.0, init_to = 0.0;
  long iter = 0, nacc = 0, nrej = 0;
  int ii = 0, e = 0;
  bool active = false, fresh = false;  // fresh: interval ii not yet set up
  while (true) {
    if (!active && next < nobj) {
      orbit = order != nullptr ? (long)order[next] : next;
      tt = t + (indiv_t ? orbit * nt : 0);
      double v[N];
#pragma unroll
      for (int k = 0; k < N; k++) v[k] = yo[orbit * N + k];
      GhPhase<V>::to_rect(v, yn);
      if (!PILOT) {
#pragma unroll
        for (int k = 0; k < N; k++) yo[orbit * N + k] = yn[k];  // galpy converts yo in place
        out = result + orbit * (long)nt * N;
        GhPhase<V>::store(out, GhPhase<V>::q(yn), GhPhase<V>::p(yn));
      }
      dt = tt[1] - tt[0];
      dt_one = dt_in;
      if (dt_in == GH_DT_SENTINEL)
        dt_one = gh_rk_estimate_step<4, GhForce<Spec>, V>(F, yn, dt, tt[0], rtol, atol);
      to = tt[0];
      gh_rk_deriv<GhForce<Spec>, V>(F, to, yn, a1);
      e = 0;
      nacc = nrej = 0;
      ii = 0;
      fresh = true;
      active = true;
    }
    if (active) {
      bool done = false;
      if (fresh) {
        if (ii >= nt - 1 || e == -3) {
          done = true;
        } else if (cancel != nullptr && (ii & 63) == 0 && *cancel) {
          e = -10;
          done = true;
        } else {
          init_dt_one = dt_one;
          init_to = to;
          iter = 0;
          fresh = false;
        }
      } else {
        bool end = !((dt >= 0.0 && to < init_to + dt) || (dt < 0.0 && to > init_to + dt));
        if (!end && ++iter > kDopr54MaxIter) {
          e = -3;
          end = true;
        }
        if (end) {
          if (!PILOT)
            GhPhase<V>::store(out + (long)N * (ii + 1), GhPhase<V>::q(yn), GhPhase<V>::p(yn));
          ii++;
          fresh = true;
        } else {
          bool accept = false;
          if (init_dt_one / dt_one > 10000.0 || dt_one != dt_one) {
            dt_one = init_dt_one / 10000.0;
            accept = true;
            if (e % 2 == 0) e += 1;
          }
          if (dt >= 0.0 && dt_one > init_to + dt - to) dt_one = init_to + dt - to;
          if (dt < 0.0 && dt_one < init_to + dt - to) dt_one = init_to + dt - to;
          bool accepted;
          dt_one = gh_dopr54_actualstep<GhForce<Spec>, V>(F, yn, dt_one, to, rtol, atol, a1,
                                                           accept, accepted);
          if (accepted)
            nacc++;
          else
            nrej++;
        }
      }
      if (done) {
        if (PILOT) {
          cost[orbit] = nacc + nrej > 0xffffffffL ? 0xffffffffu : (unsigned)(nacc + nrej);
        } else {
          err[orbit] = e;
          if (stats) {
            stats[orbit * 3 + 0] = (double)nacc;
            stats[orbit * 3 + 1] = (double)nrej;
            stats[orbit * 3 + 2] = dt_one;
          }
        }
        active = false;
        next = (long)atomicAdd(counter, 1ULL);
      }
    }
    if (!__any(active || next < nobj)) break;
  }
}

// wez_ias15 (wez_ias15.c:137-425), one orbit at a time per lane.  The first step is
// the grid's first spacing (or the caller's dt); the integration runs until
// the accepted steps add up to |nt dt|, one step length past the last output
// time as in the reference; outputs come from the accepted step's
// interpolating polynomial at each grid time it passes.  Rows the run never
// reaches (only when the caller's dt is shorter than the grid spacing:
// undefined in the reference, whose result array is uninitialised) are set
// to NaN.  Every force of a step is evaluated at the step's end time, as the
// reference does.  A run is capped at kIas15MaxAttempts step attempts
// (err = -3 then), a bound the reference does not have, so that no input
// keeps a lane busy.
constexpr long kIas15MaxAttempts = 1L << 22;

// The output row `row` of an IAS15 orbit from its position q and velocity p.
template <class V, int D>
__device__ __forceinline__ void gh_ias15_put(double* out, int row, const double* q,
                                             const double* p) {
  constexpr int N = GhPhase<V>::N;
  double y[N];
#pragma unroll
  for (int i = 0; i < D; i++) {
    y[i] = q[i];
    y[D + i] = p[i];
  }
  GhPhase<V>::store(out + (long)N * row, GhPhase<V>::q(y), GhPhase<V>::p(y));
}

// The IAS15 integration on persistent lanes, as gh_dopr54_persistent: every
// lane claims its orbits from a counter zeroed before the launch and makes
// one step attempt (with the rows it passes) per pass of the loop.  Each
// orbit ends after at most kIas15MaxAttempts attempts, so every lane leaves.
// `order` and PILOT as for gh_dopr54_persistent.
template <class Spec, class V, bool PILOT>
__global__ void __launch_bounds__(kBlock)
gh_ias15_persistent(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                    int indiv_t, double dt_in, const GhPotTable* __restrict__ T,
                    double* __restrict__ result, int* __restrict__ err,
                    double* __restrict__ stats, const volatile int* cancel,
                    unsigned long long* __restrict__ counter, const int* __restrict__ order,
                    unsigned* __restrict__ cost) {
  constexpr int N = GhPhase<V>::N;
  constexpr int D = N / 2;
  const GhForce<Spec> F{T};
  long next = (long)atomicAdd(counter, 1ULL);  // the orbit to start next
  long orbit = 0;
  const double* tt = t;
  double* out = result;
  double x[D], v[D], B[7][D], E[7][D], G[7][D], F0[D];
  double init_dt = 0.0, dt = 0.0, to = 0.0, time_remaining = 0.0;
  int row = 1, e = 0;
  long nacc = 0, nrej = 0;
  bool active = false;
  while (true) {
    if (!active && next < nobj) {
      orbit = order != nullptr ? (long)order[next] : next;
      tt = t + (indiv_t ? orbit * nt : 0);
      {
        double w[N], y[N];
#pragma unroll
        for (int k = 0; k < N; k++) w[k] = yo[orbit * N + k];
        GhPhase<V>::to_rect(w, y);
        if (!PILOT) {
#pragma unroll
          for (int k = 0; k < N; k++) yo[orbit * N + k] = y[k];  // galpy converts yo in place
        }
#pragma unroll
        for (int i = 0; i < D; i++) {
          x[i] = y[i];
          v[i] = y[D + i];
        }
      }
      if (!PILOT) {
        out = result + orbit * (long)nt * N;
        gh_ias15_put<V, D>(out, 0, x, v);
      }
#pragma unroll
      for (int k = 0; k < 7; k++)
#pragma unroll
        for (int i = 0; i < D; i++) B[k][i] = E[k][i] = 0.0;
      init_dt = tt[1] - tt[0];
      dt = dt_in == GH_DT_SENTINEL ? init_dt : dt_in;
      to = tt[0];
      time_remaining = fabs((double)nt * dt);
      row = 1;
      e = 0;
      nacc = nrej = 0;
      active = true;
    }
    if (active) {
      bool done = !(time_remaining > 0.0);
      if (!done && ((nacc + nrej) & 63) == 0) {
        if (cancel != nullptr && *cancel) {
          e = -10;
          done = true;
        } else if (nacc + nrej >= kIas15MaxAttempts) {
          e = -3;
          done = true;
        }
      }
      if (!done) {
        const double dt_temp = dt;
        const double to_temp = to + dt_temp;
        const double dt_required =
            gh_ias15_attempt<GhForce<Spec>, V, D>(F, x, v, dt_temp, to_temp, F0, B, G);
        if (fabs(dt_temp) > fabs(dt_required)) {
          dt = dt_required;
          nrej++;
        } else {
          nacc++;
          time_remaining -= fabs(dt);
          while (row < nt && (init_dt > 0.0 ? (to < tt[row] && tt[row] <= to_temp)
                                            : (to > tt[row] && tt[row] >= to_temp))) {
            const double hs = fabs(tt[row] - to) / fabs(to_temp - to);
            double xs[D], vs[D];
            gh_ias15_position<D>(xs, x, v, hs, dt_temp, F0, B);
            gh_ias15_velocity<D>(vs, v, hs, dt_temp, F0, B);
            if (!PILOT) gh_ias15_put<V, D>(out, row, xs, vs);
            row++;
          }
          gh_ias15_advance<D>(x, v, dt_temp, F0, B, E);
          to = to_temp;
          dt = dt_required;
        }
      } else {
        if (PILOT) {
          cost[orbit] = nacc + nrej > 0xffffffffL ? 0xffffffffu : (unsigned)(nacc + nrej);
        } else {
          if (e == 0) {
            double nan[D];
#pragma unroll
            for (int i = 0; i < D; i++) nan[i] = __builtin_nan("");
            for (; row < nt; row++) gh_ias15_put<V, D>(out, row, nan, nan);
          }
          err[orbit] = e;
          if (stats) {
            stats[orbit * 3 + 0] = (double)nacc;
            stats[orbit * 3 + 1] = (double)nrej;
            stats[orbit * 3 + 2] = dt;
          }
        }
        active = false;
        next = (long)atomicAdd(counter, 1ULL);
      }
    }
    if (!__any(active || next < nobj)) break;
  }
}

// dop853 (leung_dop853.c:231-517), persistent lanes.  The grid holds only as
// many wavefronts as the GPU keeps resident; lane g starts on orbit g and,
// whenever its orbit finishes, claims the next unstarted orbit from a global
// counter (one atomic per wavefront per refill, ranks from the ballot mask).
// Orbits differ ~25x in accepted-step count at fixed tolerance, so without
// refill a wavefront runs as long as its slowest lane.
//
// `order` (nullable) permutes the claim sequence: claim k integrates orbit
// order[k].  With the orbits sorted by expected cost, longest first, the
// lanes finish together instead of waiting on a few long orbits started
// last (list scheduling; the expected cost comes from a PILOT launch of the
// same kernel over the first few percent of the time range, which writes
// only cost[o] = attempted steps and no output rows).  Every orbit is
// integrated by the same code whatever the order, so results do not depend
// on it.
// Built for 1 and 2 waves per SIMD (WPE).  At 2 waves the kernel fits 256
// registers per lane without spilling (239 VGPRs for MWPotential2014) and the
// second wave hides the first one's FP64 and LDS latency: on MI355X it is
// faster at every size measured (131,072 to 1,048,576 orbits: 1.3-1.6x), so
// it is the default; GALPY_HIP_DOP_WAVES=1 selects the other instance.
// Deferred dense output.  A lane whose accepted step passes output times
// stores the step's GhDenseJobT in its wavefront's LDS queue and steps on;
// once the queue holds kDenseFlush jobs the whole wavefront computes them,
// one job per lane.  Done in line, the three extra force evaluations of the
// dense output run for nearly every wavefront iteration (some lane almost
// always needs them) with ~1 lane in 8 active.  The queue is a structure of
// arrays (field k of slot j at k * kDenseCap + j, so a wavefront's stores hit
// consecutive banks); 47 doubles + 2 ints per job, 52 jobs = 19,968 bytes per
// wavefront (13,728 for planar orbits), so 8 single-wavefront workgroups
// still fit a CU's 160 KB LDS.
constexpr int kDenseCap = 52;
// Flush at 32 queued jobs: a step adds ~7.5 jobs per wavefront on average
// (MWPotential2014, 1001 outputs), so the 20 free slots almost never
// overflow, and on MI355X 32 measured fastest (817 ms for 1M orbits, against
// 822 at 36, 864 at 44 -- where overflow is frequent -- and 830 at 24).
constexpr int kDenseFlush = 32;
// Doubles per queued job: q0, p0, qn, pn and 11 accelerations (15 vectors of
// D components), then t_old and h -- 47 for full orbits, 32 for planar ones.
template <class V>
constexpr int kJobDoubles = 15 * (GhPhase<V>::N / 2) + 2;

__device__ __forceinline__ void gh_put(double* v, GhVec3 a) {
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
}
__device__ __forceinline__ void gh_put(double* v, GhVec2 a) {
  v[0] = a.x;
  v[1] = a.y;
}
__device__ __forceinline__ void gh_put(double* v, GhVec6 a) {
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.dx;
  v[4] = a.dy;
  v[5] = a.dz;
}
__device__ __forceinline__ void gh_put(double* v, GhVec4 a) {
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.dx;
  v[3] = a.dy;
}

// the job as kJobDoubles<V> doubles, in queue field order
template <class V>
__device__ __forceinline__ void gh_job_pack(const GhDenseJobT<V>& J, double* v) {
  constexpr int D = GhPhase<V>::N / 2;
  gh_put(v, J.q0);
  gh_put(v + D, J.p0);
  gh_put(v + 2 * D, J.qn);
  gh_put(v + 3 * D, J.pn);
#pragma unroll
  for (int k = 0; k < 11; k++) gh_put(v + 4 * D + D * k, J.A[k]);
  v[15 * D] = J.t_old;
  v[15 * D + 1] = J.h;
}

template <class V>
__device__ __forceinline__ void gh_job_store(double* __restrict__ qd, int* __restrict__ qi,
                                             int slot, const GhDenseJobT<V>& J, int orbit) {
  constexpr int cap = kDenseCap;
  double v[kJobDoubles<V>];
  gh_job_pack(J, v);
#pragma unroll
  for (int k = 0; k < kJobDoubles<V>; k++) qd[k * cap + slot] = v[k];
  qi[slot] = J.done0;
  qi[cap + slot] = orbit;
}

// Jobs that did not fit the LDS queue wait in a per-wavefront area of device
// memory (kOverflowCap slots) and move into the queue after a flush.  Each
// overflow job is contiguous (its doubles, then done0 and the orbit as two
// ints), so one address plus immediate offsets reaches all its fields;
// volatile so the copy reads see the wavefront's own earlier stores.
constexpr int kOverflowCap = 64;
template <class V>
constexpr int kJobWords = kJobDoubles<V> + 1;  // doubles per overflow slot
template <class V>
constexpr size_t kOverflowBytes = (size_t)kOverflowCap * kJobWords<V> * 8;

template <class V>
__device__ __forceinline__ void gh_job_store_ovf(volatile double* od, int slot,
                                                 const GhDenseJobT<V>& J, int orbit) {
  volatile double* p = od + (long)slot * kJobWords<V>;
  double v[kJobDoubles<V>];
  gh_job_pack(J, v);
#pragma unroll
  for (int k = 0; k < kJobDoubles<V>; k++) p[k] = v[k];
  volatile int* pi = reinterpret_cast<volatile int*>(p + kJobDoubles<V>);
  pi[0] = J.done0;
  pi[1] = orbit;
}

template <class V>
__device__ __forceinline__ void gh_job_copy_in(const volatile double* od, int from,
                                               double* __restrict__ qd, int* __restrict__ qi,
                                               int slot) {
  const volatile double* p = od + (long)from * kJobWords<V>;
#pragma unroll 8
  for (int k = 0; k < kJobDoubles<V>; k++) qd[k * kDenseCap + slot] = p[k];
  const volatile int* pi = reinterpret_cast<const volatile int*>(p + kJobDoubles<V>);
  qi[slot] = pi[0];
  qi[kDenseCap + slot] = pi[1];
}

// vector at queue field k of job `slot`
__device__ __forceinline__ void gh_job_vec(const double* __restrict__ qd, int slot, int k,
                                           GhVec3& v) {
  v = GhVec3{qd[k * kDenseCap + slot], qd[(k + 1) * kDenseCap + slot],
             qd[(k + 2) * kDenseCap + slot]};
}
__device__ __forceinline__ void gh_job_vec(const double* __restrict__ qd, int slot, int k,
                                           GhVec2& v) {
  v = GhVec2{qd[k * kDenseCap + slot], qd[(k + 1) * kDenseCap + slot]};
}
__device__ __forceinline__ void gh_job_vec(const double* __restrict__ qd, int slot, int k,
                                           GhVec6& v) {
  v = GhVec6{qd[k * kDenseCap + slot],       qd[(k + 1) * kDenseCap + slot],
             qd[(k + 2) * kDenseCap + slot], qd[(k + 3) * kDenseCap + slot],
             qd[(k + 4) * kDenseCap + slot], qd[(k + 5) * kDenseCap + slot]};
}
__device__ __forceinline__ void gh_job_vec(const double* __restrict__ qd, int slot, int k,
                                           GhVec4& v) {
  v = GhVec4{qd[k * kDenseCap + slot], qd[(k + 1) * kDenseCap + slot],
             qd[(k + 2) * kDenseCap + slot], qd[(k + 3) * kDenseCap + slot]};
}
template <class V>
__device__ __forceinline__ V gh_job_vec(const double* __restrict__ qd, int slot, int k) {
  V v;
  gh_job_vec(qd, slot, k, v);
  return v;
}

// Loads the accelerations A1, A4..A13 of job `slot` into A[1], A[4..13].
template <class V>
__device__ __forceinline__ void gh_job_load_acc(const double* __restrict__ qd, int slot, V* A) {
  constexpr int D = GhPhase<V>::N / 2;
  A[1] = gh_job_vec<V>(qd, slot, 4 * D);
#pragma unroll
  for (int j = 4; j <= 13; j++) A[j] = gh_job_vec<V>(qd, slot, 4 * D + D * (j - 3));
  A[2] = A[3] = V{};  // unused by stages 14-16 and the extension
}

// Dense output of queued job `slot` (gh_dop_dense's arithmetic), in phases
// that each re-read the job from LDS: stage 14, stage 15, stage 16, then the
// extension at every requested time.  The signal fences keep the phases'
// loads apart, so only the three new accelerations are live from one phase
// to the next -- a whole job plus a force evaluation would not fit the
// register budget of two wavefronts per SIMD next to the lane's own state.
template <class F, class V>
__device__ __forceinline__ void gh_job_run(const F& T, const double* __restrict__ qd,
                                           const int* __restrict__ qi, int slot,
                                           const double* __restrict__ t0, int nt, int indiv_t,
                                           double* __restrict__ result) {
  using namespace ghdp;
  constexpr int D = GhPhase<V>::N / 2, N = GhPhase<V>::N;
  const double t_old = qd[15 * D * kDenseCap + slot], h = qd[(15 * D + 1) * kDenseCap + slot];
  V A14, A15, A16;
  {
    V A[17];
    gh_job_load_acc(qd, slot, A);
    A14 = gh_stage_acc(T, t_old + c14 * h,
                       gh_rkn_q14(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                                  gh_job_vec<V>(qd, slot, 3 * D), A),
                       h, gh_job_vec<V>(qd, slot, D), a141, A[1], a147, A[7], a148, A[8], a149, A[9], a1410, A[10], a1411, A[11], a1412, A[12], a1413, A[13]);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  {
    V A[17];
    gh_job_load_acc(qd, slot, A);
    A[14] = A14;
    A15 = gh_stage_acc(T, t_old + c15 * h,
                       gh_rkn_q15(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                                  gh_job_vec<V>(qd, slot, 3 * D), A),
                       h, gh_job_vec<V>(qd, slot, D), a151, A[1], a156, A[6], a157, A[7], a158, A[8], a1511, A[11], a1512, A[12], a1513, A[13], a1514, A[14]);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  {
    V A[17];
    gh_job_load_acc(qd, slot, A);
    A[14] = A14;
    A[15] = A15;
    A16 = gh_stage_acc(T, t_old + c16 * h,
                       gh_rkn_q16(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                                  gh_job_vec<V>(qd, slot, 3 * D), A),
                       h, gh_job_vec<V>(qd, slot, D), a161, A[1], a166, A[6], a167, A[7], a168, A[8], a169, A[9], a1613, A[13], a1614, A[14], a1615, A[15]);
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  const int o = qi[kDenseCap + slot];
  const double* __restrict__ t = t0 + (indiv_t ? (long)o * nt : 0);
  double* __restrict__ out = result + (long)o * nt * N;
  const double pos_neg = (h > 0.0) ? 1.0 : -1.0;
  const double t_cur = t_old + h;
  int done = qi[slot];
  while (done < nt - 1 && pos_neg * t[done + 1] < pos_neg * t_cur) {
    const double sd = (t[done + 1] - t_old) / h;
    const double s1 = 1.0 - sd;
    double yo[N];
    {
      V A[17];
      gh_job_load_acc(qd, slot, A);
      A[14] = A14;
      A[15] = A15;
      A[16] = A16;
      gh_rkn_dense_eval(h, sd, s1, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                        gh_job_vec<V>(qd, slot, 2 * D), gh_job_vec<V>(qd, slot, 3 * D), A, yo);
    }
    done++;
    GhPhase<V>::store(out + (long)N * done, gh_vec_at<V>(yo), gh_vec_at<V>(yo + N / 2));
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
}

template <class Spec, bool PILOT, int WPE, class V>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
gh_dop853_kernel(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                 int indiv_t, double rtol_log, double atol_log, const GhPotTable* __restrict__ T,
                 double* __restrict__ result, int* __restrict__ err, double* __restrict__ stats,
                 const volatile int* cancel, unsigned long long* __restrict__ counter,
                 const int* __restrict__ order, unsigned* __restrict__ cost,
                 char* __restrict__ overflow) {
  const GhForce<Spec> F{T};
  const double rtol = exp(rtol_log);
  const double atol = exp(atol_log);
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  long claim = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long orbit = 0;
  constexpr int N = GhPhase<V>::N;
  GhDopStateT<V> s;
  __shared__ double qd[PILOT ? 1 : kJobDoubles<V> * kDenseCap];
  __shared__ int qi[PILOT ? 1 : 2 * kDenseCap];
  int queued = 0;  // jobs in this wavefront's queue (the same in every lane)
  volatile double* od =
      PILOT ? nullptr
            : reinterpret_cast<volatile double*>(overflow + blockIdx.x * kOverflowBytes<V>);
  auto flush = [&]() {
    __syncthreads();  // the queue's stores are visible to every lane
    if (lane < queued) gh_job_run<GhForce<Spec>, V>(F, qd, qi, lane, t, nt, indiv_t, result);
    __syncthreads();  // every job is read before slots are reused
    queued = 0;
  };
  const double* tt = t;
  double* out = result;
  bool active = false, exhausted = false, stop = false;

  auto start = [&](long o) {
    tt = t + (indiv_t ? o * nt : 0);
    out = PILOT ? nullptr : result + o * (long)nt * N;
    double v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = yo[o * N + k];
    GhPhase<V>::to_rect(v, s.y);
    gh_dop_init<GhForce<Spec>, !PILOT>(F, tt, nt, rtol, atol, s, out);
  };
  auto finish = [&](long o, int e) {
    if (PILOT) {
      cost[o] = (unsigned)(s.nacc + s.nrej);
      return;
    }
#pragma unroll
    for (int k = 0; k < N; k++) yo[o * N + k] = s.y[k];  // galpy's y0 holds the last state
    err[o] = e;
    if (stats) {
      stats[o * 3 + 0] = (double)s.nacc;
      stats[o * 3 + 1] = (double)s.nrej;
      stats[o * 3 + 2] = 0.0;
    }
  };

  if (claim < nobj) {
    orbit = order ? (long)order[claim] : claim;
    start(orbit);
    active = true;
  }
  // The stepping loop never flushes the queue itself: it runs until the
  // queue is due (or jobs overflowed, or the wavefront is done), and the
  // flush runs between stepping phases, so the stage loop's register
  // allocation never has to make room for a job's dense output.
  unsigned iter = 0;
  int over = 0;  // jobs waiting in the overflow area
  bool all_done = false;
  for (;;) {
    for (;; iter++) {
      if (cancel != nullptr && (iter & 255u) == 0 && *cancel) stop = true;
      // the dense-output job of an accepted step that passes output times is
      // stored from inside the step, as soon as it is formed (never held in
      // registers): the lanes calling the sink are exactly those with a job,
      // so a ballot there ranks them.  A wavefront can produce up to 64 jobs
      // in one step, more than the queue has room for: those wait in the
      // overflow area.
      bool need = false;
      const int room = kDenseCap - queued;
      auto sink = [&](const GhDenseJobT<V>& J) {
        const int rank = __popcll(__ballot(1) & below);
        if (rank < room)
          gh_job_store(qd, qi, queued + rank, J, (int)orbit);
        else
          gh_job_store_ovf(od, rank - room, J, (int)orbit);
        need = true;
      };
      if (active) {
        if (stop) {
          finish(orbit, -10);
          active = false;
        } else if (gh_dop_step<GhForce<Spec>, PILOT ? GH_DENSE_SKIP : GH_DENSE_DEFER>(
                       F, tt, nt, rtol, atol, s, out, sink)) {
          finish(orbit, 0);
          active = false;
        }
      }
      if (!PILOT) {
        const int nnew = __popcll(__ballot(need));
        queued += min(nnew, room);
        over = max(nnew - room, 0);
      }
      const bool want = !active && !exhausted;
      const unsigned long long m = __ballot(want);
      if (m) {
        const int leader = __ffsll((long long)m) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (want) {
          claim = (long)(base + (unsigned long long)__popcll(m & below));
          if (claim < nobj) {
            orbit = order ? (long)order[claim] : claim;
            if (stop) {
              if (!PILOT) err[orbit] = -10;  // interrupted before it started
            } else {
              start(orbit);
              active = true;
            }
          } else {
            exhausted = true;
          }
        }
      }
      all_done = __all(!active && exhausted);
      if (all_done || over > 0 || queued >= kDenseFlush) {
        iter++;
        break;
      }
    }
    if (!PILOT) {
      // Every dense output runs through this one flush path, so results do
      // not depend on timing; overflowed jobs move in by rounds.
      for (int from = 0; over > 0;) {
        flush();
        const int take = min(over, kDenseCap);
        if (lane < take) gh_job_copy_in<V>(od, from + lane, qd, qi, lane);
        queued = take;
        from += take;
        over -= take;
      }
      if (queued >= kDenseFlush || (all_done && queued > 0)) flush();
    }
    if (all_done) break;
  }
}

// Pilot time grid {t0, t0 + frac (t_end - t0)} for the cost estimate.
__global__ void gh_pilot_grid(const double* __restrict__ t, int nt, double frac,
                              double* __restrict__ pt) {
  pt[0] = t[0];
  pt[1] = t[0] + frac * (t[nt - 1] - t[0]);
}

__global__ void gh_iota(int n, int* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void gh_set_counter(unsigned long long* c, unsigned long long v) { *c = v; }
// Pilot length for the DOP853 cost estimate, as a fraction of the time
// range: the pilot's step counts correlate with the full run's at ~0.99 for
// MWPotential2014 orbits, for ~2% of the work.
constexpr double kDopPilotFrac = 0.02;

// GALPY_HIP_DOP_ORDER=0 turns the longest-first ordering off.
bool dop_order_enabled() {
  static const int v = [] {
    const char* e = std::getenv("GALPY_HIP_DOP_ORDER");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  return v != 0;
}

// GALPY_HIP_DOP_WAVES=1 or 2 selects the DOP853 occupancy variant (default 2).
[[maybe_unused]] int dop_waves_forced() {
  static const int v = [] {
    const char* e = std::getenv("GALPY_HIP_DOP_WAVES");
    return (e && (e[0] == '1' || e[0] == '2') && e[1] == 0) ? e[0] - '0' : 0;
  }();
  return v;
}

// workgroups of gh_dop853_kernel<Spec> the device keeps resident at once
template <class Spec, int WPE, class V>
int dop_resident_blocks(int dev) {
  static int cached[64] = {0};
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gh_dop853_kernel<Spec, false, WPE, V>,
                                                   kBlock, 0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int n = std::max(1, per_cu) * std::max(1, ncu);
  if (dev >= 0 && dev < 64) cached[dev] = n;
  return n;
}
// DOP853 launch at WPE waves per SIMD: optional pilot + longest-first
// order, then the persistent kernel.  The temporaries (dense-output overflow
// area, pilot cost keys and order, the radix sort's temporary) live in one
// stream-ordered allocation freed after the launch, or, with a workspace,
// in the caller's memory (then nothing is allocated).
template <class Spec, int WPE, class V>
int launch_dop853(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                  const GhPotTable* d_table, double rtol, double atol, double* d_result,
                  int* d_err, double* d_stats, const int* cancel, int dev, hipStream_t stream,
                  const GhWorkspace* ws) {
  const int resident = dop_resident_blocks<Spec, WPE, V>(dev);
  if (resident <= 0) {
    set_error("could not query DOP853 occupancy on device %d", dev);
    return -1;
  }
  const unsigned nblocks =
      (unsigned)std::min<long>((long)((nobj + kBlock - 1) / kBlock), (long)resident);
  const unsigned long long lanes = (unsigned long long)nblocks * kBlock;
  // Layout of the temporaries: [counter 256 B, workspace only][overflow:
  // kOverflowBytes per workgroup][ordered: cost, sorted cost, iota, order
  // (4 n4), pilot grid (256 B), sort temporary].  The order only matters
  // when some lane gets more than one orbit.
  const bool ordered = !indiv_t && (unsigned long long)nobj > lanes && dop_order_enabled();
  size_t sort_bytes = 0;
  hipError_t e = hipSuccess;
  if (ordered)
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, (const unsigned*)nullptr,
                                                     (unsigned*)nullptr, (const int*)nullptr,
                                                     (int*)nullptr, nobj, 0, 32, stream);
  const size_t head = ws ? 256 : 0;
  const size_t ovf = ((size_t)nblocks * kOverflowBytes<V> + 255) / 256 * 256;
  const size_t n4 = ordered ? ((size_t)nobj * 4 + 255) / 256 * 256 : 0;
  const size_t need = head + ovf + (ordered ? 4 * n4 + 256 + sort_bytes : 0);
  unsigned long long* counter = nullptr;
  void* work = nullptr;
  if (ws) {
    if (need > ws->bytes) {
      set_error("DOP853 temporaries for %d orbits (%zu bytes: %zu overflow area for %u workgroups, "
                "%zu pilot order and sort) exceed the workspace's %zu bytes",
                nobj, need, ovf, nblocks, need - head - ovf, ws->bytes);
      return -1;
    }
    counter = reinterpret_cast<unsigned long long*>(ws->ptr);
    work = ws->ptr + head;
  } else {
    counter = next_counter(dev);
    if (counter == nullptr) {
      set_error("could not set up the DOP853 work counter on device %d", dev);
      return -1;
    }
    if (e == hipSuccess) e = hipMallocAsync(&work, need, stream);
  }
  char* d_overflow = static_cast<char*>(work);
  const int* d_order = nullptr;
  if (e == hipSuccess && ordered) {
    char* w = static_cast<char*>(work) + ovf;
    unsigned* d_cost = reinterpret_cast<unsigned*>(w);
    unsigned* d_cost_sorted = reinterpret_cast<unsigned*>(w + n4);
    int* d_iota = reinterpret_cast<int*>(w + 2 * n4);
    int* d_perm = reinterpret_cast<int*>(w + 3 * n4);
    double* d_pt = reinterpret_cast<double*>(w + 4 * n4);
    void* d_sort_tmp = w + 4 * n4 + 256;
    hipLaunchKernelGGL(gh_pilot_grid, dim3(1), dim3(1), 0, stream, d_t, nt, kDopPilotFrac, d_pt);
    hipLaunchKernelGGL(gh_set_counter, dim3(1), dim3(1), 0, stream, counter, lanes);
    hipLaunchKernelGGL((gh_dop853_kernel<Spec, true, WPE, V>), dim3(nblocks), dim3(kBlock), 0,
                       stream, nobj, d_yo, 2, d_pt, 0, rtol, atol, d_table, nullptr, nullptr,
                       nullptr, cancel, counter, nullptr, d_cost, nullptr);
    hipLaunchKernelGGL(gh_iota, dim3((unsigned)((nobj + 255) / 256)), dim3(256), 0, stream, nobj,
                       d_iota);
    e = hipcub::DeviceRadixSort::SortPairsDescending(d_sort_tmp, sort_bytes, d_cost, d_cost_sorted,
                                                     d_iota, d_perm, nobj, 0, 32, stream);
    d_order = d_perm;
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gh_set_counter, dim3(1), dim3(1), 0, stream, counter, lanes);
    hipLaunchKernelGGL((gh_dop853_kernel<Spec, false, WPE, V>), dim3(nblocks), dim3(kBlock), 0,
                       stream, nobj, d_yo, nt, d_t, indiv_t, rtol, atol, d_table, d_result, d_err,
                       d_stats, cancel, counter, d_order, nullptr, d_overflow);
    e = hipGetLastError();
  }
  if (work && !ws) {
    const hipError_t ef = hipFreeAsync(work, stream);
    if (e == hipSuccess) e = ef;
  }
  if (e != hipSuccess) {
    set_error("DOP853 launch failed: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

// One launch of the persistent DOPR54 (DOPR54 = true) or IAS15 kernel.
template <class Spec, class V, bool DOPR54, bool PILOT>
void persistent_launch(unsigned nblocks, hipStream_t stream, int nobj, double* d_yo, int nt,
                       const double* d_t, int indiv_t, const GhPotTable* d_table, double dt,
                       double rtol, double atol, double* d_result, int* d_err, double* d_stats,
                       const int* cancel, unsigned long long* counter, const int* order,
                       unsigned* cost) {
  hipLaunchKernelGGL(gh_set_counter, dim3(1), dim3(1), 0, stream, counter, 0ULL);
  if constexpr (DOPR54)
    hipLaunchKernelGGL((gh_dopr54_persistent<Spec, V, PILOT>), dim3(nblocks), dim3(kBlock), 0,
                       stream, nobj, d_yo, nt, d_t, indiv_t, dt, rtol, atol, d_table, d_result,
                       d_err, d_stats, cancel, counter, order, cost);
  else
    hipLaunchKernelGGL((gh_ias15_persistent<Spec, V, PILOT>), dim3(nblocks), dim3(kBlock), 0,
                       stream, nobj, d_yo, nt, d_t, indiv_t, dt, d_table, d_result, d_err,
                       d_stats, cancel, counter, order, cost);
}

// DOPR54 (DOPR54 = true) or IAS15 on persistent lanes: twice as many
// workgroups as the occupancy query says stay resident (or as the orbits
// fill), the work counter zeroed first; the counter is the workspace's first
// 256 B when one is given.  The query reports 8 workgroups of 64 per CU for
// the 120-register MWPotential2014 DOPR54 instance, half of what gfx950's 512
// registers per SIMD lane hold; the surplus costs nothing when the query is
// right, because a workgroup that starts after the queue has drained leaves
// at once.  With more orbits than resident lanes (one time grid, no
// workspace) the orbits are claimed longest first, as for DOP853: a PILOT
// launch over the first kDopPilotFrac of the time range counts each orbit's
// step attempts, a radix sort orders them, and the lanes that finish early
// take the short orbits at the end.
template <class Spec, class V, bool DOPR54>
int launch_persistent(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                      const GhPotTable* d_table, double dt, double rtol, double atol,
                      double* d_result, int* d_err, double* d_stats, const int* cancel, int dev,
                      hipStream_t stream, const GhWorkspace* ws) {
  const char* what = DOPR54 ? "DOPR54" : "IAS15";
  static int cached[64] = {0};
  int resident = dev >= 0 && dev < 64 ? cached[dev] : 0;
  if (resident <= 0) {
    int per_cu = 0, ncu = 0;
    hipError_t q;
    if constexpr (DOPR54)
      q = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, gh_dopr54_persistent<Spec, V, false>, kBlock, 0);
    else
      q = hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, gh_ias15_persistent<Spec, V, false>, kBlock, 0);
    if (q != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
      set_error("could not query %s occupancy on device %d", what, dev);
      return -1;
    }
    resident = std::max(1, per_cu) * std::max(1, ncu);
    if (dev >= 0 && dev < 64) cached[dev] = resident;
  }
  unsigned long long* counter = nullptr;
  if (ws) {
    if (ws->bytes < 256) {
      set_error("the %s work counter needs 256 bytes of workspace, %zu given", what, ws->bytes);
      return -1;
    }
    counter = reinterpret_cast<unsigned long long*>(ws->ptr);
  } else {
    counter = next_counter(dev);
    if (counter == nullptr) {
      set_error("could not set up the %s work counter on device %d", what, dev);
      return -1;
    }
  }
  const unsigned nblocks =
      (unsigned)std::min<long>((long)((nobj + kBlock - 1) / kBlock), 2L * resident);
  const bool ordered =
      !ws && !indiv_t && nt > 1 && (long)nobj > (long)resident * kBlock && dop_order_enabled();
  hipError_t e = hipSuccess;
  void* work = nullptr;
  const int* d_order = nullptr;
  if (ordered) {
    size_t sort_bytes = 0;
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, (const unsigned*)nullptr,
                                                     (unsigned*)nullptr, (const int*)nullptr,
                                                     (int*)nullptr, nobj, 0, 32, stream);
    const size_t n4 = ((size_t)nobj * 4 + 255) / 256 * 256;
    if (e == hipSuccess) e = hipMallocAsync(&work, 4 * n4 + 256 + sort_bytes, stream);
    if (e == hipSuccess) {
      char* w = static_cast<char*>(work);
      unsigned* d_cost = reinterpret_cast<unsigned*>(w);
      unsigned* d_cost_sorted = reinterpret_cast<unsigned*>(w + n4);
      int* d_iota = reinterpret_cast<int*>(w + 2 * n4);
      int* d_perm = reinterpret_cast<int*>(w + 3 * n4);
      double* d_pt = reinterpret_cast<double*>(w + 4 * n4);
      void* d_sort_tmp = w + 4 * n4 + 256;
      hipLaunchKernelGGL(gh_pilot_grid, dim3(1), dim3(1), 0, stream, d_t, nt, kDopPilotFrac, d_pt);
      persistent_launch<Spec, V, DOPR54, true>(nblocks, stream, nobj, d_yo, 2, d_pt, 0, d_table,
                                               dt, rtol, atol, nullptr, nullptr, nullptr, cancel,
                                               counter, nullptr, d_cost);
      hipLaunchKernelGGL(gh_iota, dim3((unsigned)((nobj + 255) / 256)), dim3(256), 0, stream,
                         nobj, d_iota);
      e = hipcub::DeviceRadixSort::SortPairsDescending(d_sort_tmp, sort_bytes, d_cost,
                                                       d_cost_sorted, d_iota, d_perm, nobj, 0, 32,
                                                       stream);
      d_order = d_perm;
    }
  }
  if (e == hipSuccess) {
    persistent_launch<Spec, V, DOPR54, false>(nblocks, stream, nobj, d_yo, nt, d_t, indiv_t,
                                              d_table, dt, rtol, atol, d_result, d_err, d_stats,
                                              cancel, counter, d_order, nullptr);
    e = hipGetLastError();
  }
  if (work) {
    const hipError_t ef = hipFreeAsync(work, stream);
    if (e == hipSuccess) e = ef;
  }
  if (e != hipSuccess) {
    set_error("%s launch failed: %s", what, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

}  // namespace

// Planar DOP853 runs at 2 waves per SIMD only (its state is two thirds of a
// full orbit's, so it never needs the single-wave instance).
template <class Spec, class V>
int launch_spec(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                const GhPotTable* d_table, double dt, double rtol, double atol, double* d_result,
                int* d_err, int odeint_type, double* d_stats, hipStream_t stream,
                const GhWorkspace* ws) {
  const int* cancel = ws ? ws->cancel : irq_device_ptr();
  const dim3 grid((unsigned)((nobj + kBlock - 1) / kBlock)), block(kBlock);
  switch (odeint_type) {
    case 0:
      hipLaunchKernelGGL((gh_symplectic_kernel<0, Spec, V>), grid, block, 0, stream, nobj, d_yo,
                         nt, d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                         cancel);
      break;
    case 3:
      hipLaunchKernelGGL((gh_symplectic_kernel<3, Spec, V>), grid, block, 0, stream, nobj, d_yo,
                         nt, d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                         cancel);
      break;
    case 1:  // RK4 and RK6: plain orbits, and the variational systems of
    case 2:  // integrate_dxdv (integrateFullOrbit.c:1272-1296)
      if (odeint_type == 1)
        hipLaunchKernelGGL((gh_rk_kernel<4, Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                           d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                           cancel);
      else
        hipLaunchKernelGGL((gh_rk_kernel<6, Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                           d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                           cancel);
      break;
    case 5: {  // DOPR54: plain orbits on persistent lanes, the variational systems as before
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        int dev = 0;
        GH_CHECK(hipGetDevice(&dev));
        return launch_persistent<Spec, V, true>(nobj, d_yo, nt, d_t, indiv_t, d_table, dt, rtol,
                                                atol, d_result, d_err, d_stats, cancel, dev,
                                                stream, ws);
      } else {
        hipLaunchKernelGGL((gh_dopr54_kernel<Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                           d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats, cancel);
      }
      break;
    }
    case 7:  // IAS15: plain orbits (galpy has no ias15 integrate_dxdv)
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        int dev = 0;
        GH_CHECK(hipGetDevice(&dev));
        return launch_persistent<Spec, V, false>(nobj, d_yo, nt, d_t, indiv_t, d_table, dt, rtol,
                                                 atol, d_result, d_err, d_stats, cancel, dev,
                                                 stream, ws);
      } else {
        set_error("ias15 has no integrate_dxdv (as in galpy, which refuses it upstream)");
        return -1;
      }
    case 4:  // symplec6: plain orbits, and symplec6_dxdv's tangent map
             // (bovy_symplecticode.c:1269-, step estimated from the orbit)
      hipLaunchKernelGGL((gh_symplectic_kernel<4, Spec, V>), grid, block, 0, stream, nobj, d_yo,
                         nt, d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                         cancel);
      break;
    case 6: {
      int dev = 0;
      GH_CHECK(hipGetDevice(&dev));
      if constexpr (std::is_same<V, GhVec3>::value) {
        const bool two = dop_waves_forced() != 1;
        return two ? launch_dop853<Spec, 2, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                               d_result, d_err, d_stats, cancel, dev, stream, ws)
                   : launch_dop853<Spec, 1, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                               d_result, d_err, d_stats, cancel, dev, stream, ws);
      } else if constexpr (std::is_same<V, GhVec2>::value) {
        return launch_dop853<Spec, 2, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                         d_result, d_err, d_stats, cancel, dev, stream, ws);
      } else {
        // variational states: twice the registers and LDS per job, one wave
        // per SIMD (the LDS queue allows four to six workgroups per CU)
        return launch_dop853<Spec, 1, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                         d_result, d_err, d_stats, cancel, dev, stream, ws);
      }
    }
    default:
      set_error("odeint_type %d has no HIP integrator (supported: 0 leapfrog, 1 rk4, 2 rk6, "
                "3 symplec4, 4 symplec6, 5 dopr54, 6 dop853, 7 ias15)", odeint_type);
      return -1;
  }
  GH_CHECK(hipGetLastError());
  return 0;
}

}  // namespace gh
