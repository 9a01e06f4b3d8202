// Device integrators: leapfrog, symplec4 (+ step estimators) and DOP853,
// one orbit per lane, FP64 state in registers.
//
// The arithmetic sequence follows galpy's C integrators
// (galpy/util/bovy_symplecticode.c:132-377 and :588-808 for the symplectic
// methods, galpy/util/leung_dop853.c:231-517 for DOP853) so that the GPU
// reproduces galpy's trajectories to rounding, including the per-orbit step
// size chosen by the estimators and the (long) truncation of init_dt/dt.
#pragma once
#include "potential_dev.h"

#define GH_DT_SENTINEL (-9999.99)  // integrateFullOrbit.py:666-667
#define GH_MAX_DT_REDUCE 10000.0   // bovy_symplecticode.c:35

// symplec4 coefficients (Kinoshita et al.), bovy_symplecticode.c:96-99
#define GH_S4C1 0.6756035959798289
#define GH_S4C2 (-0.1756035959798288)
#define GH_S4C3 (-0.1756035959798288)
#define GH_S4C4 0.6756035959798289
#define GH_S4D1 1.3512071919596578
#define GH_S4D2 (-1.7024143839193153)
#define GH_S4D3 1.3512071919596578

// Position, velocity or acceleration of a full (GhVec3) or planar (GhVec2)
// orbit.  The integrators are templates over the two; galpy runs the same
// C integrators with dim = 3 (integrateFullOrbit.c:1091-1132) or dim = 2
// (integratePlanarOrbit.c:788-825).
struct GhVec3 {
  double x, y, z;
};

struct GhVec2 {
  double x, y;
};

// Position (or velocity, acceleration) of an orbit together with one
// phase-space deviation, for the variational equations of galpy's
// integrate_dxdv (integrateFullOrbit.c:1234-1306, integratePlanarOrbit.c:940-1012):
// (x, y, z | dx, dy, dz) for full orbits, (x, y | dx, dy) for planar ones.
struct GhVec6 {
  double x, y, z, dx, dy, dz;
};

struct GhVec4 {
  double x, y, dx, dy;
};

// a vector from consecutive doubles in component order (x, y, z, dx, ...)
__device__ __forceinline__ void gh_vec_at(const double* w, GhVec3& v) { v = GhVec3{w[0], w[1], w[2]}; }
__device__ __forceinline__ void gh_vec_at(const double* w, GhVec2& v) { v = GhVec2{w[0], w[1]}; }
__device__ __forceinline__ void gh_vec_at(const double* w, GhVec6& v) {
  v = GhVec6{w[0], w[1], w[2], w[3], w[4], w[5]};
}
__device__ __forceinline__ void gh_vec_at(const double* w, GhVec4& v) {
  v = GhVec4{w[0], w[1], w[2], w[3]};
}
template <class V>
__device__ __forceinline__ V gh_vec_at(const double* w) {
  V v;
  gh_vec_at(w, v);
  return v;
}

// the orbit part of a state (the step estimators look at the orbit only,
// bovy_symplecticode.c:1081-1084)
__device__ __forceinline__ GhVec3 gh_base(GhVec3 v) { return v; }
__device__ __forceinline__ GhVec2 gh_base(GhVec2 v) { return v; }
__device__ __forceinline__ GhVec3 gh_base(GhVec6 v) { return GhVec3{v.x, v.y, v.z}; }
__device__ __forceinline__ GhVec2 gh_base(GhVec4 v) { return GhVec2{v.x, v.y}; }

__device__ __forceinline__ GhVec3 gh_drift(GhVec3 q, GhVec3 p, double h) {
  return GhVec3{q.x + h * p.x, q.y + h * p.y, q.z + h * p.z};
}

__device__ __forceinline__ GhVec2 gh_drift(GhVec2 q, GhVec2 p, double h) {
  return GhVec2{q.x + h * p.x, q.y + h * p.y};
}

__device__ __forceinline__ GhVec6 gh_drift(GhVec6 q, GhVec6 p, double h) {
  return GhVec6{q.x + h * p.x,   q.y + h * p.y,   q.z + h * p.z,
                q.dx + h * p.dx, q.dy + h * p.dy, q.dz + h * p.dz};
}

__device__ __forceinline__ GhVec4 gh_drift(GhVec4 q, GhVec4 p, double h) {
  return GhVec4{q.x + h * p.x, q.y + h * p.y, q.dx + h * p.dx, q.dy + h * p.dy};
}

template <class F>
__device__ __forceinline__ GhVec3 gh_accel(const F& T, double t, GhVec3 q) {
  GhVec3 a;
  T.rect(t, q.x, q.y, q.z, a.x, a.y, a.z);
  return a;
}

// planar acceleration (evalPlanarRectForce, integratePlanarOrbit.c:1015-1031)
template <class F>
__device__ __forceinline__ GhVec2 gh_accel(const F& T, double t, GhVec2 q) {
  GhVec2 a;
  T.rect(t, q.x, q.y, a.x, a.y);
  return a;
}

// acceleration with the velocity (evalRectDeriv, integrateFullOrbit.c:
// 1326-1351; evalPlanarRectDeriv, integratePlanarOrbit.c:1032-1055): the
// first-order integrators hand it to velocity-dependent forces
template <class F>
__device__ __forceinline__ GhVec3 gh_accel_v(const F& T, double t, GhVec3 q, GhVec3 p) {
  GhVec3 a;
  T.rect_v(t, q.x, q.y, q.z, p.x, p.y, p.z, a.x, a.y, a.z);
  return a;
}
template <class F>
__device__ __forceinline__ GhVec2 gh_accel_v(const F& T, double t, GhVec2 q, GhVec2 p) {
  GhVec2 a;
  T.rect_v(t, q.x, q.y, p.x, p.y, a.x, a.y);
  return a;
}
// the variational systems of integrate_dxdv: conservative forces only
template <class F, class V>
__device__ __forceinline__ V gh_accel_v(const F& T, double t, V q, V) {
  return gh_accel(T, t, q);
}

// log-space scale of the step estimators, bovy_symplecticode.c:606-623
__device__ __forceinline__ void gh_est_scales(GhVec3 q, GhVec3 p, double rtol, double atol,
                                              double& sq, double& sp) {
  const double mq = fmax(fmax(log(fabs(q.x)), log(fabs(q.y))), log(fabs(q.z)));
  const double mp = fmax(fmax(log(fabs(p.x)), log(fabs(p.y))), log(fabs(p.z)));
  double c = fmax(atol, rtol + mq);
  sq = log(exp(atol - c) + exp(rtol + mq - c)) + c;
  c = fmax(atol, rtol + mp);
  sp = log(exp(atol - c) + exp(rtol + mp - c)) + c;
}

// error norm of the estimators, bovy_symplecticode.c:641-646 / 795-800
__device__ __forceinline__ double gh_est_err(GhVec3 q11, GhVec3 q12, GhVec3 p11, GhVec3 p12,
                                             double sq, double sp) {
  double e = 0.0;
  e += exp(2.0 * log(fabs(q11.x - q12.x)) - 2.0 * sq);
  e += exp(2.0 * log(fabs(p11.x - p12.x)) - 2.0 * sp);
  e += exp(2.0 * log(fabs(q11.y - q12.y)) - 2.0 * sq);
  e += exp(2.0 * log(fabs(p11.y - p12.y)) - 2.0 * sp);
  e += exp(2.0 * log(fabs(q11.z - q12.z)) - 2.0 * sq);
  e += exp(2.0 * log(fabs(p11.z - p12.z)) - 2.0 * sp);
  return sqrt(e / 2.0 / 3.0);
}

__device__ __forceinline__ void gh_est_scales(GhVec2 q, GhVec2 p, double rtol, double atol,
                                              double& sq, double& sp) {
  const double mq = fmax(log(fabs(q.x)), log(fabs(q.y)));
  const double mp = fmax(log(fabs(p.x)), log(fabs(p.y)));
  double c = fmax(atol, rtol + mq);
  sq = log(exp(atol - c) + exp(rtol + mq - c)) + c;
  c = fmax(atol, rtol + mp);
  sp = log(exp(atol - c) + exp(rtol + mp - c)) + c;
}

__device__ __forceinline__ double gh_est_err(GhVec2 q11, GhVec2 q12, GhVec2 p11, GhVec2 p12,
                                             double sq, double sp) {
  double e = 0.0;
  e += exp(2.0 * log(fabs(q11.x - q12.x)) - 2.0 * sq);
  e += exp(2.0 * log(fabs(p11.x - p12.x)) - 2.0 * sp);
  e += exp(2.0 * log(fabs(q11.y - q12.y)) - 2.0 * sq);
  e += exp(2.0 * log(fabs(p11.y - p12.y)) - 2.0 * sp);
  return sqrt(e / 2.0 / 2.0);
}

// leapfrog_estimate_step, bovy_symplecticode.c:588-663
// The step estimators run once per orbit; they are inlined so that a kernel's
// register count is not set by a separately compiled copy (a call's callee is
// budgeted at its own, larger, register use).
template <class F, class V>
__device__ __forceinline__ double gh_leapfrog_estimate_step(const F& T, V qo, V po, double dt,
                                                            double to, double rtol, double atol) {
  double err = 2.0;
  const double init_dt = dt;
  double sq, sp;
  gh_est_scales(qo, po, rtol, atol, sq, sp);
  dt *= 2.0;
  while (err > 1.0 && init_dt / dt < GH_MAX_DT_REDUCE) {
    dt /= 2.0;
    V q12 = gh_drift(qo, po, dt / 2.0);
    V a = gh_accel(T, to + dt / 2.0, q12);
    const V p11 = gh_drift(po, a, dt);
    const V q11 = gh_drift(q12, p11, dt / 2.0);
    q12 = gh_drift(qo, po, dt / 4.0);
    a = gh_accel(T, to + dt / 4.0, q12);
    const V ptmp = gh_drift(po, a, dt / 2.0);
    const V qtmp = gh_drift(q12, ptmp, dt / 2.0);
    a = gh_accel(T, to + 3.0 * dt / 4.0, qtmp);
    const V p12 = gh_drift(ptmp, a, dt / 2.0);
    q12 = gh_drift(qtmp, p12, dt / 4.0);
    err = gh_est_err(q11, q12, p11, p12, sq, sp);
  }
  return dt;
}

// symplec4_estimate_step, bovy_symplecticode.c:665-808
template <class F, class V>
__device__ __forceinline__ double gh_symplec4_estimate_step(const F& T, V qo, V po, double dt,
                                                            double to, double rtol, double atol) {
  const double c1 = GH_S4C1, c2 = GH_S4C2, c3 = GH_S4C3, c4 = GH_S4C4;
  const double d1 = GH_S4D1, d2 = GH_S4D2, d3 = GH_S4D3;
  double err = 2.0;
  const double init_dt = dt;
  double sq, sp;
  gh_est_scales(qo, po, rtol, atol, sq, sp);
  dt *= 2.0;
  while (err > 1.0 && init_dt / dt < GH_MAX_DT_REDUCE) {
    dt /= 2.0;
    // one step of dt
    V q = gh_drift(qo, po, c1 * dt);
    to += c1 * dt;
    V p = gh_drift(po, gh_accel(T, to, q), d1 * dt);
    q = gh_drift(q, p, c2 * dt);
    to += c2 * dt;
    p = gh_drift(p, gh_accel(T, to, q), d2 * dt);
    q = gh_drift(q, p, c3 * dt);
    to += c3 * dt;
    const V p11 = gh_drift(p, gh_accel(T, to, q), d3 * dt);
    const V q11 = gh_drift(q, p11, c4 * dt);
    to += c4 * dt;
    to -= dt;
    // two steps of dt/2
    q = gh_drift(qo, po, c1 * dt / 2.0);
    to += c1 * dt / 2.0;
    p = gh_drift(po, gh_accel(T, to, q), d1 * dt / 2.0);
    q = gh_drift(q, p, c2 * dt / 2.0);
    to += c2 * dt / 2.0;
    p = gh_drift(p, gh_accel(T, to, q), d2 * dt / 2.0);
    q = gh_drift(q, p, c3 * dt / 2.0);
    to += c3 * dt / 2.0;
    p = gh_drift(p, gh_accel(T, to, q), d3 * dt / 2.0);
    q = gh_drift(q, p, (c1 + c4) * dt / 2.0);
    to += (c1 + c4) * dt / 2.0;
    p = gh_drift(p, gh_accel(T, to, q), d1 * dt / 2.0);
    q = gh_drift(q, p, c2 * dt / 2.0);
    to += c2 * dt / 2.0;
    p = gh_drift(p, gh_accel(T, to, q), d2 * dt / 2.0);
    q = gh_drift(q, p, c3 * dt / 2.0);
    to += c3 * dt / 2.0;
    p = gh_drift(p, gh_accel(T, to, q), d3 * dt / 2.0);
    q = gh_drift(q, p, c4 * dt / 2.0);
    to += c4 * dt / 2.0;
    err = gh_est_err(q11, q, p11, p, sq, sp);
    to -= dt;
  }
  return dt;
}

// symplec6 coefficients (Yoshida 1990), bovy_symplecticode.c:100-107: drifts
// c1..c8, kicks d1..d7.  The seven-kick loops below stay rolled (one force
// evaluation per loop body), so a run-time dispatch kernel does not carry
// the force switch 28 times.
static __constant__ double gh_s6c[8] = {0.392256805238780, 0.510043411918458, -0.471053385409758,
                                 0.687531682525198e-1, 0.687531682525198e-1, -0.471053385409758,
                                 0.510043411918458, 0.392256805238780};
static __constant__ double gh_s6d[7] = {0.784513610477560, 0.235573213359357, -0.117767998417887e1,
                                 0.131518632068391e1, -0.117767998417887e1, 0.235573213359357,
                                 0.784513610477560};

// symplec6_estimate_step, bovy_symplecticode.c:810-1000: one step of dt
// against two of dt/2 (middle drifts merged as (c1 + c8) dt/2), the scales
// and norm of the symplec4 estimator.
template <class F, class V>
__device__ __forceinline__ double gh_symplec6_estimate_step(const F& T, V qo, V po, double dt,
                                                            double to, double rtol, double atol) {
  double err = 2.0;
  const double init_dt = dt;
  double sq, sp;
  gh_est_scales(qo, po, rtol, atol, sq, sp);
  dt *= 2.0;
  while (err > 1.0 && init_dt / dt < GH_MAX_DT_REDUCE) {
    dt /= 2.0;
    V q = gh_drift(qo, po, gh_s6c[0] * dt);
    V p = po;
    to += gh_s6c[0] * dt;
#pragma unroll 1
    for (int k = 0; k < 7; k++) {
      p = gh_drift(p, gh_accel(T, to, q), gh_s6d[k] * dt);
      q = gh_drift(q, p, gh_s6c[k + 1] * dt);
      to += gh_s6c[k + 1] * dt;
    }
    const V q11 = q, p11 = p;
    to -= dt;
    q = gh_drift(qo, po, gh_s6c[0] * dt / 2.0);
    p = po;
    to += gh_s6c[0] * dt / 2.0;
#pragma unroll 1
    for (int k = 0; k < 14; k++) {
      const int kk = k < 7 ? k : k - 7;
      p = gh_drift(p, gh_accel(T, to, q), gh_s6d[kk] * dt / 2.0);
      const double cn = (k == 6) ? gh_s6c[0] + gh_s6c[7] : gh_s6c[kk + 1];
      q = gh_drift(q, p, cn * dt / 2.0);
      to += cn * dt / 2.0;
    }
    err = gh_est_err(q11, q, p11, p, sq, sp);
    to -= dt;
  }
  return dt;
}

// bovy_coords.c:101-115 (cylindrical -> rectangular)
__device__ __forceinline__ void gh_cyl_to_rect(const double* v, double* w) {
  double sp, cp;
  sincos(v[5], &sp, &cp);
  const double R = v[0], vR = v[1], vT = v[2];
  w[0] = R * cp;
  w[1] = R * sp;
  w[2] = v[3];
  w[3] = vR * cp - vT * sp;
  w[4] = vR * sp + vT * cp;
  w[5] = v[4];
}

// bovy_coords.c:125-139 (rectangular -> cylindrical, phi = atan2(y, x)),
// stored as one 48-byte row
__device__ __forceinline__ void gh_store_cyl(double* __restrict__ out, double x, double y,
                                             double z, double vx, double vy, double vz) {
  const double R = sqrt(x * x + y * y);
  const double phi = atan2(y, x);
  double cp = 1.0, sp = 0.0;
  if (R > 0.0) {
    cp = x / R;
    sp = y / R;
  }
  double2* o = reinterpret_cast<double2*>(out);
  o[0] = make_double2(R, vx * cp + vy * sp);
  o[1] = make_double2(-vx * sp + vy * cp, z);
  o[2] = make_double2(vz, phi);
}

// bovy_coords.c:57-69 (polar -> rectangular, planar orbits)
__device__ __forceinline__ void gh_polar_to_rect(const double* v, double* w) {
  double sp, cp;
  sincos(v[3], &sp, &cp);
  const double R = v[0], vR = v[1], vT = v[2];
  w[0] = R * cp;
  w[1] = R * sp;
  w[2] = vR * cp - vT * sp;
  w[3] = vR * sp + vT * cp;
}

// bovy_coords.c:79-91 (rectangular -> polar, phi = atan2(y, x)), stored as
// one 32-byte row
__device__ __forceinline__ void gh_store_polar(double* __restrict__ out, double x, double y,
                                               double vx, double vy) {
  const double R = sqrt(x * x + y * y);
  const double phi = atan2(y, x);
  double cp = 1.0, sp = 0.0;
  if (R > 0.0) {
    cp = x / R;
    sp = y / R;
  }
  double2* o = reinterpret_cast<double2*>(out);
  o[0] = make_double2(R, vx * cp + vy * sp);
  o[1] = make_double2(-vx * sp + vy * cp, phi);
}

// Row layout of full and planar orbits: N doubles per row, galpy's
// (R,vR,vT,z,vz,phi) / (R,vR,vT,phi) outside and the rectangular
// (q, p) = (x,y,z,vx,vy,vz) / (x,y,vx,vy) inside the integrators.
template <class V>
struct GhPhase;

template <>
struct GhPhase<GhVec3> {
  static constexpr int N = 6;
  __device__ static __forceinline__ void to_rect(const double* v, double* w) {
    gh_cyl_to_rect(v, w);
  }
  __device__ static __forceinline__ GhVec3 q(const double* w) { return GhVec3{w[0], w[1], w[2]}; }
  __device__ static __forceinline__ GhVec3 p(const double* w) { return GhVec3{w[3], w[4], w[5]}; }
  __device__ static __forceinline__ void set(double* w, GhVec3 q, GhVec3 p) {
    w[0] = q.x; w[1] = q.y; w[2] = q.z;
    w[3] = p.x; w[4] = p.y; w[5] = p.z;
  }
  __device__ static __forceinline__ void store(double* __restrict__ out, GhVec3 q, GhVec3 p) {
    gh_store_cyl(out, q.x, q.y, q.z, p.x, p.y, p.z);
  }
};

// Variational states: rows are galpy's rectangular (q, p | dq, dp)
// (save_qp_dxdv, bovy_symplecticode.c:79-90) -- no coordinate transform.
template <>
struct GhPhase<GhVec6> {
  static constexpr int N = 12;
  __device__ static __forceinline__ void to_rect(const double* v, double* w) {
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = v[k];
  }
  __device__ static __forceinline__ GhVec6 q(const double* w) {
    return GhVec6{w[0], w[1], w[2], w[6], w[7], w[8]};
  }
  __device__ static __forceinline__ GhVec6 p(const double* w) {
    return GhVec6{w[3], w[4], w[5], w[9], w[10], w[11]};
  }
  __device__ static __forceinline__ void set(double* w, GhVec6 q, GhVec6 p) {
    w[0] = q.x; w[1] = q.y; w[2] = q.z; w[3] = p.x; w[4] = p.y; w[5] = p.z;
    w[6] = q.dx; w[7] = q.dy; w[8] = q.dz; w[9] = p.dx; w[10] = p.dy; w[11] = p.dz;
  }
  __device__ static __forceinline__ void store(double* __restrict__ out, GhVec6 q, GhVec6 p) {
    double2* o = reinterpret_cast<double2*>(out);
    o[0] = make_double2(q.x, q.y);
    o[1] = make_double2(q.z, p.x);
    o[2] = make_double2(p.y, p.z);
    o[3] = make_double2(q.dx, q.dy);
    o[4] = make_double2(q.dz, p.dx);
    o[5] = make_double2(p.dy, p.dz);
  }
};

template <>
struct GhPhase<GhVec4> {
  static constexpr int N = 8;
  __device__ static __forceinline__ void to_rect(const double* v, double* w) {
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = v[k];
  }
  __device__ static __forceinline__ GhVec4 q(const double* w) {
    return GhVec4{w[0], w[1], w[4], w[5]};
  }
  __device__ static __forceinline__ GhVec4 p(const double* w) {
    return GhVec4{w[2], w[3], w[6], w[7]};
  }
  __device__ static __forceinline__ void set(double* w, GhVec4 q, GhVec4 p) {
    w[0] = q.x; w[1] = q.y; w[2] = p.x; w[3] = p.y;
    w[4] = q.dx; w[5] = q.dy; w[6] = p.dx; w[7] = p.dy;
  }
  __device__ static __forceinline__ void store(double* __restrict__ out, GhVec4 q, GhVec4 p) {
    double2* o = reinterpret_cast<double2*>(out);
    o[0] = make_double2(q.x, q.y);
    o[1] = make_double2(p.x, p.y);
    o[2] = make_double2(q.dx, q.dy);
    o[3] = make_double2(p.dx, p.dy);
  }
};

template <>
struct GhPhase<GhVec2> {
  static constexpr int N = 4;
  __device__ static __forceinline__ void to_rect(const double* v, double* w) {
    gh_polar_to_rect(v, w);
  }
  __device__ static __forceinline__ GhVec2 q(const double* w) { return GhVec2{w[0], w[1]}; }
  __device__ static __forceinline__ GhVec2 p(const double* w) { return GhVec2{w[2], w[3]}; }
  __device__ static __forceinline__ void set(double* w, GhVec2 q, GhVec2 p) {
    w[0] = q.x; w[1] = q.y;
    w[2] = p.x; w[3] = p.y;
  }
  __device__ static __forceinline__ void store(double* __restrict__ out, GhVec2 q, GhVec2 p) {
    gh_store_polar(out, q.x, q.y, p.x, p.y);
  }
};
