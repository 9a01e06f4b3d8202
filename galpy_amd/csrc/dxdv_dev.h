// Variational equations of galpy's integrate_dxdv on the device: the orbit
// plus one phase-space deviation, whose kick is K dq with K the Cartesian
// Hessian of -Phi at the orbit's position (evalRectForce_dxdv,
// integrateFullOrbit.c:1619-1681; evalPlanarRectForce_dxdv,
// integratePlanarOrbit.c:1228-1275; the same K in the variational
// right-hand sides evalRectDeriv_dxdv, integrateFullOrbit.c:1496-1600, and
// evalPlanarRectDeriv_dxdv, integratePlanarOrbit.c:1124-1175, that DOP853
// integrates).
//
// The orbit's acceleration is the plain force kernel's, so the orbit of a
// variational run is bit-identical to a plain run (as galpy's base block is,
// integrateFullOrbit.c:1640-1646); K dq is the Hessian-vector product of the
// adjoint (time-independent components with a Hessian: no bar, no triaxial
// halo -- checked on the host).  For planar orbits K is the (x, y) block of
// the full Hessian at z = 0, which is galpy's planar Hessian
// (planarR2deriv, planarphi2deriv, planarRphideriv are the full second
// derivatives in the plane).
#pragma once
#include "adjoint_dev.h"

// K v at q (and nothing else: the acceleration and parameter slots the
// adjoint sum also produces are discarded)
template <class Spec>
__device__ __forceinline__ GhVec3 gh_hvp(const GhPotTable* __restrict__ T, GhVec3 q, GhVec3 v) {
  const double r2 = q.x * q.x + q.y * q.y + q.z * q.z;
  const double rinv = rsqrt(r2);
  GhVec3 a{0.0, 0.0, 0.0}, k{0.0, 0.0, 0.0};
  double d[GH_MAX_POT][GH_ADJ_NPAR] = {};
  GhAdjSum<Spec>::add(T, q, r2, rinv, v, a, k, d);
  return k;
}

// (a(q), K(q) dq) for a full orbit and its deviation
template <class Spec>
__device__ __forceinline__ GhVec6 gh_accel(const GhForce<Spec>& F, double t, GhVec6 Q) {
  GhVec3 a;
  F.rect(t, Q.x, Q.y, Q.z, a.x, a.y, a.z);
  const GhVec3 k = gh_hvp<Spec>(F.T, GhVec3{Q.x, Q.y, Q.z}, GhVec3{Q.dx, Q.dy, Q.dz});
  return GhVec6{a.x, a.y, a.z, k.x, k.y, k.z};
}

// the same for a planar orbit
template <class Spec>
__device__ __forceinline__ GhVec4 gh_accel(const GhForce<Spec>& F, double t, GhVec4 Q) {
  GhVec2 a;
  F.rect(t, Q.x, Q.y, a.x, a.y);
  const GhVec3 k = gh_hvp<Spec>(F.T, GhVec3{Q.x, Q.y, 0.0}, GhVec3{Q.dx, Q.dy, 0.0});
  return GhVec4{a.x, a.y, k.x, k.y};
}
