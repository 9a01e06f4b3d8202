// Variational equations of galpy's integrate_dxdv on the device: the orbit
// plus one phase-space deviation, whose kick is K dq with K the Cartesian
// Hessian of -Phi at the orbit's position and time (evalRectForce_dxdv,
// integrateFullOrbit.c:1619-1681; evalPlanarRectForce_dxdv,
// integratePlanarOrbit.c:1228-1275; the same K in the variational
// right-hand sides evalRectDeriv_dxdv, integrateFullOrbit.c:1496-1600, and
// evalPlanarRectDeriv_dxdv, integratePlanarOrbit.c:1124-1175, that RK4, RK6,
// DOPR54 and DOP853 integrate).
//
// The orbit's acceleration is the plain force kernel's, so the orbit of a
// variational run is bit-identical to a plain run (as galpy's base block is,
// integrateFullOrbit.c:1640-1646); K dq is a Hessian-vector product written
// in Cartesian form from each component's potential -- galpy assembles the
// same K from six cylindrical second derivatives and the forces.  For planar
// orbits K is the (x, y) block of the full Hessian at z = 0, which is
// galpy's planar Hessian (planarR2deriv, planarphi2deriv, planarRphideriv
// are the full second derivatives in the plane).
#pragma once
#include "adjoint_dev.h"

// K v of component TYPE at (x, t), added to k.  The time-independent
// components reuse the adjoint's Hessian-vector products.
template <int TYPE>
__device__ __forceinline__ void gh_comp_hvp(const GhPotTable* __restrict__ T, int i, GhVec3 x,
                                            double r2, double rinv, double t, GhVec3 v,
                                            GhVec3& k) {
  GhVec3 a{0.0, 0.0, 0.0};
  double dth[GH_ADJ_NPAR] = {0.0, 0.0, 0.0};
  gh_comp_adj<TYPE>(T, i, x, r2, rinv, v, a, k, dth);
}

// DehnenBarPotential in Cartesian form: Phi = A(t) P(x, y) g(r) with
// A = amp smooth(t) (DehnenBarPotential.c:5-17), P = R^2 cos 2(phi - psi) =
// (x^2 - y^2) cos 2 psi + 2 x y sin 2 psi, psi = omegab t + barphi, and
// g = f(r) / r^2 with f = (r/rb)^3 - 2 inside rb, -(rb/r)^3 outside
// (DehnenBarPotential.py _evaluate).  Its Hessian is
//   H = A [g P'' + (g'/r)(P' x^T + x P'^T) + P G2 x x^T + P (g'/r) I],
//   G2 = (g'' - g'/r) / r^2,
// and K v = -H v; the branch at r <= rb is the one the force takes.
template <>
__device__ __forceinline__ void gh_comp_hvp<GH_DEHNENBAR>(const GhPotTable* __restrict__ T, int i,
                                                          GhVec3 x, double r2, double rinv,
                                                          double t, GhVec3 v, GhVec3& k) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double rb = C->p[3];
  const double A = C->p[0] * gh_bar_smooth(t, C->p[1], C->p[2]);
  const GhSinCos ps = gh_sincos(2.0 * (C->p[4] * t + C->p[5]));
  const double c2 = ps.c, s2 = ps.s;
  const double r = r2 * rinv;
  const double ri2 = rinv * rinv, ri4 = ri2 * ri2;
  double g, G1, G2;  // g, g'/r, (g'' - g'/r)/r^2
  if (r <= rb) {
    const double irb3 = C->c[0] * C->c[0] * C->c[0];  // 1/rb^3
    g = r * irb3 - 2.0 * ri2;
    G1 = irb3 * rinv + 4.0 * ri4;
    G2 = -16.0 * ri4 * ri2 - irb3 * ri2 * rinv;
  } else {
    const double rb3 = rb * rb * rb;
    const double ri5 = ri4 * rinv;
    g = -rb3 * ri5;
    G1 = 5.0 * rb3 * ri5 * ri2;
    G2 = -35.0 * rb3 * ri5 * ri4;
  }
  const double P = (x.x * x.x - x.y * x.y) * c2 + 2.0 * x.x * x.y * s2;
  const double Px = 2.0 * (x.x * c2 + x.y * s2), Py = 2.0 * (x.x * s2 - x.y * c2);
  const double xv = x.x * v.x + x.y * v.y + x.z * v.z;
  const double pv = Px * v.x + Py * v.y;
  const double u1x = 2.0 * (c2 * v.x + s2 * v.y), u1y = 2.0 * (s2 * v.x - c2 * v.y);
  const double cx = G1 * pv + P * G2 * xv;  // coefficient of x
  k.x += -A * (g * u1x + G1 * Px * xv + cx * x.x + P * G1 * v.x);
  k.y += -A * (g * u1y + G1 * Py * xv + cx * x.y + P * G1 * v.y);
  k.z += -A * (cx * x.z + P * G1 * v.z);
}

template <class Spec>
struct GhHvpSum;

template <>
struct GhHvpSum<GhSpec<>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, GhVec3 x, double r2,
                                             double rinv, double t, GhVec3 v, GhVec3& k) {
    for (int i = 0; i < T->npot; i++) {
      switch (T->comp[i].type) {
#define GH_HVP_CASE(CODE) \
  case CODE: gh_comp_hvp<CODE>(T, i, x, r2, rinv, t, v, k); break;
        GH_HVP_CASE(GH_POWERSPHCUTOFF)
        GH_HVP_CASE(GH_MIYAMOTONAGAI)
        GH_HVP_CASE(GH_NFW)
        GH_HVP_CASE(GH_HERNQUIST)
        GH_HVP_CASE(GH_LOGHALO)
        GH_HVP_CASE(GH_DEHNENBAR)
        GH_HVP_CASE(GH_POWERSPH)
        GH_HVP_CASE(GH_JAFFE)
        GH_HVP_CASE(GH_ISOCHRONE)
        GH_HVP_CASE(GH_PLUMMER)
        GH_HVP_CASE(GH_PSEUDOISO)
        GH_HVP_CASE(GH_KUZMINDISK)
        GH_HVP_CASE(GH_FLATPOWER)
        GH_HVP_CASE(GH_BURKERT)
        GH_HVP_CASE(GH_DEHNENCORESPH)
        GH_HVP_CASE(GH_DEHNENSPH)
        GH_HVP_CASE(GH_HOMOGSPHERE)
#undef GH_HVP_CASE
        default: break;
      }
    }
  }
};

template <int... Codes>
struct GhHvpSum<GhSpec<Codes...>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, GhVec3 x, double r2,
                                             double rinv, double t, GhVec3 v, GhVec3& k) {
    int i = 0;
    ((gh_comp_hvp<Codes>(T, i, x, r2, rinv, t, v, k), i++), ...);
  }
};

// K(q, t) v
template <class Spec>
__device__ __forceinline__ GhVec3 gh_hvp(const GhPotTable* __restrict__ T, double t, GhVec3 q,
                                         GhVec3 v) {
  const double r2 = q.x * q.x + q.y * q.y + q.z * q.z;
  const double rinv = rsqrt(r2);
  GhVec3 k{0.0, 0.0, 0.0};
  GhHvpSum<Spec>::add(T, q, r2, rinv, t, v, k);
  return k;
}

// (a(q), K(q) dq) for a full orbit and its deviation
template <class Spec>
__device__ __forceinline__ GhVec6 gh_accel(const GhForce<Spec>& F, double t, GhVec6 Q) {
  GhVec3 a;
  F.rect(t, Q.x, Q.y, Q.z, a.x, a.y, a.z);
  const GhVec3 k = gh_hvp<Spec>(F.T, t, GhVec3{Q.x, Q.y, Q.z}, GhVec3{Q.dx, Q.dy, Q.dz});
  return GhVec6{a.x, a.y, a.z, k.x, k.y, k.z};
}

// the same for a planar orbit
template <class Spec>
__device__ __forceinline__ GhVec4 gh_accel(const GhForce<Spec>& F, double t, GhVec4 Q) {
  GhVec2 a;
  F.rect(t, Q.x, Q.y, a.x, a.y);
  const GhVec3 k = gh_hvp<Spec>(F.T, t, GhVec3{Q.x, Q.y, 0.0}, GhVec3{Q.dx, Q.dy, 0.0});
  return GhVec4{a.x, a.y, k.x, k.y};
}
