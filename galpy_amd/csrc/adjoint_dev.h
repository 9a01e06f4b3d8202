// Reverse-time adjoint of the symplectic integrators (leapfrog, symplec4).
//
// Forward: N full steps of galpy's steppers (bovy_symplecticode.c:132-377
// with one step per output interval, i.e. t = t0 + k dt).  Each step is a
// composition of drifts q += c h p and kicks p += d h a(q, theta).
// Reverse: starting from the final state and the loss gradient
// lambda = dL/d(q, p), every sub-step is undone exactly in reverse order --
// the trajectory is recomputed by reversibility, never stored -- while
//   drift:  lambda_p += c h lambda_q
//   kick:   lambda_q += d h K(q) lambda_p,   dL/dtheta += d h (da/dtheta)^T lambda_p
// with K = da/dq the Cartesian Hessian of -Phi (symmetric).  The Hessians
// follow the second derivatives galpy uses for its tangent map
// (integrateFullOrbit.c:1619-1681, MiyamotoNagaiPotential.c:66-119,
// NFWPotential.c:63-107), written in Cartesian form.
#pragma once
#include "integrators_dev.h"

#define GH_ADJ_NPAR 3  // parameter slots per component: its packed args 0..2

// Acceleration, K.lam and (da/dtheta).lam of one component; dth[j] is the
// slot of packed argument j of the component.
template <int TYPE>
__device__ __forceinline__ void gh_comp_adj(const GhPotTable* __restrict__ T, int i, GhVec3 x,
                                            double r2, double rinv, GhVec3 lam, GhVec3& a,
                                            GhVec3& k, double* dth);

// spherical a = g(r) x:  K lam = g lam + (g'/r)(x.lam) x,  da/dtheta.lam = dg/dtheta (x.lam)
__device__ __forceinline__ void gh_sph_adj(GhVec3 x, GhVec3 lam, double g, double gpr, GhVec3& a,
                                           GhVec3& k) {
  const double xl = x.x * lam.x + x.y * lam.y + x.z * lam.z;
  a.x += g * x.x;
  a.y += g * x.y;
  a.z += g * x.z;
  k.x += g * lam.x + gpr * xl * x.x;
  k.y += g * lam.y + gpr * xl * x.y;
  k.z += g * lam.z + gpr * xl * x.z;
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_POWERSPHCUTOFF>(const GhPotTable* __restrict__ T,
                                                               int i, GhVec3 x, double r2,
                                                               double rinv, GhVec3 lam, GhVec3& a,
                                                               GhVec3& k, double* dth) {
  // g = -amp M(r)/r^3, M' = 4 pi r^(2-alpha) e^(-r^2/rc^2)  =>
  // g'/r = -amp 4 pi r^(-2-alpha) e^(-r^2/rc^2) - 3 g / r^2
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], alpha = C->p[1];
  const double xx = r2 * C->c[3];
  const double gam = gh_lower_incgamma_tab(C, T->ser[i][0], &T->tab[(int)C->c[11]][0][0], xx);
  const double g = -C->c[2] * gam * (rinv * rinv * rinv);
  const double gpr = -amp * 4.0 * M_PI * exp(-(2.0 + alpha) * log(r2 * rinv) - xx) - 3.0 * g * rinv * rinv;
  gh_sph_adj(x, lam, g, gpr, a, k);
  dth[0] += (g / amp) * (x.x * lam.x + x.y * lam.y + x.z * lam.z);
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_NFW>(const GhPotTable* __restrict__ T, int i,
                                                    GhVec3 x, double r2, double rinv, GhVec3 lam,
                                                    GhVec3& a, GhVec3& k, double* dth) {
  // g = amp h, h = 1/(r^2 (a+r)) - L/r^3, L = log(1+r/a)
  // h' = -3/(r^3 (a+r)) - 1/(r^2 (a+r)^2) + 3 L / r^4;  dh/da = 1/(a r (a+r)^2)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], aa = C->p[1];
  const double r = r2 * rinv;
  const double ar = aa + r;
  const double iar = 1.0 / ar;
  const double L = gh_log_ge1(1.0 + r * C->c[0]);
  const double ri2 = rinv * rinv, ri3 = ri2 * rinv;
  const double h = ri2 * iar - L * ri3;
  const double hp = -3.0 * ri3 * iar - ri2 * iar * iar + 3.0 * L * ri2 * ri2;
  gh_sph_adj(x, lam, amp * h, amp * hp * rinv, a, k);
  const double xl = x.x * lam.x + x.y * lam.y + x.z * lam.z;
  dth[0] += h * xl;
  dth[1] += amp * C->c[0] * rinv * iar * iar * xl;
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_HERNQUIST>(const GhPotTable* __restrict__ T, int i,
                                                          GhVec3 x, double r2, double rinv,
                                                          GhVec3 lam, GhVec3& a, GhVec3& k,
                                                          double* dth) {
  // g = -(amp/2)/(r (a+r)^2),  g'/r = (amp/2)(a+3r)/(r^3 (a+r)^3),  dg/da = amp/(r (a+r)^3)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], aa = C->p[1];
  const double r = r2 * rinv;
  const double iar = 1.0 / (aa + r);
  const double g = -0.5 * amp * rinv * iar * iar;
  const double gpr = 0.5 * amp * (aa + 3.0 * r) * rinv * rinv * rinv * iar * iar * iar;
  gh_sph_adj(x, lam, g, gpr, a, k);
  const double xl = x.x * lam.x + x.y * lam.y + x.z * lam.z;
  dth[0] += (g / amp) * xl;
  dth[1] += amp * rinv * iar * iar * iar * xl;
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_MIYAMOTONAGAI>(const GhPotTable* __restrict__ T,
                                                              int i, GhVec3 x, double r2,
                                                              double rinv, GhVec3 lam, GhVec3& a,
                                                              GhVec3& k, double* dth) {
  // a = -amp (x, y, w) D^-3/2, s = sqrt(z^2+b^2), A = a+s, D = R^2 + A^2,
  // w = z A / s, dw/dz = 1 + a b^2/s^3
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], aa = C->p[1], b = C->p[2];
  const double s2 = x.z * x.z + C->c[0];
  const double is = rsqrt(s2);
  const double s = s2 * is;
  const double A = aa + s;
  const double D = x.x * x.x + x.y * x.y + A * A;
  const double iD = rsqrt(D);
  const double D32 = iD * iD * iD;
  const double D52 = D32 * iD * iD;
  const double w = x.z * A * is;
  const double dw = 1.0 + aa * b * b * is * is * is;
  a.x += -amp * x.x * D32;
  a.y += -amp * x.y * D32;
  a.z += -amp * w * D32;
  const double u = x.x * lam.x + x.y * lam.y + w * lam.z;
  k.x += -amp * D32 * lam.x + 3.0 * amp * x.x * D52 * u;
  k.y += -amp * D32 * lam.y + 3.0 * amp * x.y * D52 * u;
  k.z += -amp * dw * D32 * lam.z + 3.0 * amp * w * D52 * u;
  const double xyl = x.x * lam.x + x.y * lam.y;
  dth[0] += -(xyl + w * lam.z) * D32;
  const double da_z_a = -amp * (x.z * is * D32 - 3.0 * w * A * D52);
  dth[1] += 3.0 * amp * A * D52 * xyl + da_z_a * lam.z;
  const double bs = b * is;
  const double da_z_b = -amp * (-aa * x.z * b * is * is * is * D32 - 3.0 * w * A * bs * D52);
  dth[2] += 3.0 * amp * A * bs * D52 * xyl + da_z_b * lam.z;
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_LOGHALO>(const GhPotTable* __restrict__ T, int i,
                                                        GhVec3 x, double r2, double rinv,
                                                        GhVec3 lam, GhVec3& a, GhVec3& k,
                                                        double* dth) {
  // Phi = amp/2 ln(x^2 + y^2/b^2 + z^2/q^2 + c) (R^2 (1 - (1 - 1/b^2) sin^2 phi)
  // = x^2 + y^2/b^2; b^2 = 1 for the axisymmetric halo, packed as 1 - 1/b^2
  // = 2): a = -amp w / D, w = (x, y/b^2, z/q^2), D the log's argument,
  // K lam = -amp (lam.x, lam.y/b^2, lam.z/q^2) / D + 2 amp w (w.lam) / D^2
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], core2 = C->p[2], e = C->p[3];
  const double iq2 = C->c[0];
  const double ib2 = (e < 1.0) ? 1.0 - e : 1.0;
  const double D = x.x * x.x + ib2 * x.y * x.y + x.z * x.z * iq2 + core2;
  const double iD = 1.0 / D;
  const GhVec3 v{x.x, ib2 * x.y, x.z * iq2};
  a.x += -amp * v.x * iD;
  a.y += -amp * v.y * iD;
  a.z += -amp * v.z * iD;
  const double vl = v.x * lam.x + v.y * lam.y + v.z * lam.z;
  k.x += -amp * lam.x * iD + 2.0 * amp * v.x * vl * iD * iD;
  k.y += -amp * ib2 * lam.y * iD + 2.0 * amp * v.y * vl * iD * iD;
  k.z += -amp * lam.z * iq2 * iD + 2.0 * amp * v.z * vl * iD * iD;
  dth[0] += -vl * iD;
}

// the spherical components of potential_dev.h (gh_sph_g): g and g'/r; the
// amplitude is the only differentiable parameter (g is linear in it)
#define GH_SPH_ADJ(CODE)                                                                      \
  template <>                                                                                 \
  __device__ __forceinline__ void gh_comp_adj<CODE>(const GhPotTable* __restrict__ T, int i,  \
                                                    GhVec3 x, double r2, double rinv,         \
                                                    GhVec3 lam, GhVec3& a, GhVec3& k,         \
                                                    double* dth) {                            \
    const GhComponent* __restrict__ C = &T->comp[i];                                          \
    double g, gpr;                                                                            \
    gh_sph_g<CODE>(C, r2 * rinv, rinv, g, gpr);                                               \
    gh_sph_adj(x, lam, g, gpr, a, k);                                                         \
    dth[0] += (g / C->p[0]) * (x.x * lam.x + x.y * lam.y + x.z * lam.z);                      \
  }
GH_SPH_ADJ(GH_POWERSPH)
GH_SPH_ADJ(GH_JAFFE)
GH_SPH_ADJ(GH_ISOCHRONE)
GH_SPH_ADJ(GH_PLUMMER)
GH_SPH_ADJ(GH_PSEUDOISO)
GH_SPH_ADJ(GH_BURKERT)
GH_SPH_ADJ(GH_DEHNENCORESPH)
GH_SPH_ADJ(GH_DEHNENSPH)
GH_SPH_ADJ(GH_HOMOGSPHERE)
#undef GH_SPH_ADJ

template <>
__device__ __forceinline__ void gh_comp_adj<GH_KUZMINDISK>(const GhPotTable* __restrict__ T, int i,
                                                           GhVec3 x, double r2, double rinv,
                                                           GhVec3 lam, GhVec3& a, GhVec3& k,
                                                           double* dth) {
  // a = -amp (x, y, w) D^-3/2, w = sign(z) (a + |z|), D = R^2 + (a + |z|)^2;
  // K as galpy's KuzminDiskPotential.c second derivatives: the z-z entry
  // uses (a + |z|)^2 also at z = 0, where sign(z), and so w, vanish
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0];
  const double az = C->p[1] + fabs(x.z);
  const double iD = rsqrt(x.x * x.x + x.y * x.y + az * az);
  const double D32 = iD * iD * iD;
  const double D52 = D32 * iD * iD;
  const double w = (x.z == 0.0) ? 0.0 : copysign(az, x.z);
  a.x += -amp * x.x * D32;
  a.y += -amp * x.y * D32;
  a.z += -amp * w * D32;
  const double u = x.x * lam.x + x.y * lam.y;
  k.x += -amp * D32 * lam.x + 3.0 * amp * x.x * D52 * (u + w * lam.z);
  k.y += -amp * D32 * lam.y + 3.0 * amp * x.y * D52 * (u + w * lam.z);
  k.z += -amp * D32 * lam.z + 3.0 * amp * D52 * (w * u + az * az * lam.z);
  dth[0] += -(u + w * lam.z) * D32;
}

template <>
__device__ __forceinline__ void gh_comp_adj<GH_FLATPOWER>(const GhPotTable* __restrict__ T, int i,
                                                          GhVec3 x, double r2, double rinv,
                                                          GhVec3 lam, GhVec3& a, GhVec3& k,
                                                          double* dth) {
  // a = g w, g = -amp m2^(-alpha/2 - 1), w = (x, y, z/q^2),
  // m2 = R^2 + z^2/q^2 + core^2; K lam = g (lam.x, lam.y, lam.z/q^2)
  // + 2 g' (w . lam) w with g' = dg/dm2 = -g (alpha/2 + 1) / m2 (the second
  // derivatives of FlattenedPowerPotential.c:83-133)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], alpha = C->p[1], iq2 = C->c[0];
  const double m2 = C->p[3] + x.x * x.x + x.y * x.y + x.z * x.z * iq2;
  const double g = -amp * exp((-0.5 * alpha - 1.0) * log(m2));
  const double gp = -g * (0.5 * alpha + 1.0) / m2;
  const GhVec3 w{x.x, x.y, x.z * iq2};
  a.x += g * w.x;
  a.y += g * w.y;
  a.z += g * w.z;
  const double wl = w.x * lam.x + w.y * lam.y + w.z * lam.z;
  k.x += g * lam.x + 2.0 * gp * wl * w.x;
  k.y += g * lam.y + 2.0 * gp * wl * w.y;
  k.z += g * lam.z * iq2 + 2.0 * gp * wl * w.z;
  dth[0] += (g / amp) * wl;
}

__device__ __forceinline__ void gh_comp_adj_dyn(const GhPotTable* __restrict__ T, int i, GhVec3 x,
                                                double r2, double rinv, GhVec3 lam, GhVec3& a,
                                                GhVec3& k, double* dth) {
  switch (T->comp[i].type) {
    case GH_POWERSPHCUTOFF: gh_comp_adj<GH_POWERSPHCUTOFF>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_MIYAMOTONAGAI: gh_comp_adj<GH_MIYAMOTONAGAI>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_NFW: gh_comp_adj<GH_NFW>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_HERNQUIST: gh_comp_adj<GH_HERNQUIST>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_LOGHALO: gh_comp_adj<GH_LOGHALO>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_POWERSPH: gh_comp_adj<GH_POWERSPH>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_JAFFE: gh_comp_adj<GH_JAFFE>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_ISOCHRONE: gh_comp_adj<GH_ISOCHRONE>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_PLUMMER: gh_comp_adj<GH_PLUMMER>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_PSEUDOISO: gh_comp_adj<GH_PSEUDOISO>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_KUZMINDISK: gh_comp_adj<GH_KUZMINDISK>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_FLATPOWER: gh_comp_adj<GH_FLATPOWER>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_BURKERT: gh_comp_adj<GH_BURKERT>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_DEHNENCORESPH: gh_comp_adj<GH_DEHNENCORESPH>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_DEHNENSPH: gh_comp_adj<GH_DEHNENSPH>(T, i, x, r2, rinv, lam, a, k, dth); break;
    case GH_HOMOGSPHERE: gh_comp_adj<GH_HOMOGSPHERE>(T, i, x, r2, rinv, lam, a, k, dth); break;
    default: break;
  }
}

template <class Spec>
struct GhAdjSum;

template <>
struct GhAdjSum<GhSpec<>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, GhVec3 x, double r2,
                                             double rinv, GhVec3 lam, GhVec3& a, GhVec3& k,
                                             double (*dth)[GH_ADJ_NPAR]) {
    for (int i = 0; i < T->npot; i++) gh_comp_adj_dyn(T, i, x, r2, rinv, lam, a, k, dth[i]);
  }
};

template <int... Codes>
struct GhAdjSum<GhSpec<Codes...>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, GhVec3 x, double r2,
                                             double rinv, GhVec3 lam, GhVec3& a, GhVec3& k,
                                             double (*dth)[GH_ADJ_NPAR]) {
    int i = 0;
    ((gh_comp_adj<Codes>(T, i, x, r2, rinv, lam, a, k, dth[i]), i++), ...);
  }
};

// Kick adjoint at q: returns a(q); adds h K lam_p to lam_q and
// h (da/dtheta)^T lam_p to dth.  Only time-independent components have a
// Hessian here (checked on the host).
template <class Spec>
__device__ __forceinline__ GhVec3 gh_kick_adjoint(const GhPotTable* __restrict__ T, GhVec3 q,
                                                  double h, GhVec3 lam_p, GhVec3& lam_q,
                                                  double (*dth)[GH_ADJ_NPAR]) {
  const double r2 = q.x * q.x + q.y * q.y + q.z * q.z;
  const double rinv = rsqrt(r2);
  GhVec3 a{0.0, 0.0, 0.0}, k{0.0, 0.0, 0.0};
  double d[GH_MAX_POT][GH_ADJ_NPAR] = {};
  GhAdjSum<Spec>::add(T, q, r2, rinv, lam_p, a, k, d);
  lam_q.x += h * k.x;
  lam_q.y += h * k.y;
  lam_q.z += h * k.z;
#pragma unroll
  for (int i = 0; i < GH_MAX_POT; i++)
#pragma unroll
    for (int j = 0; j < GH_ADJ_NPAR; j++) dth[i][j] += h * d[i][j];
  return a;
}
