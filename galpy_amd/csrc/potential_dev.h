// Device-side potential table and force kernels (FP64, one orbit per lane).
//
// galpy packs a potential as (npot, pot_type[], pot_args[])
// (galpy/orbit/integrateFullOrbit.py:43-614) and its C driver turns that into
// a per-thread array of function-pointer records that may cache values
// (galpy/orbit/orbit_c_ext/integrateFullOrbit.c:69-1045,
// galpy/potential/potential_c_ext/galpy_potentials.h:16-120).
// Here the packed arrays are decoded once on the host into a compact,
// read-only table in device memory.  Kernels index it with wavefront-uniform
// addresses, so every parameter is fetched with scalar loads through the
// scalar cache and broadcast to all 64 lanes.  Constants that galpy
// recomputes on every call (Gamma(a), 2 pi rc^(3-alpha), 1/rc^2, the series
// coefficients of the incomplete gamma function) are precomputed.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

#include "gh_math.h"

#define GH_MAX_POT 8
#define GH_MAX_WRAP 8      // wrappers per potential
#define GH_WRAP_DEPTH 3    // wrappers nested around one component
#define GH_IG_NSER 32  // terms of the incomplete-gamma series (enough for a <= 1.5)
// piecewise-polynomial table of Gamma(a,x) on [a+1, xbig): GH_TAB_PER_OCT
// segments per binary octave of x, degree GH_TAB_NC-1 in the local coordinate
#define GH_TAB_PER_OCT 8
#define GH_TAB_MAXSEG 64
#define GH_TAB_NC 11
#define GH_TAB_SLOTS 2  // PowerSphericalPotentialwCutoff components with a table
// variable-length per-component data (quadrature nodes, rotations, series
// and spline tables): doubles shared by all components of a table.  Host
// tables live on the heap and device tables are allocated to the used
// prefix (gh_table_bytes), so the bound costs nothing when unused; it fits
// a DoubleExponentialDiskPotential at galpy's default de_n = 10000 (4 x
// 10000 nodes and weights) with room to spare.
#define GH_EXT_LEN (1 << 24)

enum GhPotCode {
  GH_LOGHALO = 0,
  GH_DEHNENBAR = 1,
  GH_MIYAMOTONAGAI = 5,
  GH_POWERSPH = 7,
  GH_HERNQUIST = 8,
  GH_NFW = 9,
  GH_JAFFE = 10,
  GH_ISOCHRONE = 14,
  GH_POWERSPHCUTOFF = 15,
  GH_PLUMMER = 17,
  GH_PSEUDOISO = 18,
  GH_KUZMINDISK = 19,
  GH_BURKERT = 20,
  GH_DEHNENCORESPH = 33,
  GH_DEHNENSPH = 34,
  GH_HOMOGSPHERE = 35,
  GH_FLATPOWER = 12,
  GH_KKSTAECKEL = 16,
  GH_ELLIPDISK = 4,  // planar only (integratePlanarOrbit's code 4)
  GH_LOPSIDED = 6,   // planar only (integratePlanarOrbit's code 6)
  GH_STEADYSPIRAL = 3,  // planar only (integratePlanarOrbit's code 3)
  GH_TRANSIENTSPIRAL = 2,  // planar only (integratePlanarOrbit's code 2)
  GH_COSMPHIDISK = 28,     // planar only (integratePlanarOrbit's code 28)
  GH_HENONHEILES = 29,     // planar only (integratePlanarOrbit's code 29)
  GH_SPIRALARMS = 27,
  GH_DEDISK = 11,  // DoubleExponentialDiskPotential
  GH_SOFTNEEDLEBAR = 25,
  GH_NULLPOT = 40,
  // EllipsoidalPotential family (integrateFullOrbit.c:420-483, 549, 647, 668, 764)
  GH_TRIAXHERNQUIST = 21,
  GH_TRIAXNFW = 22,
  GH_TRIAXJAFFE = 23,
  GH_PERFECTELLIPSOID = 30,
  GH_TRIAXGAUSSIAN = 37,
  GH_POWERTRIAX = 38,
  GH_TWOPOWERTRIAX = 43,
  // spherical components with special functions (integrateFullOrbit.c:731-780)
  GH_EINASTO = 41,
  GH_TWOPOWERSPH = 42,
  GH_EXPTRUNCNFW = 46,
  // wrappers (integrateFullOrbit.c:821-895)
  GH_DEHNENSMOOTHWRAP = -1,
  GH_SOLIDBODYROTWRAP = -2,
  GH_COROTATINGROTWRAP = -4,
  GH_GAUSSIANAMPWRAP = -5,
  GH_ROTATETILTWRAP = -8,
  GH_KUZMINLIKEWRAP = -10,
  GH_CYLSEPWRAP = -12,
  GH_TIMEDEPAMPWRAP = -9,
  GH_MOVINGOBJWRAP = -6,
  GH_INTERPSPH = 36,
  GH_SCF = 24,
  GH_DISKSCF = 26,
  GH_INTERPRZ = 13,
  GH_MULTIPOLE = 44,
  GH_OBLATESTAECKELWRAP = -3,
  // velocity-dependent forces (integrateFullOrbit.c:689-721)
  GH_NONINERTIAL = 39,
  GH_NONINERTIALINTERP = 45,
  GH_CHANDRASEKHAR = -7,
  GH_FDM = -11,
};

struct GhComponent {
  int type;
  int nonaxi;    // 1 if the component exerts a phi torque
  int nwrap;     // wrappers around this component
  int wrap[GH_WRAP_DEPTH];  // their indices in GhPotTable::wrap, outermost first
  int ext;       // offset of the component's variable-length data in GhPotTable::ext
  int mapwrap;   // wrappers that move the evaluation point: 0 none, else passes (1 or 2)
  int mappot;    // the mapped walk needs the component's potential too (OblateStaeckel)
  int densonly;  // only a density (for dynamical friction): no force, no potential
  double p[8];   // raw galpy arguments, in galpy's order
  double c[14];  // precomputed constants, per type (decoded on the host)
};

// A wrapper (DehnenSmoothWrapperPotential.c, SolidBodyRotationWrapperPotential.c,
// CorotatingRotationWrapperPotential.c, GaussianAmplitudeWrapperPotential.c):
// its code and its own arguments in galpy's order.  galpy nests wrappers
// around lists of components; the host flattens the nesting so that every
// leaf component lists the wrappers around it.
struct GhWrap {
  int kind;
  int first;     // index of the first component it wraps (its constant terms are added there)
  double p[16];  // RotateAndTilt: amp, rot[9], rotSet, offsetSet, offset[3]
};

struct GhPotTable {
  int npot;
  int nonaxi;    // some component needs phi
  int has_wrap;  // some component is wrapped or heavy (extended dispatch only)
  int nwraps;
  GhComponent comp[GH_MAX_POT];
  GhWrap wrap[GH_MAX_WRAP];
  // series coefficients 1/(a (a+1) ... (a+n)) of the incomplete gamma function
  // for the force exponent a = 1.5 - alpha/2 ([0]) and the potential-value
  // exponent 1 - alpha/2 ([1]) of PowerSphericalPotentialwCutoff components
  double ser[GH_MAX_POT][2][GH_IG_NSER];
  // Gamma(a,x) tables for the force exponent of up to GH_TAB_SLOTS
  // PowerSphericalPotentialwCutoff components (slot in comp.c[11])
  double tab[GH_TAB_SLOTS][GH_TAB_MAXSEG][GH_TAB_NC];
  // variable-length data, last so that a table is uploaded up to next
  int next;  // doubles of ext in use
  int ndiss;  // velocity-dependent components (their forces need the velocity)
  double ext[GH_EXT_LEN];
};

// bytes of a table that carry data (the prefix a host-to-device copy needs)
__host__ __device__ inline size_t gh_table_bytes(const GhPotTable& T) {
  return offsetof(GhPotTable, ext) + sizeof(double) * (size_t)T.next;
}
// device allocation for a table: its data, rounded up to 256 bytes
__host__ __device__ inline size_t gh_table_alloc_bytes(const GhPotTable& T) {
  return (gh_table_bytes(T) + 255) & ~(size_t)255;
}

// GSL's natural cubic spline (gsl_interp_cspline, which galpy's C builds for
// interpSphericalPotential and MovingObjectPotential), coefficients computed
// on the host: ext[off] = n, x[0..n-1], then per interval i < n-1 the value
// y_i, b_i, c_i, d_i (y = y_i + d (b_i + d (c_i + d d_i)), d = x - x_i) and
// the integral I_i from x[0] to x[i].  The interval by binary search.
__device__ __forceinline__ int gh_spline_find(const double* __restrict__ x, int n, double t) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (x[mid] > t) hi = mid;
    else lo = mid;
  }
  return lo;
}
__device__ __forceinline__ double gh_spline_eval(const GhPotTable* __restrict__ T, int off,
                                                 double t) {
  const double* __restrict__ e = &T->ext[off];
  const int n = (int)e[0];
  const int i = gh_spline_find(e + 1, n, t);
  const double* __restrict__ q = e + 1 + n + 5 * i;
  const double d = t - e[1 + i];
  return fma(d, fma(d, fma(d, q[3], q[2]), q[1]), q[0]);
}
// derivative of the spline (gsl_spline_eval_deriv: b_i + d (2 c_i + 3 d_i d))
__device__ __forceinline__ double gh_spline_deriv(const GhPotTable* __restrict__ T, int off,
                                                  double t) {
  const double* __restrict__ e = &T->ext[off];
  const int n = (int)e[0];
  const int i = gh_spline_find(e + 1, n, t);
  const double* __restrict__ q = e + 1 + n + 5 * i;
  const double d = t - e[1 + i];
  return q[1] + d * (2.0 * q[2] + 3.0 * q[3] * d);
}
// integral of the spline from x[0] to t
__device__ __forceinline__ double gh_spline_integ(const GhPotTable* __restrict__ T, int off,
                                                  double t) {
  const double* __restrict__ e = &T->ext[off];
  const int n = (int)e[0];
  const int i = gh_spline_find(e + 1, n, t);
  const double* __restrict__ q = e + 1 + n + 5 * i;
  const double d = t - e[1 + i];
  return q[4] + d * fma(d, fma(d, fma(d, 0.25 * q[3], q[2] / 3.0), 0.5 * q[1]), q[0]);
}


// ---------------------------------------------------------------------------
// Lower incomplete gamma gamma(a, x) (replaces GSL's gsl_sf_gamma_inc_P and
// gsl_sf_gamma_inc at PowerSphericalPotentialwCutoff.c:10,23).
//   x < a+1 : gamma(a,x) = e^-x x^a sum_n x^n / (a (a+1) ... (a+n))
//             (A&S 6.5.29), Horner over host-computed coefficients: no
//             divisions, fixed trip count.
//   x >= a+1: Gamma(a,x) by Legendre's continued fraction (A&S 6.5.31)
//             1/(b1 - 1(1-a)/(b2 - 2(2-a)/(b3 - ...))), b_n = x + 2n - 1 - a,
//             evaluated with the forward (Wallis) recurrence on numerators
//             and denominators: 4 FMAs per term and a single division at the
//             end, instead of the two divisions per term of Lentz's method.
//             The loop stops once successive convergents differ by less than
//             2^-58 Gamma(a) in Gamma(a,x) (tested without dividing), so it
//             takes ~29 terms at x = 4, ~11 at x = 10 and 1-3 past x = 20;
//             gamma = Gamma(a) - Gamma(a,x).  Numerators and denominators are
//             rescaled by 2^-400 (exactly) if they grow past 2^400.
//   x >= xbig: e^-x x^(a-1) < 2^-60 Gamma(a), so the upper tail cannot change
//             the rounded result and gamma = Gamma(a).
// cf_tol = 2^-58 Gamma(a) is precomputed on the host.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double gh_incgamma_cf(double a, double x, double thr);

__device__ __forceinline__ double gh_lower_incgamma(double a, double x, double gamma_a,
                                                   double xbig, double cf_tol,
                                                   const double* __restrict__ ser) {
  if (x >= xbig) return gamma_a;
  if (x <= 0.0) return 0.0;
  const double pref = exp(a * log(x) - x);
  if (x < a + 1.0) {
    double s = ser[GH_IG_NSER - 1];
#pragma unroll
    for (int n = GH_IG_NSER - 2; n >= 0; n--) s = fma(s, x, ser[n]);
    return pref * s;
  }
  return gamma_a - pref * gh_incgamma_cf(a, x, cf_tol / pref);
}

// Legendre's continued fraction f of Gamma(a, x) = e^-x x^a f (x >= a+1),
// by the forward recurrence described above; thr = allowed change of f.
__device__ __forceinline__ double gh_incgamma_cf(double a, double x, double thr) {
  double b = x + 1.0 - a;
  double Am2 = 0.0, Bm2 = 1.0;  // A_0, B_0
  double Am1 = 1.0, Bm1 = b;    // A_1, B_1: f_1 = 1/b_1
  for (int n = 1; n < 400; n++) {
    const double an = -n * (n - a);
    b += 2.0;
    const double A = fma(b, Am1, an * Am2);
    const double B = fma(b, Bm1, an * Bm2);
    // |A/B - Am1/Bm1| * |B Bm1| = |A Bm1 - Am1 B|
    const double diff = fma(A, Bm1, -Am1 * B);
    const double den = B * Bm1;
    Am2 = Am1;
    Bm2 = Bm1;
    Am1 = A;
    Bm1 = B;
    if (fabs(diff) <= thr * fabs(den)) break;
    if (fabs(B) > 0x1p400) {
      Am2 *= 0x1p-400;
      Bm2 *= 0x1p-400;
      Am1 *= 0x1p-400;
      Bm1 *= 0x1p-400;
    }
  }
  return Am1 / Bm1;
}

// gamma(a, x) for any a > 0 with the series summed term by term (x < a+1,
// stopping below 2^-56 of the sum) instead of host coefficients; the same
// continued fraction above.
__device__ inline double gh_lower_incgamma_gen(double a, double x, double gamma_a, double cf_tol) {
  if (x <= 0.0) return 0.0;
  const double pref = exp(a * log(x) - x);
  if (x < a + 1.0) {
    double term = 1.0 / a, sum = term, ap = a;
    for (int n = 1; n < 2000; n++) {
      ap += 1.0;
      term *= x / ap;
      sum += term;
      if (term < sum * 0x1p-56) break;
    }
    return pref * sum;
  }
  return gamma_a - pref * gh_incgamma_cf(a, x, cf_tol / pref);
}

// e^x E1(x), x > 0 (replaces GSL's gsl_sf_expint_E1 at
// ExpTruncNFWPotential.c:25,32): A&S 5.1.11 below x = 1,
// -gamma_E - ln x - sum_k (-x)^k / (k k!), and the continued fraction
// A&S 5.1.22, 1/(x + 1 - 1/(x + 3 - 4/(x + 5 - ...))), by Lentz's method above.
__device__ inline double gh_e1_scaled(double x) {
  if (x <= 1.0) {
    double sum = 0.0, term = 1.0;
    for (int k = 1; k < 60; k++) {
      term *= -x / k;
      const double add = term / k;
      sum += add;
      if (fabs(add) < fabs(sum) * 0x1p-56) break;
    }
    return exp(x) * (-0.57721566490153286061 - log(x) - sum);
  }
  double b = x + 1.0, c = 1e300, d = 1.0 / b, h = d;
  for (int i = 1; i < 1000; i++) {
    const double an = -(double)i * i;
    b += 2.0;
    d = 1.0 / fma(an, d, b);
    c = b + an / c;
    const double del = c * d;
    h *= del;
    if (fabs(del - 1.0) < 0x1p-53) break;
  }
  return h;
}

// Series branch of gh_lower_incgamma (x < a+1, i.e. r < rc sqrt(a+1)), kept
// out of line: orbits rarely enter it, and every inlined copy of the force
// would otherwise carry its exp/log code through the instruction cache.
__device__ __noinline__ double gh_lower_incgamma_series(double a, double x,
                                                       const double* __restrict__ ser) {
  if (x <= 0.0) return 0.0;
  const double pref = exp(a * log(x) - x);
  double s = ser[GH_IG_NSER - 1];
#pragma unroll
  for (int n = GH_IG_NSER - 2; n >= 0; n--) s = fma(s, x, ser[n]);
  return pref * s;
}

// gamma(a,x) for the force kernel.  On [a+1, xbig) the upper tail
// Gamma(a,x) comes from a host-fitted table instead of the continued
// fraction: segment boundaries split each binary octave of x into
// GH_TAB_PER_OCT equal parts, so every segment lies at least 8 half-widths
// from the branch point of Gamma(a,x) at x = 0 and a degree-10 polynomial in
// the local coordinate reproduces it to ~1e-17 Gamma(a).  The segment index
// comes from the exponent and mantissa bits of x (frexp), so the evaluation
// is a gather of 11 coefficients and 10 FMAs, with no exp/log and no
// division.  Below a+1 the series branch of gh_lower_incgamma is used.
// c[0] = a, c[1] = Gamma(a), c[4] = xbig, c[9] = continued-fraction
// tolerance, c[11] = table slot, c[12] = frexp exponent of the first octave.
__device__ __forceinline__ double gh_lower_incgamma_tab(const GhComponent* __restrict__ C,
                                                       const double* __restrict__ ser,
                                                       const double* __restrict__ tab,
                                                       double x) {
  const double a = C->c[0];
  if (x >= C->c[4]) return C->c[1];
  if (x < a + 1.0) return gh_lower_incgamma_series(a, x, ser);
  int e;
  const double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  const double u = (m - 0.5) * (2.0 * GH_TAB_PER_OCT);  // in [0, GH_TAB_PER_OCT)
  int k = (int)u;
  k = k > GH_TAB_PER_OCT - 1 ? GH_TAB_PER_OCT - 1 : k;
  const int seg = (e - (int)C->c[12]) * GH_TAB_PER_OCT + k;
  const double t = 2.0 * (u - k) - 1.0;  // local coordinate in [-1, 1)
  const double* __restrict__ c = tab + seg * GH_TAB_NC;
  double v = c[GH_TAB_NC - 1];
#pragma unroll
  for (int n = GH_TAB_NC - 2; n >= 0; n--) v = fma(v, t, c[n]);
  return C->c[1] - v;
}

// DehnenBarPotential.c:5-17 (quintic turn-on of the bar)
__device__ __forceinline__ double gh_bar_smooth(double t, double tform, double tsteady) {
  if (t < tform) return 0.0;
  if (t < tsteady) {
    const double xi = 2.0 * (t - tform) / (tsteady - tform) - 1.0;
    const double xi2 = xi * xi;
    const double xi3 = xi2 * xi;
    return 3.0 / 16.0 * xi3 * xi2 - 5.0 / 8.0 * xi3 + 15.0 / 16.0 * xi + 0.5;
  }
  return 1.0;
}

// cos/sin of 2(phi - omegab t - barphi) from cos(2 phi), sin(2 phi) by angle
// addition (galpy forms phi with acos, integrateFullOrbit.c:1316-1320)
__device__ __forceinline__ void gh_bar_angle(double cos2p, double sin2p, double omegab,
                                             double barphi, double t, double& cang,
                                             double& sang) {
  const GhSinCos b = gh_sincos(2.0 * (omegab * t + barphi));
  const double sb = b.s, cb = b.c;
  cang = cos2p * cb + sin2p * sb;
  sang = sin2p * cb - cos2p * sb;
}

// 1/sqrt(x) for positive normal x: the hardware estimate refined by one
// Newton-Raphson step with its second-order term, y (1 + e/2 + 3e^2/8),
// e = 1 - x y^2 -- the library rsqrt's arithmetic without its selects for
// zero, infinite and denormal arguments (a force argument is never one).
__device__ __forceinline__ double gh_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// 1/x for positive normal x: hardware estimate plus two Newton-Raphson steps
// (within an ulp), instead of the correctly rounded division sequence.
__device__ __forceinline__ double gh_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Forces are accumulated in the form the rectangular acceleration needs:
// F_R / R, F_z and F_phi / R^2.  For every component here F_R carries a
// factor R and F_phi a factor R^2 (or x y), so
//   a_x = x (F_R/R) - y (F_phi/R^2),  a_y = y (F_R/R) + x (F_phi/R^2),
// which is evalRectForce (integrateFullOrbit.c:1307-1325) without forming R,
// phi, cos(phi) or sin(phi): no square root of x^2 + y^2, and 1/R^2 only for
// the non-axisymmetric components.
struct GhAcc {
  double FRoR, Fz, FphioR2;
};

// Geometry shared by all components of one force evaluation.
struct GhGeom {
  double x, y, z, R2, z2, r2, rinv, r, t;
};

// Force of one component of type TYPE (index i in the table), added to F.
// Component kernels: PowerSphericalPotentialwCutoff.c:9-81,
// MiyamotoNagaiPotential.c:16-53, NFWPotential.c:16-49,
// HernquistPotential.c:15-46, LogarithmicHaloPotential.c:21-91,
// DehnenBarPotential.c:18-124.
template <int TYPE>
__device__ __forceinline__ void gh_comp_force(const GhPotTable* __restrict__ T, int i,
                                              const GhGeom& g, GhAcc& F);

template <>
__device__ __forceinline__ void gh_comp_force<GH_POWERSPHCUTOFF>(const GhPotTable* __restrict__ T,
                                                                 int i, const GhGeom& g, GhAcc& F) {
  // F = -amp M(r2) / r^3 (R, z), M = 2 pi rc^(3-alpha) gamma(a, r2/rc^2)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double x = g.r2 * C->c[3];
  const double gam = gh_lower_incgamma_tab(C, T->ser[i][0], &T->tab[(int)C->c[11]][0][0], x);
  const double f = -C->c[2] * gam * (g.rinv * g.rinv * g.rinv);
  F.FRoR += f;
  F.Fz += f * g.z;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_MIYAMOTONAGAI>(const GhPotTable* __restrict__ T,
                                                                int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], a = C->p[1];
  const double s2 = g.z2 + C->c[0];
  const double isbz = gh_rsqrt(s2);
  const double asqrtbz = a + s2 * isbz;
  const double D = g.R2 + asqrtbz * asqrtbz;
  const double iD = gh_rsqrt(D);
  const double Dm32 = iD * iD * iD;
  F.FRoR += -amp * Dm32;
  F.Fz += (a == 0.0) ? -amp * g.z * Dm32 : -amp * g.z * asqrtbz * isbz * Dm32;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_NFW>(const GhPotTable* __restrict__ T, int i,
                                                      const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], a = C->p[1];
  const double rinv2 = g.rinv * g.rinv;
  // log(1 + r/a) as NFWPotential.c:26 forms it, with the short log of gh_math.h
  const double br = rinv2 * gh_rcp(a + g.r) - gh_log_ge1(1.0 + g.r * C->c[0]) * (rinv2 * g.rinv);
  F.FRoR += amp * br;
  F.Fz += amp * g.z * br;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_HERNQUIST>(const GhPotTable* __restrict__ T, int i,
                                                            const GhGeom& g, GhAcc& F) {
  // F = -amp x / (2 a^2 r) (1 + r/a)^-2 = -(amp/2) x / (r (a + r)^2)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double apr = C->p[1] + g.r;
  const double f = -C->c[0] * g.rinv * gh_rcp(apr * apr);
  F.FRoR += f;
  F.Fz += f * g.z;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_LOGHALO>(const GhPotTable* __restrict__ T, int i,
                                                          const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], core2 = C->p[2], onem1overb2 = C->p[3];
  const double zq2 = g.z2 * C->c[0];  // (z/q)^2
  // Each branch fills scalars and F is updated once below: conditional
  // updates of F's fields make the compiler select a field address, which
  // keeps the accumulator in scratch memory.
  double fr, fz, fp;
  if (onem1overb2 < 1.0) {
    // R_t^2 = R^2 (1 - (1 - 1/b^2) sin^2 phi) = R^2 - (1 - 1/b^2) y^2
    const double Rt2 = g.R2 - onem1overb2 * g.y * g.y;
    const double inv = gh_rcp(Rt2 + zq2 + core2);
    const double iR2 = gh_rcp(g.R2);
    fr = -amp * Rt2 * inv * iR2;
    fz = -amp * g.z * C->c[0] * inv;
    // F_phi = amp R^2 inv sin(phi) cos(phi) (1 - 1/b^2) = amp inv x y (1 - 1/b^2)
    fp = amp * inv * g.x * g.y * onem1overb2 * iR2;
  } else {
    const double inv = gh_rcp(g.R2 + zq2 + core2);
    fr = -amp * inv;
    fz = -amp * g.z * C->c[0] * inv;
    fp = 0.0;
  }
  F.FRoR += fr;
  F.Fz += fz;
  F.FphioR2 += fp;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_DEHNENBAR>(const GhPotTable* __restrict__ T, int i,
                                                            const GhGeom& g, GhAcc& F) {
  // args amp*Af, tform, tsteady, rb, omegab, barphi
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], rb = C->p[3];
  const double smooth = gh_bar_smooth(g.t, C->p[1], C->p[2]);
  const double iR2 = gh_rcp(g.R2);
  const double cos2p = (g.x * g.x - g.y * g.y) * iR2;  // cos(2 phi)
  const double sin2p = 2.0 * g.x * g.y * iR2;          // sin(2 phi)
  double cang, sang;
  gh_bar_angle(cos2p, sin2p, C->p[4], C->p[5], g.t, cang, sang);
  const double r2inv = g.rinv * g.rinv;
  const double r4inv = r2inv * r2inv;
  const double as = amp * smooth;
  double fr, fp, fz;  // filled per branch, added once (see the logarithmic halo)
  if (g.r <= rb) {
    const double u = g.r * C->c[0];  // r/rb
    const double u3 = u * u * u;
    fr = -as * cang * (u3 * (3.0 * g.R2 + 2.0 * g.z2) - 4.0 * g.z2) * r4inv;
    fp = 2.0 * as * sang * (u3 - 2.0) * r2inv;
    fz = -as * cang * (u3 + 4.0) * g.R2 * g.z * r4inv;
  } else {
    const double u = rb * g.rinv;
    const double u3 = u * u * u;
    fr = -as * cang * u3 * r4inv * (3.0 * g.R2 - 2.0 * g.z2);
    fp = -2.0 * as * sang * u3 * r2inv;
    fz = -5.0 * as * cang * u3 * g.R2 * g.z * r4inv;
  }
  F.FRoR += fr;
  F.FphioR2 += fp;
  F.Fz += fz;
}

// ---------------------------------------------------------------------------
// Spherical components (PowerSphericalPotential.c, JaffePotential.c,
// IsochronePotential.c, PlummerPotential.c, PseudoIsothermalPotential.c,
// BurkertPotential.c, DehnenCorePotential.c, DehnenPotential.c,
// HomogeneousSpherePotential.c): F = g(r) x with g = -Phi'(r)/r, and
// gpr = g'(r)/r for Hessian-vector products (K v = g v + gpr (x.v) x).
// Each file writes Rforce = -Phi'(r) R/r with its own closed form of Phi'.
// ---------------------------------------------------------------------------
template <int TYPE>
__device__ __forceinline__ void gh_sph_g(const GhComponent* __restrict__ C, double r, double rinv,
                                         double& g, double& gpr);

template <>
__device__ __forceinline__ void gh_sph_g<GH_POWERSPH>(const GhComponent* __restrict__ C, double r,
                                                      double rinv, double& g, double& gpr) {
  // Phi' = amp r^(1-alpha)
  const double amp = C->p[0], alpha = C->p[1];
  const double ra = exp(-alpha * log(r));
  g = -amp * ra;
  gpr = alpha * amp * ra * rinv * rinv;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_JAFFE>(const GhComponent* __restrict__ C, double r,
                                                   double rinv, double& g, double& gpr) {
  // Phi' = amp / (r (r + a))
  const double amp = C->p[0], a = C->p[1];
  const double iar = gh_rcp(a + r);
  const double ri2 = rinv * rinv;
  g = -amp * ri2 * iar;
  gpr = amp * (3.0 * r + 2.0 * a) * ri2 * ri2 * iar * iar;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_ISOCHRONE>(const GhComponent* __restrict__ C, double r,
                                                       double rinv, double& g, double& gpr) {
  // Phi' = amp r / (rb (b + rb)^2), rb = sqrt(r^2 + b^2)
  const double amp = C->p[0], b = C->p[1];
  const double irb = gh_rsqrt(r * r + C->c[0]);
  const double rb = (r * r + C->c[0]) * irb;
  const double ib = gh_rcp(b + rb);
  g = -amp * irb * ib * ib;
  gpr = amp * (b + 3.0 * rb) * irb * irb * irb * ib * ib * ib;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_PLUMMER>(const GhComponent* __restrict__ C, double r,
                                                     double rinv, double& g, double& gpr) {
  // Phi' = amp r (r^2 + b^2)^(-3/2)
  const double amp = C->p[0];
  const double is = gh_rsqrt(r * r + C->c[0]);
  const double is3 = is * is * is;
  g = -amp * is3;
  gpr = 3.0 * amp * is3 * is * is;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_PSEUDOISO>(const GhComponent* __restrict__ C, double r,
                                                       double rinv, double& g, double& gpr) {
  // Phi' = amp (1/r - a atan(r/a)/r^2) / a, Phi'' from planarR2deriv;
  // gpr = (Phi' - r Phi'') / r^3
  const double amp = C->p[0], a = C->p[1], ia = C->c[0];
  const double at = atan(r * ia);
  const double Phip = amp * (rinv - a * rinv * rinv * at) * ia;
  const double Phipp =
      amp * rinv * rinv * (2.0 * a * rinv * at - (2.0 * a * a + r * r) / (a * a + r * r)) * ia;
  g = -Phip * rinv;
  gpr = (Phip - r * Phipp) * rinv * rinv * rinv;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_BURKERT>(const GhComponent* __restrict__ C, double r,
                                                     double rinv, double& g, double& gpr) {
  // Phi' = -amp a pi B / x^2, B = pi - 2 atan(1/x) - 2 log(1+x) - log(1+x^2),
  // x = r/a; Phi'' from planarR2deriv
  const double amp = C->p[0], a = C->p[1];
  const double x = r * C->c[0];
  const double B = M_PI - 2.0 * atan(1.0 / x) - 2.0 * log1p(x) - log1p(x * x);
  const double Phip = -amp * a * M_PI * B / (x * x);
  const double r5 = r * r * r * r * r;
  const double Phipp =
      -amp * M_PI * a * a * a / r5 * (-4.0 * r5 / (a * a + r * r) / (a + r) - 2.0 * r * r * B);
  g = -Phip * rinv;
  gpr = (Phip - r * Phipp) * rinv * rinv * rinv;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_DEHNENCORESPH>(const GhComponent* __restrict__ C,
                                                           double r, double rinv, double& g,
                                                           double& gpr) {
  // Phi' = amp r / (3 (a + r)^3)
  const double amp = C->p[0], a = C->p[1];
  const double iar = gh_rcp(a + r);
  const double iar3 = iar * iar * iar;
  g = -amp * iar3 * (1.0 / 3.0);
  gpr = amp * rinv * iar3 * iar;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_DEHNENSPH>(const GhComponent* __restrict__ C, double r,
                                                       double rinv, double& g, double& gpr) {
  // Phi' = amp r (1 + a/r)^alpha / ((a + r)^3 (3 - alpha));
  // g'/g = -(alpha a + 3 r) / (r (r + a))
  const double amp = C->p[0], a = C->p[1], alpha = C->p[2];
  const double iar = gh_rcp(a + r);
  const double fac = exp(alpha * log(1.0 + a * rinv));
  g = -amp * fac * iar * iar * iar * C->c[0];
  gpr = -g * (alpha * a + 3.0 * r) * rinv * rinv * iar;
}

template <>
__device__ __forceinline__ void gh_sph_g<GH_HOMOGSPHERE>(const GhComponent* __restrict__ C,
                                                         double r, double rinv, double& g,
                                                         double& gpr) {
  // inside (r^2 < R^2): Phi' = 2 amp r; outside: Phi' = 2 amp R^3 / r^2
  const double amp = C->p[0], R2 = C->p[1], R3 = C->p[2];
  if (r * r < R2) {
    g = -2.0 * amp;
    gpr = 0.0;
  } else {
    const double ri3 = rinv * rinv * rinv;
    g = -2.0 * amp * R3 * ri3;
    gpr = 6.0 * amp * R3 * ri3 * rinv * rinv;
  }
}

#define GH_SPH_FORCE(CODE)                                                                  \
  template <>                                                                               \
  __device__ __forceinline__ void gh_comp_force<CODE>(const GhPotTable* __restrict__ T, int i, \
                                                      const GhGeom& g, GhAcc& F) {          \
    double gr, gpr;                                                                         \
    gh_sph_g<CODE>(&T->comp[i], g.r, g.rinv, gr, gpr);                                      \
    F.FRoR += gr;                                                                           \
    F.Fz += gr * g.z;                                                                       \
  }
GH_SPH_FORCE(GH_POWERSPH)
GH_SPH_FORCE(GH_JAFFE)
GH_SPH_FORCE(GH_ISOCHRONE)
GH_SPH_FORCE(GH_PLUMMER)
GH_SPH_FORCE(GH_PSEUDOISO)
GH_SPH_FORCE(GH_BURKERT)
GH_SPH_FORCE(GH_DEHNENCORESPH)
GH_SPH_FORCE(GH_DEHNENSPH)
GH_SPH_FORCE(GH_HOMOGSPHERE)
#undef GH_SPH_FORCE

template <>
__device__ __forceinline__ void gh_comp_force<GH_KUZMINDISK>(const GhPotTable* __restrict__ T,
                                                             int i, const GhGeom& g, GhAcc& F) {
  // KuzminDiskPotential.c:16-37: Phi = -amp (R^2 + (a + |z|)^2)^(-1/2);
  // zforce carries sign(z) (0 in the plane)
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0];
  const double az = C->p[1] + fabs(g.z);
  const double iD = gh_rsqrt(g.R2 + az * az);
  const double D32 = iD * iD * iD;
  const double w = (g.z == 0.0) ? 0.0 : copysign(az, g.z);
  F.FRoR += -amp * D32;
  F.Fz += -amp * w * D32;
}

// SoftenedNeedleBarPotential.c:5-12, 30-63 (Long & Murali 1992): in the frame
// rotated by pa + omegab t (x_b = R cos(phi - pa - omegab t)), with T+- =
// sqrt((a +- x_b)^2 + y_b^2 + (b + sqrt(z^2 + c^2))^2), Phi = amp/(2a)
// log((x_b - a + T-) / (x_b + a + T+)); the bar-frame force is rotated back
// and added as F_R/R and F_phi/R^2 of the inertial (x, y).  p: amp, a, b,
// c^2, pa, omegab (the reference's seven cache slots are not read).
__device__ __forceinline__ void gh_needle_frame(const GhComponent* __restrict__ C, double x,
                                                double y, double t, double& cp, double& sp,
                                                double& xb, double& yb) {
  const GhSinCos a = gh_sincos(C->p[4] + C->p[5] * t);
  sp = a.s;
  cp = a.c;
  xb = cp * x + sp * y;
  yb = cp * y - sp * x;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_SOFTNEEDLEBAR>(const GhPotTable* __restrict__ T,
                                                                int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], a = C->p[1], b = C->p[2];
  double cp, sp, xb, yb;
  gh_needle_frame(C, g.x, g.y, g.t, cp, sp, xb, yb);
  const double zc = sqrt(g.z2 + C->p[3]);
  const double bz = b + zc;
  const double second = yb * yb + bz * bz;
  const double Tp = sqrt((a + xb) * (a + xb) + second), Tm = sqrt((a - xb) * (a - xb) + second);
  const double S = Tp + Tm, iTT = 1.0 / (Tp * Tm);
  const double Fx = -2.0 * xb * iTT / S;
  const double common = 0.5 * iTT * (S - 4.0 * xb * xb / S) / second;
  const double Fy = -yb * common;
  const double Ax = amp * (cp * Fx - sp * Fy), Ay = amp * (sp * Fx + cp * Fy);
  const double iR2 = gh_rcp(g.R2);
  F.FRoR += (g.x * Ax + g.y * Ay) * iR2;
  F.FphioR2 += (g.x * Ay - g.y * Ax) * iR2;
  F.Fz += -amp * g.z * common * bz / zc;
}

// EllipticalDiskPotential.c:4-52 (planar only): Phi = amp smooth(t)
// twophio/2 R^p cos(2 (phi - phib)), smooth Dehnen's growth polynomial
// between tform and tsteady (1 when tform is NaN, i.e. None in galpy).
// p: amp (divided by r1^p in galpy), tform, tsteady, twophio, p, phib;
// c[0] = cos(2 phib), c[1] = sin(2 phib).
__device__ __forceinline__ double gh_ellipdisk_smooth(double t, double tform, double tsteady) {
  if (tform != tform) return 1.0;
  if (t < tform) return 0.0;
  if (t < tsteady) {
    const double xi = 2.0 * (t - tform) / (tsteady - tform) - 1.0;
    const double xi3 = xi * xi * xi;
    return 3.0 / 16.0 * xi3 * xi * xi - 5.0 / 8.0 * xi3 + 15.0 / 16.0 * xi + 0.5;
  }
  return 1.0;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_ELLIPDISK>(const GhPotTable* __restrict__ T, int i,
                                                            const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double as = C->p[0] * gh_ellipdisk_smooth(g.t, C->p[1], C->p[2]);
  const double p = C->p[4];
  const double iR2 = gh_rcp(g.R2);
  const double cos2p = (g.x * g.x - g.y * g.y) * iR2, sin2p = 2.0 * g.x * g.y * iR2;
  const double cosd = cos2p * C->c[0] + sin2p * C->c[1];  // cos(2 (phi - phib))
  const double sind = sin2p * C->c[0] - cos2p * C->c[1];  // sin(2 (phi - phib))
  const double Rpm2 = pow(g.R2, 0.5 * p - 1.0);           // R^(p-2)
  F.FRoR += -as * p * C->p[3] * 0.5 * Rpm2 * cosd;
  F.FphioR2 += as * C->p[3] * Rpm2 * sind;
}

// LopsidedDiskPotential.c (planar only): Phi = amp phio R^p cos(phi - phib);
// p: amp, phio, p, phib; c[0] = cos(phib), c[1] = sin(phib).
template <>
__device__ __forceinline__ void gh_comp_force<GH_LOPSIDED>(const GhPotTable* __restrict__ T, int i,
                                                           const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double ap = C->p[0] * C->p[1], p = C->p[2];
  const double Rcos = g.x * C->c[0] + g.y * C->c[1];  // R cos(phi - phib)
  const double Rsin = g.y * C->c[0] - g.x * C->c[1];  // R sin(phi - phib)
  const double Rpm3 = pow(g.R2, 0.5 * p - 1.5);      // R^(p-3)
  F.FRoR += -ap * p * Rpm3 * Rcos;
  F.FphioR2 += ap * Rpm3 * Rsin;
}

// SteadyLogSpiralPotential.c (planar only): Phi = amp smooth(t) A/alpha
// cos(alpha ln R - m (phi - omegas t - gamma)), smooth the growth polynomial
// of gh_ellipdisk_smooth; p: amp, tform, tsteady, A, alpha, m, omegas, gamma.
__device__ __forceinline__ double gh_spiral_arg(const GhComponent* __restrict__ C, double R2,
                                                double phi, double t) {
  return C->p[4] * 0.5 * log(R2) - C->p[5] * (phi - C->p[6] * t - C->p[7]);
}
// phi as galpy's planar force evaluation forms it, acos(x/R) and 2 pi - that
// for y < 0 (integratePlanarOrbit.c:1022-1025): in [0, 2 pi), which decides
// the spiral's phase for a non-integer multiplicity m.
__device__ __forceinline__ double gh_phi_acos(double x, double y, double R2) {
  const double phi = acos(x / sqrt(R2));
  return (y < 0.0) ? 2.0 * M_PI - phi : phi;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_STEADYSPIRAL>(const GhPotTable* __restrict__ T,
                                                               int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double aA = C->p[0] * gh_ellipdisk_smooth(g.t, C->p[1], C->p[2]) * C->p[3];
  const double sn = gh_sincos(gh_spiral_arg(C, g.R2, gh_phi_acos(g.x, g.y, g.R2), g.t)).s;
  const double iR2 = gh_rcp(g.R2);
  F.FRoR += aA * sn * iR2;
  F.FphioR2 += -aA / C->p[4] * C->p[5] * sn * iR2;
}

// TransientLogSpiralPotential.c:5-42 (planar only): Phi = amp A
// exp(-(t - to)^2 / (2 sigma^2)) / alpha cos(alpha ln R - m (phi - omegas t
// - gamma)), the spiral's argument as gh_spiral_arg forms it (p[4..7] are
// alpha, m, omegas, gamma in both spirals); p: amp, A, to, sigma^2, alpha,
// m, omegas, gamma.
__device__ __forceinline__ double gh_transient_amp(const GhComponent* __restrict__ C, double t) {
  const double d = t - C->p[2];
  return C->p[0] * C->p[1] * exp(-d * d / 2.0 / C->p[3]);
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_TRANSIENTSPIRAL>(const GhPotTable* __restrict__ T,
                                                                  int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double aA = gh_transient_amp(C, g.t);
  const double sn = gh_sincos(gh_spiral_arg(C, g.R2, gh_phi_acos(g.x, g.y, g.R2), g.t)).s;
  const double iR2 = gh_rcp(g.R2);
  F.FRoR += aA * sn * iR2;
  F.FphioR2 += -aA / C->p[4] * C->p[5] * sn * iR2;
}

// CosmphiDiskPotential.c:4-37 (planar only): Phi = amp mphio/m R^p
// cos(m phi - mphib) outside the break radius rb, amp mphio/m rbp (2 r1p -
// rbp R^-p) cos(m phi - mphib) inside (F_R continuous at
// rb; the potential too when r1 = 1); p: amp, mphio, p, mphib, m, rb, rbp,
// rb2p; c[0] = r1p (galpy's ninth argument), c[2] = mphio / m.
template <>
__device__ __forceinline__ void gh_comp_force<GH_COSMPHIDISK>(const GhPotTable* __restrict__ T,
                                                              int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], pp = C->p[2];
  const GhSinCos sc = gh_sincos(C->p[4] * gh_phi_acos(g.x, g.y, g.R2) - C->p[3]);
  const double iR2 = gh_rcp(g.R2);
  const double Rp = exp(0.5 * pp * log(g.R2));  // R^p
  if (sqrt(g.R2) <= C->p[5]) {  // R <= rb as galpy compares it
    F.FRoR += -amp * pp * C->c[2] * C->p[7] / Rp * iR2 * sc.c;
    F.FphioR2 += amp * C->p[1] * C->p[6] * (2.0 * C->c[0] - C->p[6] / Rp) * iR2 * sc.s;
  } else {
    F.FRoR += -amp * pp * C->c[2] * Rp * iR2 * sc.c;
    F.FphioR2 += amp * C->p[1] * Rp * iR2 * sc.s;
  }
}

// HenonHeilesPotential.c:4-14 (planar only): Phi = amp/2 (R^2 + 2/3 R^3
// sin 3 phi) = amp/2 (x^2 + y^2) + amp/3 (3 x^2 y - y^3); p: amp.
template <>
__device__ __forceinline__ void gh_comp_force<GH_HENONHEILES>(const GhPotTable* __restrict__ T,
                                                              int i, const GhGeom& g, GhAcc& F) {
  const double amp = T->comp[i].p[0];
  const double iR2 = gh_rcp(g.R2);
  const double s3 = (3.0 * g.x * g.x - g.y * g.y) * g.y;  // R^3 sin 3 phi
  const double c3 = (g.x * g.x - 3.0 * g.y * g.y) * g.x;  // R^3 cos 3 phi
  F.FRoR += -amp * (1.0 + s3 * iR2);
  F.FphioR2 += -amp * c3 * iR2;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_KKSTAECKEL>(const GhPotTable* __restrict__ T,
                                                             int i, const GhGeom& g, GhAcc& F) {
  // KuzminKutuzovStaeckelPotential.c:24-85: Phi = -amp / (sqrt(l) + sqrt(n))
  // in the spheroidal coordinates l >= -alpha >= n >= -gamma of (R, z)
  // (focal distance Delta, gamma = Delta^2 / (1 - ac^2), alpha = gamma -
  // Delta^2); forces by the chain rule through l and n.  c[0] = gamma,
  // c[1] = alpha, c[2] = Delta^2.
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], gam = C->c[0], alp = C->c[1], D2 = C->c[2];
  const double r2 = g.R2 + g.z2;
  const double dm = r2 - D2;
  const double s = sqrt(dm * dm + 4.0 * D2 * g.R2);
  const double is = 1.0 / s;
  const double term = r2 - alp - gam;
  const double sl = sqrt(0.5 * (term + s)), sn = sqrt(0.5 * (term - s));
  const double isum = 1.0 / (sl + sn);
  const double dVdl = 0.5 / sl * isum * isum, dVdn = 0.5 / sn * isum * isum;
  const double uR = (r2 + D2) * is, uz = dm * is;
  F.FRoR += -amp * ((1.0 + uR) * dVdl + (1.0 - uR) * dVdn);
  F.Fz += -amp * g.z * ((1.0 + uz) * dVdl + (1.0 - uz) * dVdn);
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_FLATPOWER>(const GhPotTable* __restrict__ T,
                                                            int i, const GhGeom& g, GhAcc& F) {
  // FlattenedPowerPotential.c:19-64: F = -amp m^(-alpha/2 - 1) (x, y, z/q^2),
  // m^2 = R^2 + z^2/q^2 + core^2 (the alpha = 0 branch is the same
  // expression, -amp / m^2); c[0] = 1/q^2
  const GhComponent* __restrict__ C = &T->comp[i];
  const double amp = C->p[0], alpha = C->p[1], iq2 = C->c[0];
  const double m2 = C->p[3] + g.R2 + g.z2 * iq2;
  const double f = (alpha == 0.0) ? -amp * gh_rcp(m2) : -amp * exp((-0.5 * alpha - 1.0) * log(m2));
  F.FRoR += f;
  F.Fz += f * g.z * iq2;
}

// ---------------------------------------------------------------------------
// EllipsoidalPotential family: the Gauss-Legendre force sum of
// EllipsoidalPotentialxyzforces_xyz (EllipsoidalPotential.c:40-77) in the
// frame of the ellipsoid,
//   F = amp sum_k w_k rho(m_k) (x u1_k, y u2_k, z u3_k),
//   m_k^2 = x^2 u1_k + y^2 u2_k + z^2 u3_k,
// u1 = 1/(1+tau), u2 = 1/(b^2+tau), u3 = 1/(c^2+tau), tau = 1/s_k^2 - 1,
// w_k galpy's packed weight (-4 pi b c glw / D folded in), rotated back by
// the transpose of the frame's rotation (galpy_potentials.c:391-410).  The
// host stores per node (u1, u2, u3, w), so the loop multiplies where galpy
// divides, and the node data of a wavefront are one scalar load stream.
// Densities rho(m) (TwoPowerTriaxialPotential.c:13-117,
// PerfectEllipsoidPotential.c:9-20, TriaxialGaussianPotential.c:8-16,
// PowerTriaxialPotential.c:9-20), written on m^2:
//   21 TriaxialHernquist  a^4 / (m (m+a)^3)     p = amp, a, a^4
//   22 TriaxialNFW        a^3 / (m (m+a)^2)     p = amp, a, a^3
//   23 TriaxialJaffe      a^4 / (m (m+a))^2     p = amp, a, a^2
//   43 TwoPowerTriaxial   (a/m)^alpha (1+m/a)^(alpha-beta)
//                          p = amp, a, alpha, beta, beta-alpha, ...
//   30 PerfectEllipsoid   1 / (a^2 + m^2)^2     p = amp, a^2
//   37 TriaxialGaussian   exp(m^2 / (-2 sigma^2))  p = amp, -2 sigma^2
//   38 PowerTriaxial      m^-alpha              p = amp, alpha
// c[0] = b^2, c[1] = c^2, c[2] = aligned, c[3] = glorder, c[4] = 1/a for 43;
// ext = rotation (9), then (u1, u2, u3, w) per node.
// ---------------------------------------------------------------------------
template <int TYPE>
__device__ __forceinline__ double gh_ellip_mdens(const GhComponent* __restrict__ C, double m2) {
  if constexpr (TYPE == GH_TRIAXHERNQUIST || TYPE == GH_TRIAXNFW || TYPE == GH_TRIAXJAFFE) {
    const double im = gh_rsqrt(m2);
    const double mpa = m2 * im + C->p[1];
    if constexpr (TYPE == GH_TRIAXHERNQUIST) return C->p[2] * im * gh_rcp(mpa * mpa * mpa);
    if constexpr (TYPE == GH_TRIAXNFW) return C->p[2] * im * gh_rcp(mpa * mpa);
    const double q = im * gh_rcp(mpa);
    return C->p[2] * C->p[2] * (q * q);
  } else if constexpr (TYPE == GH_TWOPOWERTRIAX) {
    const double moa = sqrt(m2) * C->c[4];
    return exp(-C->p[2] * log(moa) - C->p[4] * log1p(moa));
  } else if constexpr (TYPE == GH_PERFECTELLIPSOID) {
    const double q = gh_rcp(C->p[1] + m2);
    return q * q;
  } else if constexpr (TYPE == GH_TRIAXGAUSSIAN) {
    return exp(m2 / C->p[1]);
  } else if constexpr (TYPE == GH_POWERTRIAX) {
    return exp(-0.5 * C->p[1] * log(m2));
  } else {
    return 0.0;
  }
}

// run-time member (generic kernels): one loop, a wavefront-uniform branch
// per node
__device__ __forceinline__ double gh_ellip_mdens_dyn(const GhComponent* __restrict__ C,
                                                     double m2) {
  switch (C->type) {
    case GH_TRIAXHERNQUIST: return gh_ellip_mdens<GH_TRIAXHERNQUIST>(C, m2);
    case GH_TRIAXNFW: return gh_ellip_mdens<GH_TRIAXNFW>(C, m2);
    case GH_TRIAXJAFFE: return gh_ellip_mdens<GH_TRIAXJAFFE>(C, m2);
    case GH_TWOPOWERTRIAX: return gh_ellip_mdens<GH_TWOPOWERTRIAX>(C, m2);
    case GH_PERFECTELLIPSOID: return gh_ellip_mdens<GH_PERFECTELLIPSOID>(C, m2);
    case GH_TRIAXGAUSSIAN: return gh_ellip_mdens<GH_TRIAXGAUSSIAN>(C, m2);
    default: return gh_ellip_mdens<GH_POWERTRIAX>(C, m2);
  }
}

// Cartesian force (times amp) of component C at (x, y, z); TYPE 0 picks the
// member at run time.
template <int TYPE>
__device__ __forceinline__ void gh_ellip_xyz(const GhPotTable* __restrict__ T,
                                             const GhComponent* __restrict__ C, double x,
                                             double y, double z, double& Ax, double& Ay,
                                             double& Az) {
  const double* __restrict__ e = &T->ext[C->ext];
  const bool aligned = C->c[2] != 0.0;
  double xe = x, ye = y, ze = z;
  if (!aligned) {
    xe = e[0] * x + e[1] * y + e[2] * z;
    ye = e[3] * x + e[4] * y + e[5] * z;
    ze = e[6] * x + e[7] * y + e[8] * z;
  }
  const double x2 = xe * xe, y2 = ye * ye, z2 = ze * ze;
  const double* __restrict__ nd = e + 9;
  const int n = (int)C->c[3];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int k = 0; k < n; k++, nd += 4) {
    const double u1 = nd[0], u2 = nd[1], u3 = nd[2];
    const double m2 = fma(x2, u1, fma(y2, u2, z2 * u3));
    const double rho = (TYPE == 0) ? gh_ellip_mdens_dyn(C, m2) : gh_ellip_mdens<TYPE>(C, m2);
    const double td = nd[3] * rho;
    fx = fma(td * u1, xe, fx);
    fy = fma(td * u2, ye, fy);
    fz = fma(td * u3, ze, fz);
  }
  if (!aligned) {
    const double gx = e[0] * fx + e[3] * fy + e[6] * fz;
    const double gy = e[1] * fx + e[4] * fy + e[7] * fz;
    const double gz = e[2] * fx + e[5] * fy + e[8] * fz;
    fx = gx;
    fy = gy;
    fz = gz;
  }
  const double amp = C->p[0];
  Ax = amp * fx;
  Ay = amp * fy;
  Az = amp * fz;
}

// In the accumulated form: F_R / R = (x Ax + y Ay) / R^2, F_phi / R^2 =
// (x Ay - y Ax) / R^2 (EllipsoidalPotentialRforce / phitorque,
// EllipsoidalPotential.c:85-105, with cos phi = x / R, sin phi = y / R).
template <int TYPE>
__device__ __forceinline__ void gh_ellip_force(const GhPotTable* __restrict__ T, int i,
                                               const GhGeom& g, GhAcc& F) {
  double Ax, Ay, Az;
  gh_ellip_xyz<TYPE>(T, &T->comp[i], g.x, g.y, g.z, Ax, Ay, Az);
  const double iR2 = gh_rcp(g.R2);
  F.FRoR += (g.x * Ax + g.y * Ay) * iR2;
  F.FphioR2 += (g.x * Ay - g.y * Ax) * iR2;
  F.Fz += Az;
}
#define GH_ELLIP_FORCE(CODE)                                                                  \
  template <>                                                                                 \
  __device__ __forceinline__ void gh_comp_force<CODE>(const GhPotTable* __restrict__ T, int i, \
                                                      const GhGeom& g, GhAcc& F) {            \
    gh_ellip_force<CODE>(T, i, g, F);                                                         \
  }
GH_ELLIP_FORCE(GH_TRIAXHERNQUIST)
GH_ELLIP_FORCE(GH_TRIAXNFW)
GH_ELLIP_FORCE(GH_TRIAXJAFFE)
GH_ELLIP_FORCE(GH_TWOPOWERTRIAX)
GH_ELLIP_FORCE(GH_PERFECTELLIPSOID)
GH_ELLIP_FORCE(GH_TRIAXGAUSSIAN)
GH_ELLIP_FORCE(GH_POWERTRIAX)
#undef GH_ELLIP_FORCE

// ---------------------------------------------------------------------------
// Spherical components whose enclosed mass needs a special function; each
// adds g(r) x, g = -Phi'(r)/r, like the closed-form spherical components.
//
// EinastoPotential.c:27-47: Phi' = amp 4 pi h n s^-2 gamma(3n, s^(1/n)),
// s = r/h (galpy writes -Gamma(3n) (Q(3n, x) - 1), which cancels for
// x << 3n; gamma(3n, x) is the same function without the cancellation).
// c[0] = 3n, c[1] = Gamma(3n), c[2] = 2^-58 Gamma(3n), c[3] = 2n,
// c[4] = Gamma(2n), c[5] = 2^-58 Gamma(2n), c[6] = 4 pi h^3 n, c[7] = 1/h,
// c[8] = 1/n.
__device__ __forceinline__ double gh_einasto_x(const GhComponent* __restrict__ C, double r) {
  return exp(C->c[8] * log(r * C->c[7]));  // (r/h)^(1/n)
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_EINASTO>(const GhPotTable* __restrict__ T, int i,
                                                          const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double gam = gh_lower_incgamma_gen(C->c[0], gh_einasto_x(C, g.r), C->c[1], C->c[2]);
  const double gr = -C->p[0] * C->c[6] * gam * (g.rinv * g.rinv * g.rinv);
  F.FRoR += gr;
  F.Fz += gr * g.z;
}

// ExpTruncNFWPotential.c:11-28,53-58: Phi' = amp F(r) / r^2 with
// F = (1 + alpha) (e^alpha E1(alpha) - e^alpha E1(beta)) - 1 + a e^(-r/rc)/(a + r),
// alpha = a/rc, beta = (a + r)/rc, and e^alpha E1(beta) = e^(-r/rc) e^beta
// E1(beta); below 1e-3 min(a, rc) galpy's quartic series.
// c[0] = alpha, c[1] = e^alpha E1(alpha), c[2] = 1/rc, c[3] = series
// threshold, c[4..7] = series coefficients, c[8] = 1/a^2.
__device__ inline double gh_exptruncnfw_F(const GhComponent* __restrict__ C, double r,
                                          double& ex, double& e1b) {
  const double a = C->p[1];
  ex = exp(-r * C->c[2]);
  if (r < C->c[3]) {
    e1b = 0.0;
    return (r * r * C->c[8]) * (C->c[4] + r * (C->c[5] + r * (C->c[6] + r * C->c[7])));
  }
  e1b = gh_e1_scaled((a + r) * C->c[2]);
  return (1.0 + C->c[0]) * (C->c[1] - ex * e1b) - 1.0 + a * ex / (a + r);
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_EXPTRUNCNFW>(const GhPotTable* __restrict__ T,
                                                              int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  double ex, e1b;
  const double gr = -C->p[0] * gh_exptruncnfw_F(C, g.r, ex, e1b) * (g.rinv * g.rinv * g.rinv);
  F.FRoR += gr;
  F.Fz += gr * g.z;
}

// TwoPowerSphericalPotential.c:37-49: Phi' = amp r^(1-alpha) a^(alpha-3) /
// (3-alpha) 2F1(3-alpha, beta-alpha; 4-alpha; -r/a) = amp I(r/a) / r^2 with
// I(x) = int_0^x u^(A-1) (1+u)^-B du, A = 3 - alpha, B = beta - alpha (the
// enclosed mass over amp).  galpy evaluates 2F1 with the xsf library; here
//   x <= 3: I = x^A/A (1+x)^-B sum_n s_n w^n, w = x/(1+x), s_n =
//           (B)_n/(A+1)_n (Pfaff's transformation; positive terms);
//   x > 3 : I = K + x^(3-beta) sum_k d_k x^-k [+ c* ln x], the binomial
//           series of (1+1/u)^-B integrated from 3 (ln for the u^-1 term when
//           3 - beta is a nonnegative integer), K collecting the constants.
// The host computes s_n, d_k, K and c* in long double (ext: s_n then d_k).
// c[0] = A, c[1] = B, c[2] = 3 - beta, c[3] = #s, c[4] = #d, c[5] = K,
// c[6] = c*, c[7] = 1/a, c[8] = 1/A.
__device__ inline double gh_twopower_mass(const GhPotTable* __restrict__ T,
                                          const GhComponent* __restrict__ C, double x) {
  const double* __restrict__ e = &T->ext[C->ext];
  const int ns = (int)C->c[3], nd = (int)C->c[4];
  if (x <= 3.0) {
    const double w = x / (1.0 + x);
    double sum = e[ns - 1];
    for (int n = ns - 2; n >= 0; n--) sum = fma(sum, w, e[n]);
    return exp(C->c[0] * log(x) - C->c[1] * log1p(x)) * C->c[8] * sum;
  }
  const double* __restrict__ d = e + ns;
  const double u = 1.0 / x, lx = log(x);
  double p = d[nd - 1];
  for (int k = nd - 2; k >= 0; k--) p = fma(p, u, d[k]);
  return C->c[5] + exp(C->c[2] * lx) * p + C->c[6] * lx;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_TWOPOWERSPH>(const GhPotTable* __restrict__ T,
                                                              int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double gr = -C->p[0] * gh_twopower_mass(T, C, g.r * C->c[7]) * (g.rinv * g.rinv * g.rinv);
  F.FRoR += gr;
  F.Fz += gr * g.z;
}

// DoubleExponentialDiskPotential.c: the disk's Hankel integrals by Ogata's
// double-exponential quadrature over the packed nodes x_k and weights w_k
// (ext: J0 pairs (x_k, w_k), then J1 pairs), each sum stopped once a term is
// at most 1e-15 of the running sum (galpy divides the term by the sum and
// compares; here the comparison is multiplied out) or the nodes run out:
//   F_R = A/R sum_J1 w u (alpha^2 + u^2)^-3/2 (beta e^-u|z| - u e^-beta|z|) / (beta^2 - u^2)
//   F_z = +-A beta/R sum_J0 w u (alpha^2 + u^2)^-3/2 (e^-u|z| - e^-beta|z|) / (beta^2 - u^2)
//   Phi = A/R sum_J0 w (alpha^2 + u^2)^-3/2 (beta e^-u|z| - u e^-beta|z|) / (beta^2 - u^2)
// u = x_k / R, A = -4 pi alpha amp (galpy's second argument); planar
// tables use galpy's in-plane form of F_R, w u (alpha^2 + u^2)^-3/2 /
// (beta + u) (PlanarRforce, :63-86).
// p: amp, A, alpha, beta; c[0] = de_n, c[1] = alpha^2, c[2] = beta^2,
// c[3] = 1 for the planar packing.
template <int KIND>  // 0: F_R (J1), 1: F_z (J0), 2: Phi (J0), 3: planar F_R (J1)
__device__ __forceinline__ double gh_dedisk_sum(const GhPotTable* __restrict__ T,
                                                const GhComponent* __restrict__ C, double iR,
                                                double fz) {
  const int n = (int)C->c[0];
  const double* __restrict__ xw = &T->ext[C->ext + ((KIND == 0 || KIND == 3) ? 2 * n : 0)];
  const double a2 = C->c[1], b2 = C->c[2], beta = C->p[3];
  const double ebz = exp(-beta * fz);
  double out = 0.0;
  for (int k = 0; k < n; k++) {
    const double u = xw[2 * k] * iR;
    const double rs = gh_rsqrt(a2 + u * u);
    const double w = xw[2 * k + 1] * rs * rs * rs;
    double term;
    if (KIND == 3) {
      term = w * u / (beta + u);
    } else {
      const double eu = exp(-u * fz);
      const double num = (KIND == 1) ? u * (eu - ebz) : beta * eu - u * ebz;
      term = ((KIND == 0) ? w * u : w) * num / (b2 - u * u);
    }
    out += term;
    if (!(fabs(term) > 1e-15 * fabs(out))) break;
  }
  return out;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_DEDISK>(const GhPotTable* __restrict__ T, int i,
                                                         const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double iR = gh_rsqrt(g.R2), A = C->p[1];
  const double fz = fabs(g.z);
  if (C->c[3] != 0.0) {  // planar packing
    F.FRoR += A * gh_dedisk_sum<3>(T, C, iR, 0.0) * iR * iR;
    return;
  }
  F.FRoR += A * gh_dedisk_sum<0>(T, C, iR, fz) * iR * iR;
  const double sz = A * C->p[3] * gh_dedisk_sum<1>(T, C, iR, fz) * iR;
  F.Fz += (g.z > 0.0) ? sz : -sz;
}

// SpiralArmsPotential.c (Cox & Gomez 2002 arms), all three forces in one
// pass over the terms n = 1..nCs (the reference recomputes K, B, D in each
// of Rforce :61-126, zforce :128-170 and phitorque :172-210).  Per term,
// with kappa = H K_n = H N n / (R sin alpha):
//   B_n = kappa (1 + 0.4 kappa), D_n = (0.3 kappa^2 + kappa + 1) / (0.3 kappa + 1),
//   dK/dR = -K/R, dB/dR = -kappa (1 + 0.8 kappa) / R,
//   dD/dR = kappa [0.3 (kappa + 0.3 kappa^2 + 1) / (0.3 kappa + 1)^2
//                  - (1 + 0.6 kappa) / (0.3 kappa + 1)] / R,
// u = z K/B; tanh u and log cosh u from e = exp(-2|u|) (one exp and one
// log1p), sech^B = exp(-B log cosh u); cos n g, sin n g by the angle-addition
// recurrence from sincos g, g = N (phi - omega t - phi_ref - ln(R/r_ref) /
// tan alpha) with phi from acos as galpy's evalRectForce forms it.
// p: amp, N, sin alpha, tan alpha, r_ref, phi_ref, Rs, H; c[0] = omega,
// c[1] = nCs, c[2] = 1/Rs, c[3] = 1/tan alpha, c[4] = ln r_ref,
// c[5] = H N / sin alpha; ext = C_1..C_nCs.
struct GhSpiralSums {
  double sr, sz, sp, pot;
};
template <bool POT>
__device__ __forceinline__ GhSpiralSums gh_spiral_sums(const GhPotTable* __restrict__ T,
                                                       const GhComponent* __restrict__ C, double R,
                                                       double z, double phi, double t) {
  const double* __restrict__ Cs = &T->ext[C->ext];
  const int nCs = (int)C->c[1];
  const double N = C->p[1], iR = 1.0 / R;
  const double g = N * (phi - C->c[0] * t - C->p[5] - (log(R) - C->c[4]) * C->c[3]);
  const double dg = -N * iR * C->c[3];
  const GhSinCos sc1 = gh_sincos(g);
  double cng = sc1.c, sng = sc1.s;
  const double K1 = N * iR / C->p[2], kap1 = C->c[5] * iR, az = fabs(z);
  GhSpiralSums S = {0.0, 0.0, 0.0, 0.0};
  for (int k = 1; k <= nCs; k++) {
    const double n = k;
    const double Kn = n * K1, kap = n * kap1;
    const double q = 0.3 * kap + 1.0, iq = 1.0 / q;
    const double Bn = kap * (1.0 + 0.4 * kap);
    const double Dn = (0.3 * kap * kap + kap + 1.0) * iq;
    const double u = az / (C->p[7] * (1.0 + 0.4 * kap));  // |z| K / B
    const double e = exp(-2.0 * u);
    const double lch = u + log1p(e) - M_LN2;               // log cosh u
    const double th = copysign((1.0 - e) / (1.0 + e), z);  // tanh(z K / B)
    const double sechB = exp(-Bn * lch);
    const double Cn = Cs[k - 1], iK = 1.0 / Kn, iD = 1.0 / Dn;
    if (POT) {
      S.pot += Cn * iK * iD * cng * sechB;
    } else {
      const double dKoK = -iR, dBoB = -(1.0 + 0.8 * kap) * iR / (1.0 + 0.4 * kap);
      const double dB = Bn * dBoB;
      const double dD = kap * iR * (0.3 * (kap + 0.3 * kap * kap + 1.0) * iq * iq - (1.0 + 0.6 * kap) * iq);
      S.sr += Cn * sechB * iD *
              (n * dg * iK * sng +
               cng * (z * th * (dKoK - dBoB) + dB * iK * lch + dKoK * iK + dD * iD * iK) +
               cng * iK * C->c[2]);
      S.sz += Cn * iD * cng * th * sechB;
      S.sp += N * n * Cn * iD * iK * sechB * sng;
    }
    const double c2 = cng * sc1.c - sng * sc1.s;
    sng = sng * sc1.c + cng * sc1.s;
    cng = c2;
  }
  return S;
}
__device__ __forceinline__ double gh_spiral_pre(const GhComponent* __restrict__ C, double R) {
  return -C->p[0] * C->p[7] * exp(-(R - C->p[4]) * C->c[2]);
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_SPIRALARMS>(const GhPotTable* __restrict__ T,
                                                             int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double R = sqrt(g.R2);
  const GhSpiralSums S = gh_spiral_sums<false>(T, C, R, g.z, gh_phi_acos(g.x, g.y, g.R2), g.t);
  const double pre = gh_spiral_pre(C, R), iR2 = gh_rcp(g.R2);
  F.FRoR += pre * S.sr * (R * iR2);
  F.Fz += pre * S.sz;
  F.FphioR2 += pre * S.sp * iR2;
}

// interpSphericalPotential.c:5-37 through SphericalPotential.c:13-38: the
// radial force is the natural spline of the tabulated force on [rmin, rmax),
// 0 inside rmin, -M/r^2 outside rmax.  p: amp, rmin, rmax, M(<rmax), Phi0,
// Phimax; ext = the spline.
__device__ __forceinline__ double gh_interpsph_rforce(const GhPotTable* __restrict__ T,
                                                      const GhComponent* __restrict__ C,
                                                      double r) {
  if (r >= C->p[2]) return -C->p[3] / (r * r);
  return (r < C->p[1]) ? 0.0 : gh_spline_eval(T, C->ext, r);
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_INTERPSPH>(const GhPotTable* __restrict__ T,
                                                            int i, const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double f = C->p[0] * gh_interpsph_rforce(T, C, g.r) * g.rinv;
  F.FRoR += f;
  F.Fz += f * g.z;
}

// SCFPotential.c (Hernquist & Ostriker 1992 basis): the sums of
// sum_spher_forces / sum_expansion (:366-438) without the per-call arrays:
// the outer loop runs over m (cos m phi, sin m phi and P_m^m by recurrence
// from the previous m), the next over l >= m (P_l^m and dP_l^m/dx by the
// three-term recurrence, (a r / (a + r)^2)^l by products), the inner over n,
// where C_n^(2l+3/2)(xi) and C_(n-1)^(2l+5/2)(xi) (dC/dxi / (2 alpha)) are
// run by their recurrences.  Coefficients (ext): Acos[N L M] and Asin, or,
// for a time-dependent expansion, tgrid[Nt] and per coefficient the cubic in
// t - tgrid[i] (Horner).  p: a, isNonAxi, N, L, M, Nt.
// Returns (dPhi/dr, dPhi/dtheta, dPhi/dphi) with the reference's signs in
// F[0..2] (minus, minus, plus) and the potential.
template <bool POT>
__device__ inline void gh_scf_sums(const GhPotTable* __restrict__ T,
                                   const GhComponent* __restrict__ C, double R, double z,
                                   double cphi, double sphi, double t, double* F, double& V) {
  const double a = C->p[0];
  const bool nonaxi = C->p[1] != 0.0;
  const int N = (int)C->p[2], L = (int)C->p[3], M = (int)C->p[4], Nt = (int)C->p[5];
  const int NLM = N * L * M;
  const double* __restrict__ e = &T->ext[C->ext];
  int it = 0;
  double dt = 0.0;
  const double* __restrict__ ca = e;
  const double* __restrict__ cs_ = e + NLM;
  if (Nt > 0) {  // scf_find_time_interval, clamped to the boundary intervals
    const double* __restrict__ tg = e;
    if (t <= tg[0]) it = 0;
    else if (t >= tg[Nt - 1]) it = Nt - 2;
    else {
      int lo = 0, hi = Nt - 2;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tg[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      it = lo;
    }
    dt = t - tg[it];
    ca = e + Nt;
    cs_ = ca + (size_t)NLM * 4 * (Nt - 1);
  }
  const double r2 = R * R + z * z, r = sqrt(r2);
  const double x = z / r, sth = R / r;
  const double ar = a + r, xi = (r - a) / ar, q = r * a / (ar * ar);
  const double iar = 1.0 / ar, ird = 1.0 / (r * ar * ar * ar), ix2m1 = 1.0 / (x * x - 1.0);
  double F0 = 0.0, F1 = 0.0, F2 = 0.0, pot = 0.0;
  const int Meff = nonaxi ? (M < L ? M : L) : 1;
  double cm = 1.0, sm = 0.0, pmm = 1.0, qm = 1.0;
  for (int m = 0; m < Meff; m++) {
    if (m > 0) {
      const double c2 = cm * cphi - sm * sphi;
      sm = sm * cphi + cm * sphi;
      cm = c2;
      pmm *= -(2.0 * m - 1.0) * sth;
      qm *= q;
    }
    double plm1 = 0.0, plm2 = 0.0, ql = qm;
    for (int l = m; l < L; l++) {
      double pl;
      if (l == m) pl = pmm;
      else if (l == m + 1) pl = (2.0 * m + 1.0) * x * pmm;
      else pl = ((2.0 * l - 1.0) * x * plm1 - (l + m - 1.0) * plm2) / (l - m);
      const double dpl = (l == 0) ? 0.0 : (l * x * pl - (l + m) * ((l == m) ? 0.0 : plm1)) * ix2m1;
      if (l > m) plm2 = plm1;
      plm1 = pl;
      const double rt = -ql * iar, rd = ql * ird;
      const double lin = (2 * l + 1) * r * ar - l * ar * ar;
      const double alpha = 2.0 * l + 1.5;
      double c0 = 1.0, c1 = 0.0, h0 = 1.0, h1 = 0.0;  // C_n^alpha, C_(n-1)^(alpha+1)
      for (int n = 0; n < N; n++) {
        double Cn, Hn;  // C_n^alpha(xi), C_(n-1)^(alpha+1)(xi)
        if (n == 0) {
          Cn = 1.0;
          Hn = 0.0;
        } else if (n == 1) {
          Cn = 2.0 * alpha * xi;
          Hn = 1.0;
        } else {
          Cn = (2.0 * (n - 1.0 + alpha) * xi * c0 - (n - 2.0 + 2.0 * alpha) * c1) / n;
          const double k = n - 1;  // C_k^(alpha+1) from C_(k-1), C_(k-2)
          Hn = (k == 1) ? 2.0 * (alpha + 1.0) * xi
                        : (2.0 * (k - 1.0 + alpha + 1.0) * xi * h0 - (k - 2.0 + 2.0 * (alpha + 1.0)) * h1) / k;
        }
        if (n >= 1) {
          c1 = c0;
          c0 = Cn;
        } else {
          c0 = Cn;
        }
        if (n >= 2) {
          h1 = h0;
          h0 = Hn;
        } else if (n == 1) {
          h0 = Hn;
        }
        const int ci = n * L * M + l * M + m;
        double Ac, As = 0.0;
        if (Nt == 0) {
          Ac = ca[ci];
          if (nonaxi) As = cs_[ci];
        } else {
          const double* cp = ca + ((size_t)it * NLM + ci) * 4;
          Ac = fma(fma(fma(cp[0], dt, cp[1]), dt, cp[2]), dt, cp[3]);
          if (nonaxi) {
            const double* sp = cs_ + ((size_t)it * NLM + ci) * 4;
            As = fma(fma(fma(sp[0], dt, sp[1]), dt, sp[2]), dt, sp[3]);
          }
        }
        const double csum = Ac * cm + As * sm;
        const double phT = rt * Cn;
        if (POT) {
          pot += csum * pl * phT;
        } else {
          const double sdiff = Ac * sm - As * cm;
          const double dC = 2.0 * alpha * Hn;
          const double dphT = rd * (lin * Cn - 2.0 * a * r * dC);
          F0 -= csum * pl * dphT;
          F1 -= csum * dpl * phT;
          F2 += m * sdiff * pl * phT;
        }
      }
      ql *= q;
    }
  }
  const double s4 = 3.5449077018110320546;  // sqrt(4 pi)
  F[0] = F0 * s4;
  F[1] = F1 * s4 * (-sth);
  F[2] = F2 * s4;
  V = pot * s4;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_SCF>(const GhPotTable* __restrict__ T, int i,
                                                      const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double R = sqrt(g.R2);
  const double cp = (R > 0.0) ? g.x / R : 1.0, sp = (R > 0.0) ? g.y / R : 0.0;
  double S[3], V;
  gh_scf_sums<false>(T, C, R, g.z, cp, sp, g.t, S, V);
  const double ir2 = g.rinv * g.rinv, iR = 1.0 / R;
  F.FRoR += (S[0] * g.rinv + S[1] * g.z * ir2 * iR);  // F_R / R
  F.Fz += S[0] * g.z * g.rinv - S[1] * R * ir2;
  F.FphioR2 += S[2] * iR * iR;
}

// DiskSCFPotential.c:33-171: one [Sigma_i, h_i] approximation pair,
// Phi = Sigma(r) H(z); F_R = -Sigma'(r) H R / r, F_z = -Sigma'(r) H z / r -
// Sigma(r) H'(z); the planar packing's force -Sigma'(R) H(0).  Sigma: amp
// exp(-r/h) (type 0) or amp exp(-Rhole/r - r/h) (type 1); H: exponential
// (0) or sech^2 (1).  p: sigma type, amp, h, Rhole, hz type, hz h; c[0] = 1/h,
// c[1] = 1/hz, c[2] = 1 for the planar packing.
__device__ __forceinline__ void gh_dscf_sigma(const GhComponent* __restrict__ C, double r,
                                              double& S, double& dS) {
  if (C->p[0] == 0.0) {
    S = C->p[1] * exp(-r * C->c[0]);
    dS = -S * C->c[0];
  } else {
    const double ir = 1.0 / r;
    S = C->p[1] * exp(-C->p[3] * ir - r * C->c[0]);
    dS = S * (C->p[3] * ir * ir - C->c[0]);
  }
}
__device__ __forceinline__ void gh_dscf_hz(const GhComponent* __restrict__ C, double z, double& H,
                                           double& dH) {
  const double fz = fabs(z), h = C->p[5], ih = C->c[1];
  const double e = exp(-fz * ih);
  if (C->p[4] == 0.0) {
    H = 0.5 * (e - 1.0 + fz * ih) * h;
    dH = copysign(0.5 * (1.0 - e), z);
  } else {
    H = h * (log1p(e) + 0.5 * fz * ih - M_LN2);
    dH = copysign(0.5 * (1.0 - e) / (1.0 + e), z);  // tanh(z / (2 h)) / 2
  }
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_DISKSCF>(const GhPotTable* __restrict__ T, int i,
                                                          const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  double S, dS, H, dH;
  if (C->c[2] != 0.0) {  // planar packing
    const double R = sqrt(g.R2);
    gh_dscf_sigma(C, R, S, dS);
    gh_dscf_hz(C, 0.0, H, dH);
    F.FRoR += -dS * H / R;
    return;
  }
  gh_dscf_sigma(C, g.r, S, dS);
  gh_dscf_hz(C, g.z, H, dH);
  const double f = -dS * H * g.rinv;
  F.FRoR += f;
  F.Fz += f * g.z - S * dH;
}

// interpRZPotential.c:6-70: cubic B-spline interpolation on an (R or
// log R, |z|) grid (util/interp_2d/interp_2d.c:92-114 and Thevenaz's
// cubic_bspline_2d_interpol with mirror boundaries); R force, z force (odd
// in z) and potential from their own coefficient grids at the same indices
// and weights.  p: amp, logR; c[0] = nR, c[1] = nz; ext: R grid, z grid,
// then the potential's, R force's and z force's coefficients [nR][nz].
__device__ __forceinline__ int gh_mirror(long k, long n) {
  const long n2 = 2L * n - 2L;
  if (n == 1L) return 0;
  k = (k < 0L) ? (-k - n2 * ((-k) / n2)) : (k - n2 * (k / n2));
  return (int)((n <= k) ? n2 - k : k);
}
struct GhBspl {
  int xi[4], yi[4];
  double xw[4], yw[4];
};
__device__ __forceinline__ void gh_bspl_weights(double u, int n, int* idx, double* w) {
  const long i = (long)floor(u) - 1L;
  const double d = u - (double)(i + 1L);
  w[3] = (1.0 / 6.0) * d * d * d;
  w[0] = (1.0 / 6.0) + 0.5 * d * (d - 1.0) - w[3];
  w[2] = d + w[0] - 2.0 * w[3];
  w[1] = 1.0 - w[0] - w[2] - w[3];
  for (int k = 0; k < 4; k++) idx[k] = gh_mirror(i + k, n);
}
__device__ __forceinline__ GhBspl gh_interprz_setup(const GhPotTable* __restrict__ T,
                                                    const GhComponent* __restrict__ C, double R,
                                                    double z) {
  const int nR = (int)C->c[0], nz = (int)C->c[1];
  const double* __restrict__ Rg = &T->ext[C->ext];
  const double* __restrict__ zg = Rg + nR;
  double x = (C->p[1] == 1.0) ? ((R > 0.0) ? log(R) : -20.72326583694641) : R;
  double y = fabs(z);
  x = x > Rg[nR - 1] ? Rg[nR - 1] : (x < Rg[0] ? Rg[0] : x);
  y = y > zg[nz - 1] ? zg[nz - 1] : (y < zg[0] ? zg[0] : y);
  const int ix = gh_spline_find(Rg, nR, x), iy = gh_spline_find(zg, nz, y);
  GhBspl B;
  gh_bspl_weights(ix + (x - Rg[ix]) / (Rg[ix + 1] - Rg[ix]), nR, B.xi, B.xw);
  gh_bspl_weights(iy + (y - zg[iy]) / (zg[iy + 1] - zg[iy]), nz, B.yi, B.yw);
  return B;
}
__device__ __forceinline__ double gh_interprz_sum(const GhPotTable* __restrict__ T,
                                                  const GhComponent* __restrict__ C,
                                                  const GhBspl& B, int grid) {
  const int nR = (int)C->c[0], nz = (int)C->c[1];
  const double* __restrict__ co = &T->ext[C->ext + nR + nz + grid * nR * nz];
  double out = 0.0;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) out += co[B.xi[a] * nz + B.yi[b]] * B.xw[a] * B.yw[b];
  return out;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_INTERPRZ>(const GhPotTable* __restrict__ T, int i,
                                                           const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double R = sqrt(g.R2);
  const GhBspl B = gh_interprz_setup(T, C, R, g.z);
  const double amp = C->p[0];
  F.FRoR += amp * gh_interprz_sum(T, C, B, 1) / R;
  const double fz = amp * gh_interprz_sum(T, C, B, 2);
  F.Fz += (g.z < 0.0) ? -fz : fz;
}

// MultipoleExpansionPotential.c:411-507, 719-1029: Phi = amp sum_lm
// P_l^m(cos theta) [R^c_lm(r) cos m phi + R^s_lm(r) sin m phi] with R_lm =
// r^-(l+1) I_inner(r) + r^l I_outer(r), the cumulative integrals quintic
// piecewise polynomials on the radial grid (value, first and second
// derivative matched at the nodes); below the grid the density is held at
// its innermost value, above it only I_inner(rmax) r^-(l+1) remains.
// p[0] = amp; c: Nr, L, M (1 when axisymmetric), Nt, rmin, rmax, nblk (the
// (l, m, cos|sin) blocks, l-major, cos before sin).  ext: rgrid[Nr]
// (tgrid[Nt]), then per radial interval (and time interval, outermost) and
// block the 12 coefficients of I_inner and I_outer, highest power first, or
// with a time grid their 48 cubic-in-t coefficients [12][4].  Summed
// m-outer with register recurrences for P_l^m (Condon-Shortley phase),
// dP_l^m/dx, cos/sin m phi and r^l, r^-(l+1).
// Not inlined: one copy serves every call site of the extended instances.
struct GhMpSums {
  double F0, F1, F2, V;
};
__device__ __noinline__ GhMpSums gh_multipole_sums(const GhPotTable* __restrict__ T,
                                                   const GhComponent* __restrict__ C, double R,
                                                   double z, double cphi, double sphi, double t,
                                                   bool POT) {
  GhMpSums out = {0.0, 0.0, 0.0, 0.0};
  const int Nr = (int)C->c[0], L = (int)C->c[1], Meff = (int)C->c[2], Nt = (int)C->c[3];
  const int nblk = (int)C->c[6];
  const double rmin = C->c[4], rmax = C->c[5];
  const double* __restrict__ rg = &T->ext[C->ext];
  const double* __restrict__ blk = rg + Nr;
  const double r = sqrt(R * R + z * z);
  if (!(r > 0.0) || !isfinite(r)) return out;
  const double ir = 1.0 / r, x = z * ir, sth = R * ir;
  const int region = (r < rmin) ? 0 : ((r <= rmax) ? 1 : 2);
  int ii = 0;
  double dr = 0.0;
  if (region == 1) {
    ii = gh_spline_find(rg, Nr, r);
    dr = r - rg[ii];
  } else if (region == 2) {
    ii = Nr - 2;
    dr = rmax - rg[Nr - 2];
  }
  int it = 0;
  double dt = 0.0;
  const int stride = (Nt > 0) ? 48 : 12;
  if (Nt > 0) {
    const double* __restrict__ tg = blk;
    it = gh_spline_find(tg, Nt, t);
    dt = t - tg[it];
    blk = tg + Nt + (size_t)it * (Nr - 1) * nblk * 48;
  }
  blk += (size_t)ii * nblk * stride;
  const double ix2m1 = 1.0 / (x * x - 1.0);
  double S0 = 0.0, S1 = 0.0, S2 = 0.0, pot = 0.0;
  double cm = 1.0, sm = 0.0, pmm = 1.0, rm = 1.0, rnm = ir;
  for (int m = 0; m < Meff; m++) {
    if (m > 0) {
      const double c2 = cm * cphi - sm * sphi;
      sm = sm * cphi + cm * sphi;
      cm = c2;
      pmm *= -(2.0 * m - 1.0) * sth;
      rm *= r;
      rnm *= ir;
    }
    double plm1 = 0.0, plm2 = 0.0, rl = rm, rnl = rnm;
    for (int l = m; l < L; l++) {
      double pl;
      if (l == m) pl = pmm;
      else if (l == m + 1) pl = (2.0 * m + 1.0) * x * pmm;
      else pl = ((2.0 * l - 1.0) * x * plm1 - (l + m - 1.0) * plm2) / (l - m);
      const double dpl = (l == 0) ? 0.0 : (l * x * pl - (l + m) * ((l == m) ? 0.0 : plm1)) * ix2m1;
      if (l > m) plm2 = plm1;
      plm1 = pl;
      // block of (l, m, cos): sum over l' < l of 2 min(l' + 1, M) - 1, then 2 m - 1
      const int lb = (l <= Meff) ? l * l : Meff * Meff + (l - Meff) * (2 * Meff - 1);
      const int k0 = lb + ((m == 0) ? 0 : 2 * m - 1);
      for (int trig = 0; trig < ((m == 0) ? 1 : 2); trig++) {
        const double* __restrict__ q = blk + (size_t)(k0 + trig) * stride;
        double c[12];
#pragma unroll
        for (int j = 0; j < 12; j++)
          c[j] = (Nt > 0) ? ((q[4 * j] * dt + q[4 * j + 1]) * dt + q[4 * j + 2]) * dt + q[4 * j + 3] : q[j];
        double Rv, dRv;
        if (region == 1) {
          const double Ii = ((((c[0] * dr + c[1]) * dr + c[2]) * dr + c[3]) * dr + c[4]) * dr + c[5];
          const double Io = ((((c[6] * dr + c[7]) * dr + c[8]) * dr + c[9]) * dr + c[10]) * dr + c[11];
          Rv = rnl * Ii + rl * Io;
          if (!POT) {
            const double dIi = (((5.0 * c[0] * dr + 4.0 * c[1]) * dr + 3.0 * c[2]) * dr + 2.0 * c[3]) * dr + c[4];
            const double dIo = (((5.0 * c[6] * dr + 4.0 * c[7]) * dr + 3.0 * c[8]) * dr + 2.0 * c[9]) * dr + c[10];
            dRv = -(l + 1) * rnl * ir * Ii + rnl * dIi + l * rl * ir * Io + rl * dIo;
          }
        } else if (region == 2) {
          const double Ii = ((((c[0] * dr + c[1]) * dr + c[2]) * dr + c[3]) * dr + c[4]) * dr + c[5];
          Rv = Ii * rnl;
          dRv = -(l + 1) * Ii * rnl * ir;
        } else {
          const double P0 = c[4] / pow(rmin, l + 2.0);
          const double Iie = P0 / (l + 3.0) * pow(r, l + 3.0);
          const double extra = (l == 2) ? P0 * log(rmin / r)
                                        : P0 / (2.0 - l) * (pow(rmin, 2.0 - l) - pow(r, 2.0 - l));
          const double Ioe = c[11] + extra;
          const double a = pow(r, -(l + 1.0)), b = pow(r, (double)l);
          Rv = a * Iie + b * Ioe;
          dRv = -(l + 1) * a * ir * Iie + l * b * ir * Ioe;
        }
        const double cs = trig ? sm : cm, sn = trig ? -cm : sm;  // d(cs)/dphi = -m sn
        if (POT) {
          pot += pl * Rv * cs;
        } else {
          S0 += pl * cs * dRv;
          S1 += dpl * cs * Rv;
          S2 -= pl * m * sn * Rv;
        }
      }
      rl *= r;
      rnl *= ir;
    }
  }
  const double amp = C->p[0];
  out.F0 = -amp * S0;       // -dPhi/dr
  out.F1 = amp * S1 * sth;  // -dPhi/dtheta (dP/dtheta = -sin theta dP/dx)
  out.F2 = -amp * S2;       // -dPhi/dphi
  out.V = amp * pot;
  return out;
}
template <>
__device__ __forceinline__ void gh_comp_force<GH_MULTIPOLE>(const GhPotTable* __restrict__ T, int i,
                                                            const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double R = sqrt(g.R2);
  const double cp = (R > 0.0) ? g.x / R : 1.0, sp = (R > 0.0) ? g.y / R : 0.0;
  const GhMpSums S = gh_multipole_sums(T, C, R, g.z, cp, sp, g.t, false);
  const double ir2 = g.rinv * g.rinv, iR = 1.0 / R;
  F.FRoR += (S.F0 * g.rinv + S.F1 * g.z * ir2 * iR);  // F_R / R
  F.Fz += S.F0 * g.z * g.rinv - S.F1 * R * ir2;
  F.FphioR2 += S.F2 * iR * iR;
}

template <>
__device__ __forceinline__ void gh_comp_force<GH_NULLPOT>(const GhPotTable* __restrict__, int,
                                                          const GhGeom&, GhAcc&) {}

// run-time dispatch on the component type (generic kernels)
__device__ __forceinline__ void gh_comp_force_dyn(const GhPotTable* __restrict__ T, int i,
                                                  const GhGeom& g, GhAcc& F) {
  switch (T->comp[i].type) {
    case GH_POWERSPHCUTOFF: gh_comp_force<GH_POWERSPHCUTOFF>(T, i, g, F); break;
    case GH_MIYAMOTONAGAI: gh_comp_force<GH_MIYAMOTONAGAI>(T, i, g, F); break;
    case GH_NFW: gh_comp_force<GH_NFW>(T, i, g, F); break;
    case GH_HERNQUIST: gh_comp_force<GH_HERNQUIST>(T, i, g, F); break;
    case GH_LOGHALO: gh_comp_force<GH_LOGHALO>(T, i, g, F); break;
    case GH_DEHNENBAR: gh_comp_force<GH_DEHNENBAR>(T, i, g, F); break;
    case GH_POWERSPH: gh_comp_force<GH_POWERSPH>(T, i, g, F); break;
    case GH_JAFFE: gh_comp_force<GH_JAFFE>(T, i, g, F); break;
    case GH_ISOCHRONE: gh_comp_force<GH_ISOCHRONE>(T, i, g, F); break;
    case GH_PLUMMER: gh_comp_force<GH_PLUMMER>(T, i, g, F); break;
    case GH_PSEUDOISO: gh_comp_force<GH_PSEUDOISO>(T, i, g, F); break;
    case GH_KUZMINDISK: gh_comp_force<GH_KUZMINDISK>(T, i, g, F); break;
    case GH_FLATPOWER: gh_comp_force<GH_FLATPOWER>(T, i, g, F); break;
    case GH_KKSTAECKEL: gh_comp_force<GH_KKSTAECKEL>(T, i, g, F); break;
    case GH_ELLIPDISK: gh_comp_force<GH_ELLIPDISK>(T, i, g, F); break;
    case GH_LOPSIDED: gh_comp_force<GH_LOPSIDED>(T, i, g, F); break;
    case GH_STEADYSPIRAL: gh_comp_force<GH_STEADYSPIRAL>(T, i, g, F); break;
    case GH_TRANSIENTSPIRAL: gh_comp_force<GH_TRANSIENTSPIRAL>(T, i, g, F); break;
    case GH_COSMPHIDISK: gh_comp_force<GH_COSMPHIDISK>(T, i, g, F); break;
    case GH_HENONHEILES: gh_comp_force<GH_HENONHEILES>(T, i, g, F); break;
    case GH_SOFTNEEDLEBAR: gh_comp_force<GH_SOFTNEEDLEBAR>(T, i, g, F); break;
    case GH_BURKERT: gh_comp_force<GH_BURKERT>(T, i, g, F); break;
    case GH_DEHNENCORESPH: gh_comp_force<GH_DEHNENCORESPH>(T, i, g, F); break;
    case GH_DEHNENSPH: gh_comp_force<GH_DEHNENSPH>(T, i, g, F); break;
    case GH_HOMOGSPHERE: gh_comp_force<GH_HOMOGSPHERE>(T, i, g, F); break;
    default: break;
  }
}

// Components whose forces are quadratures or expansions over many terms
// (the EllipsoidalPotential family): inlined into the run-time dispatch
// they would raise its register demand for every composition, so they run
// only in the extended dispatch instances (with the wrappers).
__host__ __device__ inline bool gh_is_heavy(int type) {
  return type == GH_TRIAXHERNQUIST || type == GH_TRIAXNFW || type == GH_TRIAXJAFFE ||
         type == GH_TWOPOWERTRIAX || type == GH_PERFECTELLIPSOID || type == GH_TRIAXGAUSSIAN ||
         type == GH_POWERTRIAX || type == GH_EINASTO || type == GH_TWOPOWERSPH ||
         type == GH_EXPTRUNCNFW || type == GH_SPIRALARMS || type == GH_DEDISK ||
         type == GH_INTERPSPH || type == GH_SCF || type == GH_DISKSCF || type == GH_INTERPRZ ||
         type == GH_MULTIPOLE;
}

// The heavy components' forces, out of line: one copy serves every force
// call site of the extended instances (inlined, each call site would carry
// every quadrature and expansion, which multiplies code size and compile
// time); the call's cost is small beside the work of any of them.
__device__ __noinline__ GhAcc gh_heavy_force(const GhPotTable* __restrict__ T, int i, GhGeom g) {
  GhAcc F = {0.0, 0.0, 0.0};
  switch (T->comp[i].type) {
    case GH_TRIAXHERNQUIST:
    case GH_TRIAXNFW:
    case GH_TRIAXJAFFE:
    case GH_TWOPOWERTRIAX:
    case GH_PERFECTELLIPSOID:
    case GH_TRIAXGAUSSIAN:
    case GH_POWERTRIAX: gh_ellip_force<0>(T, i, g, F); break;
    case GH_EINASTO: gh_comp_force<GH_EINASTO>(T, i, g, F); break;
    case GH_EXPTRUNCNFW: gh_comp_force<GH_EXPTRUNCNFW>(T, i, g, F); break;
    case GH_TWOPOWERSPH: gh_comp_force<GH_TWOPOWERSPH>(T, i, g, F); break;
    case GH_SPIRALARMS: gh_comp_force<GH_SPIRALARMS>(T, i, g, F); break;
    case GH_DEDISK: gh_comp_force<GH_DEDISK>(T, i, g, F); break;
    case GH_INTERPSPH: gh_comp_force<GH_INTERPSPH>(T, i, g, F); break;
    case GH_SCF: gh_comp_force<GH_SCF>(T, i, g, F); break;
    case GH_DISKSCF: gh_comp_force<GH_DISKSCF>(T, i, g, F); break;
    case GH_INTERPRZ: gh_comp_force<GH_INTERPRZ>(T, i, g, F); break;
    case GH_MULTIPOLE: gh_comp_force<GH_MULTIPOLE>(T, i, g, F); break;
    default: break;
  }
  return F;
}

// run-time dispatch over every component type (extended instances)
__device__ __forceinline__ void gh_comp_force_ext(const GhPotTable* __restrict__ T, int i,
                                                  const GhGeom& g, GhAcc& F) {
  if (gh_is_heavy(T->comp[i].type)) {
    const GhAcc H = gh_heavy_force(T, i, g);
    F.FRoR += H.FRoR;
    F.Fz += H.Fz;
    F.FphioR2 += H.FphioR2;
  } else {
    gh_comp_force_dyn(T, i, g, F);
  }
}

// A caller's time function A(t) (galpy's pot_tfuncs, Python callbacks that
// the reference calls from C) as a piecewise Chebyshev table built on the
// host from those callbacks over the integration's time range: ext[off] =
// t_a, t_b, nseg, ncoef, nseg / (t_b - t_a), then ncoef coefficients per
// segment; evaluated by Clenshaw's recurrence (segments clamped at the ends).
__host__ __device__ inline double gh_tfunc_eval(const GhPotTable* __restrict__ T, int off,
                                                double t) {
  const double* __restrict__ e = &T->ext[off];
  const int nseg = (int)e[2], nc = (int)e[3];
  const double s = (t - e[0]) * e[4];
  int k = (int)floor(s);
  k = k < 0 ? 0 : (k >= nseg ? nseg - 1 : k);
  const double u = 2.0 * (s - k) - 1.0, u2 = 2.0 * u;
  const double* __restrict__ c = e + 5 + k * nc;
  double b1 = 0.0, b2 = 0.0;
  for (int j = nc - 1; j >= 1; j--) {
    const double b0 = fma(u2, b1, c[j] - b2);
    b2 = b1;
    b1 = b0;
  }
  return fma(u, b1, c[0] - b2);
}

// The wrappers around component C at cylindrical radius R and time t: the
// product of their amplitudes and time factors, the total angle offset (the
// wrapped component is evaluated at phi - offset) and d offset / dR.  Nested
// wrappers compose by multiplying factors and adding offsets, as galpy's
// recursion through calcRforce/calcphitorque does:
//   DehnenSmooth (DehnenSmoothWrapperPotential.c:4-16): amp * smooth(t),
//     Dehnen's quintic between tform and tsteady (1 - smooth when decaying);
//   GaussianAmplitude (GaussianAmplitudeWrapperPotential.c:3-5):
//     amp * exp(-(t - to)^2 / (2 sigma^2));
//   SolidBodyRotation (SolidBodyRotationWrapperPotential.c:4-9):
//     offset omega t + pa;
//   CorotatingRotation (CorotatingRotationWrapperPotential.c:3-15): offset
//     vpo R^(beta-1) (t - to) + pa, whose R derivative adds
//     -phitorque * d offset / dR to the radial force.
__device__ __forceinline__ void gh_wrap_transform(const GhPotTable* __restrict__ T,
                                                  const GhComponent* __restrict__ C, double R,
                                                  double t, double& fac, double& off,
                                                  double& doff) {
  fac = 1.0;
  off = 0.0;
  doff = 0.0;
  for (int k = 0; k < C->nwrap; k++) {
    const GhWrap* __restrict__ W = &T->wrap[C->wrap[k]];
    fac *= W->p[0];
    switch (W->kind) {
      case GH_DEHNENSMOOTHWRAP: {
        const double sm = gh_bar_smooth(t, W->p[1], W->p[2]);
        fac *= (W->p[3] != 0.0) ? sm : 1.0 - sm;
        break;
      }
      case GH_GAUSSIANAMPWRAP: {
        const double d = t - W->p[1];
        fac *= exp(-0.5 * d * d / W->p[2]);
        break;
      }
      case GH_TIMEDEPAMPWRAP: fac *= gh_tfunc_eval(T, (int)W->p[1], t); break;
      case GH_SOLIDBODYROTWRAP: off += W->p[1] * t + W->p[2]; break;
      case GH_COROTATINGROTWRAP: {
        const double dt = t - W->p[4];
        const double Rb = pow(R, W->p[2] - 1.0);
        off += W->p[1] * Rb * dt + W->p[3];
        doff += W->p[1] * (W->p[2] - 1.0) * (Rb / R) * dt;
        break;
      }
      default: break;
    }
  }
}

// Wrappers that move the point where the wrapped components are evaluated,
// walked from the outermost in (gh_wrap_forward: the point, time and
// amplitude factor handed inwards) and back out (gh_wrap_backward: the
// Cartesian force handed outwards); galpy recurses through the same maps:
//   SolidBodyRotation / CorotatingRotation: the point turned by -offset
//     about z; the force turned back, the corotating wrapper adding
//     -F_phi d offset / dR along R (CorotatingRotationWrapperPotential.c:16-40);
//   RotateAndTilt (RotateAndTiltWrapperPotential.c:21-62): x' = rot x +
//     offset; F = rot^T F';
//   KuzminLike (KuzminLikeWrapperPotential.c:5-92): the point (xi, 0, 0),
//     xi = sqrt(R^2 + (a + sqrt(z^2 + b^2))^2); its radial force F' gives
//     F_R = F' R / xi and F_z = F' (a + s) z / (xi s), s = sqrt(z^2 + b^2);
//   CylindricallySeparable (CylindricallySeparablePotentialWrapper.c:14-27):
//     two passes at t = 0, the radial force at (R, 0, 0) and the vertical
//     force at (Rp, 0, z); every map is linear in the force, so the passes'
//     forces add;
//   MovingObject (MovingObjectPotential.c:12-88): the wrapped potential at
//     (Rdist, phi, z_obj - z), Rdist the cylindrical distance to the object
//     (positions from natural splines over the normalised time); its radial
//     force F' gives F = -F' (x_obj - x, y_obj - y) / Rdist, F_z = -F_z';
//   DehnenSmooth, GaussianAmplitude, TimeDependentAmplitude: factors only.
//   OblateStaeckel: the point on the v = v0 curve at the orbit's u (pass 0)
//     or on the u = u0 curve at its v (pass 1), phi = t = 0; the wrapper's
//     forces are linear in the component's potential and forces there
//     (OblateStaeckelWrapperPotential.c:31-229).
// rec (when not null, GH_WREC doubles) keeps what the way back needs: {cos,
// sin, d offset/dR, x, y} for the rotations, {x, y, (a + s) z / s, xi} for
// KuzminLike, {x, y, R} for CylindricallySeparable, {cos, sin, the
// coefficients of (Phi, F_R, F_z) in F_R and in F_z, the constant terms,
// the potential's coefficient and constant} for OblateStaeckel.
#define GH_WREC 12
__device__ __forceinline__ void gh_wrap_forward(const GhPotTable* __restrict__ T,
                                                const GhWrap* __restrict__ W, int pass,
                                                double& x, double& y, double& z, double& t,
                                                double& fac, double* rec) {
  fac *= W->p[0];
  switch (W->kind) {
    case GH_DEHNENSMOOTHWRAP: {
      const double sm = gh_bar_smooth(t, W->p[1], W->p[2]);
      fac *= (W->p[3] != 0.0) ? sm : 1.0 - sm;
      break;
    }
    case GH_GAUSSIANAMPWRAP: {
      const double d = t - W->p[1];
      fac *= exp(-0.5 * d * d / W->p[2]);
      break;
    }
    case GH_TIMEDEPAMPWRAP: fac *= gh_tfunc_eval(T, (int)W->p[1], t); break;
    case GH_SOLIDBODYROTWRAP:
    case GH_COROTATINGROTWRAP: {
      double off, doff = 0.0;
      if (W->kind == GH_SOLIDBODYROTWRAP) {
        off = W->p[1] * t + W->p[2];
      } else {
        const double R = sqrt(x * x + y * y), dt = t - W->p[4];
        const double Rb = pow(R, W->p[2] - 1.0);
        off = W->p[1] * Rb * dt + W->p[3];
        doff = W->p[1] * (W->p[2] - 1.0) * (Rb / R) * dt;
      }
      const GhSinCos sc = gh_sincos(off);
      if (rec) {
        rec[0] = sc.c;
        rec[1] = sc.s;
        rec[2] = doff;
        rec[3] = x;
        rec[4] = y;
      }
      const double xn = x * sc.c + y * sc.s;
      y = y * sc.c - x * sc.s;
      x = xn;
      break;
    }
    case GH_ROTATETILTWRAP: {
      if (W->p[10] != 0.0) {
        const double* __restrict__ m = &W->p[1];
        const double xn = m[0] * x + m[1] * y + m[2] * z;
        const double yn = m[3] * x + m[4] * y + m[5] * z;
        z = m[6] * x + m[7] * y + m[8] * z;
        x = xn;
        y = yn;
      }
      if (W->p[11] != 0.0) {
        x += W->p[12];
        y += W->p[13];
        z += W->p[14];
      }
      break;
    }
    case GH_MOVINGOBJWRAP: {  // p: amp, t0, tf, spline offsets of x, y, z (z < 0: planar)
      double d = (t - W->p[1]) / (W->p[2] - W->p[1]);
      d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
      const double dx = gh_spline_eval(T, (int)W->p[3], d) - x;
      const double dy = gh_spline_eval(T, (int)W->p[4], d) - y;
      const double oz = (W->p[5] >= 0.0) ? gh_spline_eval(T, (int)W->p[5], d) : 0.0;
      const double R = sqrt(x * x + y * y), Rd = sqrt(dx * dx + dy * dy);
      const double cp = (R > 0.0) ? x / R : 1.0, sp = (R > 0.0) ? y / R : 0.0;
      if (rec) {
        rec[0] = dx;
        rec[1] = dy;
        rec[2] = Rd;
        rec[3] = cp;
        rec[4] = sp;
      }
      x = Rd * cp;  // the wrapped potential at (Rdist, phi, z_obj - z)
      y = Rd * sp;
      z = oz - z;
      break;
    }
    case GH_KUZMINLIKEWRAP: {
      const double sz = sqrt(z * z + W->p[2]), as = W->p[1] + sz;
      const double xi = sqrt(x * x + y * y + as * as);
      if (rec) {
        rec[0] = x;
        rec[1] = y;
        rec[2] = as * z / sz;
        rec[3] = xi;
      }
      x = xi;
      y = 0.0;
      z = 0.0;
      break;
    }
    case GH_OBLATESTAECKELWRAP: {  // p: amp, delta, u0, v0, refpot, planar
      const double d = W->p[1], u0 = W->p[2], v0 = W->p[3];
      const bool planar = W->p[5] != 0.0;
      const double R = sqrt(x * x + y * y), zz = planar ? 0.0 : z;
      const double d1 = sqrt((zz + d) * (zz + d) + R * R), d2 = sqrt((zz - d) * (zz - d) + R * R);
      const double u = acosh(0.5 / d * (d1 + d2)), v = acos(0.5 / d * (d1 - d2));
      const double sh = sinh(u), ch = cosh(u), sv = sin(v), cv = cos(v);
      double xm, zm;
      if (pass == 0) {
        xm = d * sh * sin(v0);
        zm = d * ch * cos(v0);
      } else {
        xm = d * sinh(u0) * sv;
        zm = d * cosh(u0) * cv;
      }
      if (rec) {
        const double th = tanh(u), tv = tan(v);
        const double pre = sh * sh + sv * sv, dpu = 2.0 * sh * ch, dpv = 2.0 * sv * cv;
        const double K = 1.0 / ((d * pre) * (d * pre));
        const double A1 = d * sv * ch, A2 = th * zz, B1 = R / tv, B2 = A1;
        const double G = (dpu * A1 + dpv * A2) / pre, H = (dpu / tv * R - dpv * A1) / pre;
        rec[0] = (R > 0.0) ? x / R : 1.0;
        rec[1] = (R > 0.0) ? y / R : 0.0;
        if (pass == 0) {  // U = ch^2 Phi, dU/du = 2 ch sh Phi - ch^2 (F_R R / (th + 1e-12) + F_z z th)
          const double c2 = ch * ch, fr = c2 * xm / (th + 1e-12), fz = planar ? 0.0 : c2 * zm * th;
          rec[2] = K * (-2.0 * A1 * ch * sh + G * c2);
          rec[3] = K * A1 * fr;
          rec[4] = K * A1 * fz;
          rec[5] = K * (-2.0 * B1 * ch * sh + H * c2);
          rec[6] = K * B1 * fr;
          rec[7] = K * B1 * fz;
          rec[8] = rec[9] = 0.0;
          rec[10] = c2 / pre;
          rec[11] = 0.0;
        } else if (planar) {  // the planar force has no V
          for (int j = 2; j < 10; j++) rec[j] = 0.0;
          const double P0 = sinh(u0) * sinh(u0) + sv * sv;
          rec[10] = P0 / pre;
          rec[11] = -W->p[4] / pre;
        } else {  // V = refpot - P0 Phi, dV/dv = -2 sv cv Phi + P0 (F_R R / tv - F_z z tv)
          const double P0 = sinh(u0) * sinh(u0) + sv * sv, refpot = W->p[4];
          const double fr = P0 * xm / tv, fz = -P0 * zm * tv;
          rec[2] = K * (-2.0 * A2 * sv * cv + G * P0);
          rec[3] = K * A2 * fr;
          rec[4] = K * A2 * fz;
          rec[5] = K * (2.0 * B2 * sv * cv + H * P0);
          rec[6] = -K * B2 * fr;
          rec[7] = -K * B2 * fz;
          rec[8] = -K * G * refpot;
          rec[9] = -K * H * refpot;
          rec[10] = P0 / pre;
          rec[11] = -refpot / pre;
        }
      }
      x = xm;
      y = 0.0;
      z = zm;
      t = 0.0;
      break;
    }
    case GH_CYLSEPWRAP: {
      const double R = sqrt(x * x + y * y);
      if (rec) {
        rec[0] = x;
        rec[1] = y;
        rec[2] = R;
      }
      x = (pass == 0) ? R : W->p[1];
      y = 0.0;
      if (pass == 0) z = 0.0;
      t = 0.0;
      break;
    }
    default: break;
  }
}
__device__ __forceinline__ void gh_wrap_backward(const GhWrap* __restrict__ W, int pass,
                                                 const double* rec, double pot, bool first,
                                                 double& fx, double& fy, double& fz) {
  switch (W->kind) {
    case GH_OBLATESTAECKELWRAP: {  // fx, fz: the component's F_R, F_z at (x, 0, z)
      double FR = rec[2] * pot + rec[3] * fx + rec[4] * fz;
      double Fz = rec[5] * pot + rec[6] * fx + rec[7] * fz;
      if (first) {
        FR += rec[8];
        Fz += rec[9];
      }
      fx = FR * rec[0];
      fy = FR * rec[1];
      fz = Fz;
      break;
    }
    case GH_SOLIDBODYROTWRAP:
    case GH_COROTATINGROTWRAP: {
      const double c = rec[0], s = rec[1];
      const double gx = c * fx - s * fy;
      fy = s * fx + c * fy;
      fx = gx;
      if (W->kind == GH_COROTATINGROTWRAP) {
        const double xk = rec[3], yk = rec[4];
        const double corr = -(xk * fy - yk * fx) * rec[2] * gh_rsqrt(xk * xk + yk * yk);
        fx += corr * xk;
        fy += corr * yk;
      }
      break;
    }
    case GH_ROTATETILTWRAP: {
      if (W->p[10] != 0.0) {
        const double* __restrict__ m = &W->p[1];
        const double gx = m[0] * fx + m[3] * fy + m[6] * fz;
        const double gy = m[1] * fx + m[4] * fy + m[7] * fz;
        fz = m[2] * fx + m[5] * fy + m[8] * fz;
        fx = gx;
        fy = gy;
      }
      break;
    }
    case GH_MOVINGOBJWRAP: {  // MovingObjectPotential.c:12-88
      const double FRw = rec[3] * fx + rec[4] * fy;  // wrapped radial force at (Rdist, phi)
      const double f = -FRw / rec[2];
      fx = f * rec[0];
      fy = f * rec[1];
      fz = -fz;
      break;
    }
    case GH_KUZMINLIKEWRAP: {  // the inner point is (xi, 0, 0): its radial force is fx
      const double f = fx / rec[3];
      fx = f * rec[0];
      fy = f * rec[1];
      fz = f * rec[2];
      break;
    }
    case GH_CYLSEPWRAP: {
      if (pass == 0) {  // radial force at (R, 0, 0)
        const double f = fx / rec[2];
        fx = f * rec[0];
        fy = f * rec[1];
        fz = 0.0;
      } else {  // vertical force at (Rp, 0, z)
        fx = 0.0;
        fy = 0.0;
      }
      break;
    }
    default: break;
  }
}

// Force of a component whose wrappers move the evaluation point: the leaf's
// force at the moved point, mapped back out through the wrappers (one pass,
// two with a CylindricallySeparable wrapper).
__device__ __forceinline__ double gh_comp_potential(const GhPotTable* __restrict__ T, int i,
                                                    double R, double z, double cosphi,
                                                    double sinphi, double t);
// Out of line: one copy serves every call site (the walk carries the full
// dispatch and, for OblateStaeckel, the potential's).
__device__ __noinline__ GhAcc gh_comp_force_mapped(const GhPotTable* __restrict__ T, int i,
                                                   GhGeom g) {
  GhAcc F = {0.0, 0.0, 0.0};
  const GhComponent* __restrict__ C = &T->comp[i];
  const int nw = C->nwrap;
  double Fx = 0.0, Fy = 0.0, Fz = 0.0;
  for (int pass = 0; pass < C->mapwrap; pass++) {
    double x = g.x, y = g.y, z = g.z, t = g.t, fac = 1.0;
    double rec[GH_WRAP_DEPTH][GH_WREC];
#pragma unroll
    for (int k = 0; k < GH_WRAP_DEPTH; k++)
      if (k < nw) gh_wrap_forward(T, &T->wrap[C->wrap[k]], pass, x, y, z, t, fac, rec[k]);
    GhGeom gl;
    gl.x = x;
    gl.y = y;
    gl.z = z;
    gl.R2 = x * x + y * y;
    gl.z2 = z * z;
    gl.r2 = gl.R2 + gl.z2;
    gl.rinv = gh_rsqrt(gl.r2);
    gl.r = gl.r2 * gl.rinv;
    gl.t = t;
    GhAcc L = {0.0, 0.0, 0.0};
    gh_comp_force_ext(T, i, gl, L);
    double fx = x * L.FRoR - y * L.FphioR2, fy = y * L.FRoR + x * L.FphioR2, fz = L.Fz;
    double pot = 0.0;
    if (C->mappot) {
      const double R = sqrt(gl.R2);
      pot = gh_comp_potential(T, i, R, z, (R > 0.0) ? x / R : 1.0, (R > 0.0) ? y / R : 0.0, t);
    }
#pragma unroll
    for (int k = GH_WRAP_DEPTH - 1; k >= 0; k--)
      if (k < nw) {
        const GhWrap* __restrict__ W = &T->wrap[C->wrap[k]];
        gh_wrap_backward(W, pass, rec[k], pot, W->first == i, fx, fy, fz);
      }
    Fx += fac * fx;
    Fy += fac * fy;
    Fz += fac * fz;
  }
  const double iR2 = gh_rcp(g.R2);
  F.FRoR = (g.x * Fx + g.y * Fy) * iR2;
  F.FphioR2 = (g.x * Fy - g.y * Fx) * iR2;
  F.Fz = Fz;
  return F;
}

// Force of component i with the wrappers around it: the component's own
// force at the position rotated by -offset (its F_R, F_z and F_phi are the
// same cylindrical components in either frame), scaled by the wrappers'
// factor, with the corotating term -F_phi d offset / dR in F_R (in the
// accumulated form: F_R / R gains -(F_phi / R^2) R d offset / dR).  One
// inlined dispatch serves wrapped and bare components alike, so the
// run-time dispatch kernels carry no call.
__device__ __forceinline__ void gh_comp_force_any(const GhPotTable* __restrict__ T, int i,
                                                  const GhGeom& g, GhAcc& F) {
  const GhComponent* __restrict__ C = &T->comp[i];
  if (C->densonly) return;
  if (C->mapwrap) {
    const GhAcc M = gh_comp_force_mapped(T, i, g);
    F.FRoR += M.FRoR;
    F.Fz += M.Fz;
    F.FphioR2 += M.FphioR2;
    return;
  }
  GhGeom gw = g;
  double fac = 1.0, corr = 0.0;
  const bool wrapped = C->nwrap != 0;
  if (wrapped) {
    const double R = sqrt(g.R2);
    double off, doff;
    gh_wrap_transform(T, C, R, g.t, fac, off, doff);
    const GhSinCos sc = gh_sincos(off);
    const double so = sc.s, co = sc.c;
    gw.x = g.x * co + g.y * so;
    gw.y = g.y * co - g.x * so;
    corr = R * doff;
  }
  GhAcc L = {0.0, 0.0, 0.0};
  gh_comp_force_ext(T, i, gw, L);
  // scalars per branch, F updated once (no field-address select)
  const double fr = wrapped ? fac * (L.FRoR - L.FphioR2 * corr) : L.FRoR;
  const double fz = wrapped ? fac * L.Fz : L.Fz;
  const double fp = wrapped ? fac * L.FphioR2 : L.FphioR2;
  F.FRoR += fr;
  F.Fz += fz;
  F.FphioR2 += fp;
}

// A potential composition known at compile time: GhSpec<15, 5, 9> is
// MWPotential2014; GhSpec<> is any composition, dispatched at run time.
// Specialised kernels inline only the components they use, which keeps the
// many inlined force evaluations of DOP853 inside the instruction cache.
template <int... Codes>
struct GhSpec {};

template <class Spec>
struct GhAccSum;

template <>
struct GhAccSum<GhSpec<>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, const GhGeom& g,
                                             GhAcc& F) {
    const int npot = T->npot;
    for (int i = 0; i < npot; i++) gh_comp_force_dyn(T, i, g, F);
  }
};

// Any composition with wrapped or heavy components: its own kernel
// instances, so that neither the wrapper transform nor the quadrature
// forces add to the run-time dispatch kernels of the other compositions.
struct GhSpecWrapped {};

template <>
struct GhAccSum<GhSpecWrapped> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, const GhGeom& g,
                                             GhAcc& F) {
    const int npot = T->npot;
    for (int i = 0; i < npot; i++) gh_comp_force_any(T, i, g, F);
  }
};

template <int... Codes>
struct GhAccSum<GhSpec<Codes...>> {
  __device__ static __forceinline__ void add(const GhPotTable* __restrict__ T, const GhGeom& g,
                                             GhAcc& F) {
    int i = 0;
    (gh_comp_force<Codes>(T, i++, g, F), ...);  // component order = summation order
  }
};

// Sum over components in index order (galpy_potentials.c:120-179), from the
// rectangular position.
template <class Spec = GhSpec<>>
__device__ __forceinline__ GhAcc gh_acc_forces(const GhPotTable* __restrict__ T, double x,
                                               double y, double z, double t) {
  GhGeom g;
  g.x = x;
  g.y = y;
  g.z = z;
  g.R2 = x * x + y * y;
  g.z2 = z * z;
  g.r2 = g.R2 + g.z2;
  g.rinv = gh_rsqrt(g.r2);  // shared by the spherical components
  g.r = g.r2 * g.rinv;
  g.t = t;
  GhAcc F = {0.0, 0.0, 0.0};
  GhAccSum<Spec>::add(T, g, F);
  return F;
}

// Cylindrical forces (F_R, F_z, F_phi) at (R, z, phi); cosphi/sinphi = x/R, y/R.
struct GhCyl {
  double FR, Fz, Fphi;
};

template <class Spec = GhSpec<>>
__device__ __forceinline__ GhCyl gh_cyl_forces(const GhPotTable* __restrict__ T, double R,
                                               double R2, double z, double cosphi,
                                               double sinphi, double t) {
  const GhAcc F = gh_acc_forces<Spec>(T, R * cosphi, R * sinphi, z, t);
  return GhCyl{R * F.FRoR, F.Fz, R2 * F.FphioR2};
}

// evalRectForce (integrateFullOrbit.c:1307-1325): rectangular acceleration.
template <class Spec = GhSpec<>>
__device__ __forceinline__ void gh_rect_force(const GhPotTable* __restrict__ T, double t,
                                              double x, double y, double z, double& ax,
                                              double& ay, double& az) {
  const GhAcc F = gh_acc_forces<Spec>(T, x, y, z, t);
  ax = x * F.FRoR - y * F.FphioR2;
  ay = y * F.FRoR + x * F.FphioR2;
  az = F.Fz;
}

// NonInertialFrameForce.c:27-170: the fictitious forces of a frame whose
// origin accelerates by a0(t) and that rotates at Omega(t) (centrifugal,
// Coriolis and Euler terms; with the origin's x0, v0 when it both rotates
// and accelerates), in Cartesian form times amp.  p: amp, rot_acc, lin_acc,
// omegaz_only, const_freq, Omega_as_func, interp (45), tmin; c: Omega[3],
// Omega^2, Omegadot[3], tmax; ext[C->ext + k]: the table (pot_type 39: the
// caller's function k tabulated, Omegadot its own) or spline (45: natural
// spline k, Omegadot its derivative, times clamped to [tmin, tmax]) of time
// input k in galpy's order (a0, x0, v0, then Omega).
__device__ __noinline__ void gh_noninertial_xyz(const GhPotTable* __restrict__ T,
                                                const GhComponent* __restrict__ C, double t,
                                                double x, double y, double z, double vx,
                                                double vy, double vz, double& Fx, double& Fy,
                                                double& Fz) {
  const double* __restrict__ e = &T->ext[C->ext];
  const bool interp = C->p[6] != 0.0;
  const double tq = interp ? fmin(fmax(t, C->p[7]), C->c[7]) : t;
  auto tdep = [&](int k) {
    return interp ? gh_spline_eval(T, (int)e[k], tq) : gh_tfunc_eval(T, (int)e[k], t);
  };
  auto tdepd = [&](int sk, int tk) {
    return interp ? gh_spline_deriv(T, (int)e[sk], tq) : gh_tfunc_eval(T, (int)e[tk], t);
  };
  const bool rot_acc = C->p[1] != 0.0, lin_acc = C->p[2] != 0.0, as_func = C->p[5] != 0.0;
  const int la = lin_acc ? 9 : 0;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  if (rot_acc) {
    const bool zonly = C->p[3] != 0.0, const_freq = C->p[4] != 0.0;
    if (zonly) {
      double Oz, O2, Odz = 0.0;
      if (as_func) {
        Oz = tdep(la);
        O2 = Oz * Oz;
      } else {
        Oz = C->c[2];
        if (!const_freq) {
          Oz += C->c[6] * t;
          O2 = Oz * Oz;
        } else {
          O2 = C->c[3];
        }
      }
      fx += 2.0 * Oz * vy + O2 * x;
      fy += -2.0 * Oz * vx + O2 * y;
      if (!const_freq) {
        Odz = as_func ? tdepd(la, la + 1) : C->c[6];
        fx += Odz * y;
        fy -= Odz * x;
      }
      if (lin_acc) {
        const double x0x = tdep(3), x0y = tdep(4), v0x = tdep(6), v0y = tdep(7);
        fx += 2.0 * Oz * v0y + O2 * x0x;
        fy += -2.0 * Oz * v0x + O2 * x0y;
        if (!const_freq) {
          fx += Odz * x0y;
          fy -= Odz * x0x;
        }
      }
    } else {
      double Ox, Oy, Oz, O2, Odx = 0.0, Ody = 0.0, Odz = 0.0;
      if (as_func) {
        Ox = tdep(la);
        Oy = tdep(la + 1);
        Oz = tdep(la + 2);
        O2 = Ox * Ox + Oy * Oy + Oz * Oz;
      } else {
        Ox = C->c[0];
        Oy = C->c[1];
        Oz = C->c[2];
        if (!const_freq) {
          Ox += C->c[4] * t;
          Oy += C->c[5] * t;
          Oz += C->c[6] * t;
          O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        } else {
          O2 = C->c[3];
        }
      }
      double Ov = Ox * x + Oy * y + Oz * z;
      fx += 2.0 * (Oz * vy - Oy * vz) + O2 * x - Ox * Ov;
      fy += -2.0 * (Oz * vx - Ox * vz) + O2 * y - Oy * Ov;
      fz += 2.0 * (Oy * vx - Ox * vy) + O2 * z - Oz * Ov;
      if (!const_freq) {
        if (as_func) {
          Odx = tdepd(la, la + 3);
          Ody = tdepd(la + 1, la + 4);
          Odz = tdepd(la + 2, la + 5);
        } else {
          Odx = C->c[4];
          Ody = C->c[5];
          Odz = C->c[6];
        }
        fx -= -Odz * y + Ody * z;
        fy -= Odz * x - Odx * z;
        fz -= -Ody * x + Odx * y;
      }
      if (lin_acc) {
        const double x0x = tdep(3), x0y = tdep(4), x0z = tdep(5);
        const double v0x = tdep(6), v0y = tdep(7), v0z = tdep(8);
        Ov = Ox * x0x + Oy * x0y + Oz * x0z;
        fx += 2.0 * (Oz * v0y - Oy * v0z) + O2 * x0x - Ox * Ov;
        fy += -2.0 * (Oz * v0x - Ox * v0z) + O2 * x0y - Oy * Ov;
        fz += 2.0 * (Oy * v0x - Ox * v0y) + O2 * x0z - Oz * Ov;
        if (!const_freq) {
          fx -= -Odz * x0y + Ody * x0z;
          fy -= Odz * x0x - Odx * x0z;
          fz -= -Ody * x0x + Odx * x0y;
        }
      }
    }
  }
  if (lin_acc) {
    fx -= tdep(0);
    fy -= tdep(1);
    fz -= tdep(2);
  }
  Fx = C->p[0] * fx;
  Fy = C->p[0] * fy;
  Fz = C->p[0] * fz;
}

// The densities dynamical friction needs (the Dens functions of
// HernquistPotential.c, NFWPotential.c, PlummerPotential.c,
// MiyamotoNagaiPotential.c, PowerSphericalPotentialwCutoff.c,
// LogarithmicHaloPotential.c, JaffePotential.c, IsochronePotential.c,
// PowerSphericalPotential.c, BurkertPotential.c), from galpy's arguments p.
__host__ __device__ inline bool gh_has_density(int type) {
  return type == GH_HERNQUIST || type == GH_NFW || type == GH_PLUMMER ||
         type == GH_MIYAMOTONAGAI || type == GH_POWERSPHCUTOFF || type == GH_LOGHALO ||
         type == GH_JAFFE || type == GH_ISOCHRONE || type == GH_POWERSPH || type == GH_BURKERT;
}
__device__ inline double gh_comp_density(const GhComponent* __restrict__ C, double R, double z,
                                         double sphi, double cphi) {
  const double* __restrict__ p = C->p;
  const double r2 = R * R + z * z, r = sqrt(r2);
  switch (C->type) {
    case GH_HERNQUIST: return p[0] * M_1_PI / 4.0 / p[1] / p[1] / r * pow(1.0 + r / p[1], -3.0);
    case GH_NFW: return p[0] * M_1_PI / 4.0 / p[1] / p[1] / (1.0 + r / p[1]) / (1.0 + r / p[1]) / r;
    case GH_PLUMMER: {
      const double b2 = p[1] * p[1];
      return 3.0 * p[0] * M_1_PI / 4.0 * b2 * pow(r2 + b2, -2.5);
    }
    case GH_MIYAMOTONAGAI: {
      const double a = p[1], b2 = p[2] * p[2], sbz = sqrt(b2 + z * z);
      const double asbz = (a + sbz) * (a + sbz);
      if (a == 0.0) return 3.0 * p[0] * M_1_PI / 4.0 * b2 * pow(R * R + sbz * sbz, -2.5);
      return p[0] * M_1_PI / 4.0 * b2 * (a * R * R + (a + 3.0 * sbz) * asbz) *
             pow(R * R + asbz, -2.5) * pow(sbz, -3.0);
    }
    case GH_POWERSPHCUTOFF: return p[0] * pow(r, -p[1]) * exp(-r2 / p[2] / p[2]);
    case GH_LOGHALO: {
      const double amp = p[0], q = p[1], c = p[2], om = p[3], zq = z / q, q2 = q * q;
      if (om < 1.0) {
        const double R2 = R * R, Rt2 = R2 * (1.0 - om * sphi * sphi);
        const double dn = 1.0 / (Rt2 + zq * zq + c), dn2 = dn * dn;
        const double s2p = 2.0 * sphi * cphi, c2p = cphi * cphi - sphi * sphi;
        return amp * M_1_PI / 4.0 *
               (2.0 * Rt2 / R2 * (dn - Rt2 * dn2) + dn / q2 - 2.0 * zq * zq * dn2 / q2 -
                om * (2.0 * R2 * s2p * s2p / 4.0 * om * dn2 + dn * c2p));
      }
      return amp * M_1_PI / 4.0 / q2 * ((2.0 * q2 + 1.0) * c + R * R + (2.0 - 1.0 / q2) * z * z) /
             ((R * R + zq * zq + c) * (R * R + zq * zq + c));
    }
    case GH_JAFFE: return p[0] * M_1_PI / 4.0 / p[1] * pow(r * (1.0 + r / p[1]), -2.0);
    case GH_ISOCHRONE: {
      const double b = p[1], rb = sqrt(r2 + b * b), brbrb = (b + rb) * rb;
      return p[0] * M_1_PI / 4.0 * (3.0 * brbrb * rb - r2 * (b + 3.0 * rb)) * pow(brbrb, -3.0);
    }
    case GH_POWERSPH: return p[0] * M_1_PI / 4.0 * (3.0 - p[1]) * pow(r2, -0.5 * p[1]);
    case GH_BURKERT: {
      const double x = r / p[1];
      return p[0] / (1.0 + x) / (1.0 + x * x);
    }
    default: return 0.0;
  }
}
// ChandrasekharDynamicalFrictionForce.c:13-140: F = A v, A = -amp Xfactor(X)
// lnLambda rho / v^3, X = v / (sqrt 2 sigma_r(r)) with sigma_r a natural
// spline over (r - ro) / (rf - ro) clamped to [0, 1], Xfactor = erf X -
// (2/sqrt pi) X e^-X^2, lnLambda constant or 0.5 ln(1 + r^2 / (gamma^2
// max(rhm, GM/v^2)^2)); zero inside minr.  W->p: amp, ms, rhm, gamma^2,
// lnLambda (< 0: variable), minr^2, ro, rf, spline offset, the number of
// density components (from W->first).
__device__ __noinline__ double gh_chandrasekhar_amp(const GhPotTable* __restrict__ T,
                                                    const GhWrap* __restrict__ W, double x,
                                                    double y, double z, double v2) {
  const double R2 = x * x + y * y, r2 = R2 + z * z;
  if (r2 < W->p[5]) return 0.0;
  const double amp = W->p[0], ms = W->p[1], rhm = W->p[2], gamma2 = W->p[3];
  const double ro = W->p[6], rf = W->p[7];
  double lnL = W->p[4];
  const double r = sqrt(r2), v = sqrt(v2);
  if (lnL < 0.0) {
    const double GMvs = ms / v / v;
    lnL = (GMvs < rhm) ? 0.5 * log(1.0 + r2 / gamma2 / rhm / rhm)
                       : 0.5 * log(1.0 + r2 / gamma2 / GMvs / GMvs);
  }
  double d = (r - ro) / (rf - ro);
  d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
  const double sr = gh_spline_eval(T, (int)W->p[8], d);
  const double X = M_SQRT1_2 * v / sr;
  const double Xf = erf(X) - M_2_SQRTPI * X * exp(-X * X);
  const double R = sqrt(R2);
  const double sphi = (R > 0.0) ? y / R : 0.0, cphi = (R > 0.0) ? x / R : 1.0;
  double dens = 0.0;
  const int n = (int)W->p[9];
  for (int k = 0; k < n; k++) dens += gh_comp_density(&T->comp[W->first + k], R, z, sphi, cphi);
  return -amp * Xf * lnL / v2 / v * dens;
}

// The cosine integral Ci(x), x > 0 (gsl_sf_Ci, which
// FDMDynamicalFrictionForce.c calls): the power series gamma + ln x +
// sum_k (-x^2)^k / (2k (2k)!) for x <= 4, else -Re E1(i x) from E1's
// continued fraction (A&S 5.1.22) by the modified Lentz method.
__device__ inline double gh_cos_integral(double x) {
  const double euler = 0.57721566490153286061;
  if (x <= 4.0) {
    double term = 1.0, sum = 0.0;
    for (int k = 1; k < 60; k++) {
      term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
      const double add = term / (2.0 * k);
      sum += add;
      if (fabs(add) < 1e-17 * fabs(sum)) break;
    }
    return euler + log(x) + sum;
  }
  // complex (re, im) arithmetic; z = i x, b = z + 1, c = 1/tiny, d = h = 1/b
  double br = 1.0, bi = x;
  double cr = 1e300, ci = 0.0;
  double inv = 1.0 / (br * br + bi * bi);
  double dr = br * inv, di = -bi * inv, hr = dr, hi = di;
  for (int n = 1; n < 1000; n++) {
    const double an = -double(n) * n;
    br += 2.0;
    // d = 1 / (an d + b)
    double tr = an * dr + br, ti = an * di + bi;
    inv = 1.0 / (tr * tr + ti * ti);
    dr = tr * inv;
    di = -ti * inv;
    // c = b + an / c
    inv = 1.0 / (cr * cr + ci * ci);
    tr = br + an * cr * inv;
    ti = bi - an * ci * inv;
    cr = tr;
    ci = ti;
    const double er = cr * dr - ci * di, ei = cr * di + ci * dr;  // del = c d
    tr = hr * er - hi * ei;
    hi = hr * ei + hi * er;
    hr = tr;
    if (fabs(er - 1.0) + fabs(ei) < 1e-16) break;
  }
  // -Re(h e^{-i x})
  return -(hr * cos(x) + hi * sin(x));
}
// FDMDynamicalFrictionForce.c:14-80: FDMfactor / CDMfactor for the
// Chandrasekhar entry W (p[10] = m/hbar, p[11] = constant factor or < 0).
__device__ __noinline__ double gh_fdm_ratio(const GhPotTable* __restrict__ T,
                                            const GhWrap* __restrict__ W, double r2, double v2) {
  const double ms = W->p[1], rhm = W->p[2], gamma2 = W->p[3], ro = W->p[6], rf = W->p[7];
  const double mhbar = W->p[10], cf = W->p[11];
  double lnL = W->p[4];
  const double r = sqrt(r2), v = sqrt(v2);
  const double GMvs = ms / v / v;
  if (lnL < 0.0)
    lnL = (GMvs < rhm) ? 0.5 * log(1.0 + r2 / gamma2 / rhm / rhm)
                       : 0.5 * log(1.0 + r2 / gamma2 / GMvs / GMvs);
  double d = (r - ro) / (rf - ro);
  d = d < 0.0 ? 0.0 : (d > 1.0 ? 1.0 : d);
  const double sr = gh_spline_eval(T, (int)W->p[8], d);
  const double X = M_SQRT1_2 * v / sr;
  const double Xf = erf(X) - M_2_SQRTPI * X * exp(-X * X);
  const double CDM = lnL * Xf;
  const double euler = 0.57721566490153286061;
  double FDM;
  if (cf < 0.0) {
    const double kr = 2.0 * mhbar * v * r, Ms = v / sr;
    if (kr < Ms) {
      FDM = euler + log(kr) - gh_cos_integral(kr) + sin(kr) / kr - 1.0;
    } else if (kr > 4 * Ms) {
      FDM = log(kr / Ms) * Xf;
    } else {
      const double Cf = euler + log(Ms) - gh_cos_integral(Ms) + sin(Ms) / Ms - 1.0;
      const double Cd = log(4.0) * Xf;
      const double mu = (2 * Ms - kr / 2) / (2 * Ms - Ms / 2);
      FDM = mu * Cf + (1 - mu) * Cd;
    }
  } else {
    FDM = cf;
  }
  return FDM / CDM;
}

// The velocity-dependent components' Cartesian forces at (x, v, t)
// (calcRforce / calczforce / calcphitorque with velocities,
// galpy_potentials.c), added to a.
__device__ __noinline__ void gh_diss_rect(const GhPotTable* __restrict__ T, double t, double x,
                                          double y, double z, double vx, double vy, double vz,
                                          double& ax, double& ay, double& az) {
  for (int i = 0; i < T->npot; i++) {
    const GhComponent* __restrict__ C = &T->comp[i];
    if (C->type == GH_NONINERTIAL || C->type == GH_NONINERTIALINTERP) {
      double fx, fy, fz;
      gh_noninertial_xyz(T, C, t, x, y, z, vx, vy, vz, fx, fy, fz);
      ax += fx;
      ay += fy;
      az += fz;
    }
  }
  for (int w = 0; w < T->nwraps; w++) {
    const GhWrap* __restrict__ W = &T->wrap[w];
    if (W->kind != GH_CHANDRASEKHAR) continue;
    const double v2 = vx * vx + vy * vy + vz * vz;
    double A = gh_chandrasekhar_amp(T, W, x, y, z, v2);
    if (W->p[12] != 0.0 && A != 0.0) {  // FDMDynamicalFrictionForceAmplitude (:82-99)
      const double ratio = gh_fdm_ratio(T, W, x * x + y * y + z * z, v2);
      if (ratio < 1.0 || W->p[11] >= 0.0) A *= ratio;
    }
    ax += A * vx;
    ay += A * vy;
    az += A * vz;
  }
}

// Force functor handed to the integrators: the table plus the compile-time
// composition.
template <class Spec>
struct GhForce {
  const GhPotTable* __restrict__ T;
  // velocity-dependent components present (only the extended instances
  // carry them)
  __device__ __forceinline__ bool diss() const {
    return std::is_same<Spec, GhSpecWrapped>::value && T->ndiss != 0;
  }
  // evalRectDeriv's acceleration (integrateFullOrbit.c:1326-1351): the
  // conservative force plus the velocity-dependent ones
  __device__ __forceinline__ void rect_v(double t, double x, double y, double z, double vx,
                                         double vy, double vz, double& ax, double& ay,
                                         double& az) const {
    gh_rect_force<Spec>(T, t, x, y, z, ax, ay, az);
    if (diss()) gh_diss_rect(T, t, x, y, z, vx, vy, vz, ax, ay, az);
  }
  __device__ __forceinline__ void rect_v(double t, double x, double y, double vx, double vy,
                                         double& ax, double& ay) const {
    rect(t, x, y, ax, ay);
    if (diss()) {
      double az = 0.0;
      gh_diss_rect(T, t, x, y, 0.0, vx, vy, 0.0, ax, ay, az);
    }
  }
  __device__ __forceinline__ void rect(double t, double x, double y, double z, double& ax,
                                       double& ay, double& az) const {
    gh_rect_force<Spec>(T, t, x, y, z, ax, ay, az);
  }
  // Planar orbits (evalPlanarRectForce, integratePlanarOrbit.c:1015-1031).
  // galpy's planar kernels (e.g. NFWPotentialPlanarRforce, NFWPotential.c:28-37;
  // DehnenBarPotentialPlanarRforce, DehnenBarPotential.c:41-59) are the full
  // forces at z = 0, so the full kernels run with z = 0 and their z terms
  // fold away.
  __device__ __forceinline__ void rect(double t, double x, double y, double& ax,
                                       double& ay) const {
    const GhAcc F = gh_acc_forces<Spec>(T, x, y, 0.0, t);
    ax = x * F.FRoR - y * F.FphioR2;
    ay = y * F.FRoR + x * F.FphioR2;
  }
};

// EllipsoidalPotentialEval (EllipsoidalPotential.c:6-39): -amp/2 sum_k w_k
// psi(m_k) with galpy's packed weights; psi of each member
// (TwoPowerTriaxialPotential.c:7-89, PerfectEllipsoid.c:4-8,
// TriaxialGaussian.c:4-7, PowerTriaxial.c:4-8).  TwoPowerTriaxial's psi needs
// the Gauss hypergeometric function and has no device form here (NaN).
__device__ inline double gh_ellip_potential(const GhPotTable* __restrict__ T,
                                            const GhComponent* __restrict__ C, double x, double y,
                                            double z) {
  const double* __restrict__ e = &T->ext[C->ext];
  double xe = x, ye = y, ze = z;
  if (C->c[2] == 0.0) {
    xe = e[0] * x + e[1] * y + e[2] * z;
    ye = e[3] * x + e[4] * y + e[5] * z;
    ze = e[6] * x + e[7] * y + e[8] * z;
  }
  const double* __restrict__ nd = e + 9;
  const int n = (int)C->c[3];
  double out = 0.0;
  for (int k = 0; k < n; k++, nd += 4) {
    const double m2 = xe * xe * nd[0] + ye * ye * nd[1] + ze * ze * nd[2];
    double psi;
    switch (C->type) {
      case GH_TRIAXHERNQUIST: {
        const double mpa = sqrt(m2) + C->p[1];
        psi = -C->p[2] / (mpa * mpa);
        break;
      }
      case GH_TRIAXNFW: psi = -2.0 * C->p[2] / (sqrt(m2) + C->p[1]); break;
      case GH_TRIAXJAFFE: {
        const double m = sqrt(m2), mpa = m + C->p[1];
        psi = 2.0 * C->p[2] * (C->p[1] / mpa + log(m / mpa));
        break;
      }
      case GH_PERFECTELLIPSOID: psi = -1.0 / (C->p[1] + m2); break;
      case GH_TRIAXGAUSSIAN: psi = C->p[1] * exp(m2 / C->p[1]); break;
      case GH_POWERTRIAX: psi = 2.0 / (2.0 - C->p[1]) * pow(m2, 1.0 - 0.5 * C->p[1]); break;
      default: psi = NAN; break;
    }
    out += nd[3] * psi;
  }
  return -0.5 * C->p[0] * out;
}

// Potential value of component i alone at (R, z, phi, t), cosphi/sinphi =
// cos/sin phi (per-component Eval functions; DehnenBar from
// DehnenBarPotential.py:173-200, which has no C Eval).
__device__ __forceinline__ double gh_comp_potential(const GhPotTable* __restrict__ T, int i,
                                                    double R, double z, double cosphi,
                                                    double sinphi, double t) {
  const GhComponent* __restrict__ C = &T->comp[i];
  const double R2 = R * R;
  const double z2 = z * z;
  const double r2 = R2 + z2;
  double pot = 0.0;
  switch (C->type) {
    case GH_POWERSPHCUTOFF: {
      // PowerSphericalPotentialwCutoff.c:12-24
      const double r = sqrt(r2);
      const double x = r2 * C->c[3];
      const double g2 = gh_lower_incgamma(C->c[5], x, C->c[6], C->c[7], C->c[10], T->ser[i][1]);
      const double g1 = gh_lower_incgamma(C->c[0], x, C->c[1], C->c[4], C->c[9], T->ser[i][0]);
      pot += C->c[2] / r * (r * C->c[8] * g2 - g1);
      break;
    }
    case GH_MIYAMOTONAGAI: {
      const double as = C->p[1] + sqrt(z2 + C->c[0]);
      pot += -C->p[0] / sqrt(R2 + as * as);
      break;
    }
    case GH_NFW: {
      const double r = sqrt(r2);
      pot += -C->p[0] * gh_log_ge1(1.0 + r * C->c[0]) / r;
      break;
    }
    case GH_HERNQUIST: {
      const double r = sqrt(r2);
      pot += -C->p[0] / (1.0 + r / C->p[1]) / 2.0 / C->p[1];
      break;
    }
    case GH_LOGHALO: {
      const double amp = C->p[0], core2 = C->p[2], onem1overb2 = C->p[3];
      const double zq2 = z2 * C->c[0];
      const double Rt2 = (onem1overb2 < 1.0) ? R2 * (1.0 - onem1overb2 * sinphi * sinphi) : R2;
      pot += 0.5 * amp * log(Rt2 + zq2 + core2);
      break;
    }
    case GH_FLATPOWER: {  // FlattenedPowerPotential.c:3-18
      const double alpha = C->p[1], m2 = C->p[3] + R2 + z2 * C->c[0];
      pot += (alpha == 0.0) ? 0.5 * C->p[0] * log(m2)
                            : -C->p[0] * pow(m2, -0.5 * alpha) / alpha;
      break;
    }
    case GH_SOFTNEEDLEBAR: {  // SoftenedNeedleBarPotential.c:13-29
      double cp, sp, xb, yb;
      gh_needle_frame(C, R * cosphi, R * sinphi, t, cp, sp, xb, yb);
      const double a = C->p[1], bz = C->p[2] + sqrt(z2 + C->p[3]);
      const double second = yb * yb + bz * bz;
      const double Tp = sqrt((a + xb) * (a + xb) + second), Tm = sqrt((a - xb) * (a - xb) + second);
      pot += 0.5 * C->p[0] * log((xb - a + Tm) / (xb + a + Tp)) / a;
      break;
    }
    case GH_STEADYSPIRAL:  // SteadyLogSpiralPotential.py _evaluate (planar only)
      pot += C->p[0] * gh_ellipdisk_smooth(t, C->p[1], C->p[2]) * C->p[3] / C->p[4] *
             cos(gh_spiral_arg(C, R2, atan2(sinphi, cosphi), t));
      break;
    case GH_INTERPRZ:  // interpRZPotential.c:6-20 (Eval)
      pot += C->p[0] * gh_interprz_sum(T, C, gh_interprz_setup(T, C, R, z), 0);
      break;
    case GH_DISKSCF: {  // DiskSCFPotential.c:108-120 (Eval)
      double S, dS, H, dH;
      gh_dscf_sigma(C, sqrt(r2), S, dS);
      gh_dscf_hz(C, z, H, dH);
      pot += S * H;
      break;
    }
    case GH_MULTIPOLE: {  // MultipoleExpansionPotential.c:719-829 (Eval)
      pot += gh_multipole_sums(T, C, R, z, cosphi, sinphi, t, true).V;
      break;
    }
    case GH_SCF: {  // SCFPotential.c:828-863 (Eval)
      double S[3], V;
      gh_scf_sums<true>(T, C, R, z, cosphi, sinphi, t, S, V);
      pot += V;
      break;
    }
    case GH_INTERPSPH: {  // interpSphericalPotential.c:5-20 (revaluate)
      const double r = sqrt(r2);
      const double v = (r >= C->p[2]) ? -C->p[3] / r + C->p[5]
                       : (r < C->p[1]) ? 0.0 : -gh_spline_integ(T, C->ext, r) + C->p[4];
      pot += C->p[0] * v;
      break;
    }
    case GH_DEDISK:  // DoubleExponentialDiskPotential.c:6-34 (Eval)
      pot += C->p[1] * gh_dedisk_sum<2>(T, C, 1.0 / R, fabs(z)) / R;
      break;
    case GH_SPIRALARMS:  // SpiralArmsPotential.c:21-58 (Eval)
      pot += gh_spiral_pre(C, R) *
             gh_spiral_sums<true>(T, C, R, z, atan2(sinphi, cosphi), t).pot;
      break;
    case GH_TRANSIENTSPIRAL:  // TransientLogSpiralPotential.py _evaluate (planar only)
      pot += gh_transient_amp(C, t) / C->p[4] * cos(gh_spiral_arg(C, R2, atan2(sinphi, cosphi), t));
      break;
    case GH_COSMPHIDISK: {  // CosmphiDiskPotential.py _evaluate (planar only)
      const double cm = cos(C->p[4] * atan2(sinphi, cosphi) - C->p[3]);
      const double Rp = pow(R, C->p[2]);
      pot += C->p[0] * C->c[2] * cm * ((R < C->p[5]) ? C->p[6] * (2.0 * C->c[0] - C->p[6] / Rp) : Rp);
      break;
    }
    case GH_HENONHEILES:  // HenonHeilesPotential.py _evaluate (planar only)
      pot += C->p[0] * 0.5 * R2 * (1.0 + 2.0 / 3.0 * R * (3.0 * cosphi * cosphi - sinphi * sinphi) * sinphi);
      break;
    case GH_LOPSIDED:  // CosmphiDiskPotential.py _evaluate at m = 1 (planar only)
      pot += C->p[0] * C->p[1] * pow(R2, 0.5 * C->p[2]) * (cosphi * C->c[0] + sinphi * C->c[1]);
      break;
    case GH_ELLIPDISK: {  // EllipticalDiskPotential.py _evaluate (planar only)
      const double c2 = cosphi * cosphi - sinphi * sinphi, s2 = 2.0 * sinphi * cosphi;
      pot += C->p[0] * gh_ellipdisk_smooth(t, C->p[1], C->p[2]) * C->p[3] * 0.5 *
             pow(R2, 0.5 * C->p[4]) * (c2 * C->c[0] + s2 * C->c[1]);
      break;
    }
    case GH_KKSTAECKEL: {  // KuzminKutuzovStaeckelPotential.c:5-23 (n clamped at 0)
      const double dm = r2 - C->c[2];
      const double s = sqrt(dm * dm + 4.0 * C->c[2] * R2), term = r2 - C->c[1] - C->c[0];
      const double n = 0.5 * (term - s);
      pot += -C->p[0] / (sqrt(0.5 * (term + s)) + sqrt(n > 0.0 ? n : 0.0));
      break;
    }
    case GH_KUZMINDISK: {  // KuzminDiskPotential.c:5-15
      const double az = C->p[1] + fabs(z);
      pot += -C->p[0] / sqrt(R2 + az * az);
      break;
    }
    case GH_POWERSPH: {  // PowerSphericalPotential.c Eval
      const double alpha = C->p[1];
      pot += (alpha == 2.0) ? 0.5 * C->p[0] * log(r2)
                            : -C->p[0] * pow(r2, 1.0 - 0.5 * alpha) / (alpha - 2.0);
      break;
    }
    case GH_JAFFE: {
      const double r = sqrt(r2);
      pot += -C->p[0] * log(1.0 + C->p[1] / r) / C->p[1];
      break;
    }
    case GH_ISOCHRONE: pot += -C->p[0] / (C->p[1] + sqrt(r2 + C->c[0])); break;
    case GH_PLUMMER: pot += -C->p[0] / sqrt(r2 + C->c[0]); break;
    case GH_PSEUDOISO: {
      const double a = C->p[1], r = sqrt(r2);
      pot += C->p[0] * (0.5 * log(1.0 + r2 / (a * a)) + a / r * atan(r / a)) / a;
      break;
    }
    case GH_BURKERT: {
      const double a = C->p[1], x = sqrt(r2) / a;
      pot += -C->p[0] * a * a * M_PI / x *
             (-M_PI + 2.0 * (1.0 + x) * atan(1.0 / x) + 2.0 * (1.0 + x) * log1p(x) +
              (1.0 - x) * log1p(x * x));
      break;
    }
    case GH_DEHNENCORESPH: {
      const double a = C->p[1], r = sqrt(r2), u = r / (r + a);
      pot += -C->p[0] * (1.0 - u * u) / (6.0 * a);
      break;
    }
    case GH_DEHNENSPH: {
      const double a = C->p[1], alpha = C->p[2], r = sqrt(r2);
      if (alpha == 2.0)  // the Jaffe limit of the closed form (0/0 at alpha = 2)
        pot += -C->p[0] * log1p(a / r) / a;
      else
        pot += -C->p[0] * (1.0 - pow(r / (r + a), 2.0 - alpha)) / (a * (2.0 - alpha) * (3.0 - alpha));
      break;
    }
    case GH_HOMOGSPHERE: {
      const double R2s = C->p[1], R3s = C->p[2];
      pot += (r2 < R2s) ? C->p[0] * (r2 - 3.0 * R2s) : -2.0 * C->p[0] * R3s / sqrt(r2);
      break;
    }
    case GH_DEHNENBAR: {
      const double amp = C->p[0], rb = C->p[3];
      const double smooth = gh_bar_smooth(t, C->p[1], C->p[2]);
      const double cos2p = cosphi * cosphi - sinphi * sinphi;
      const double sin2p = 2.0 * sinphi * cosphi;
      double cang, sang;
      gh_bar_angle(cos2p, sin2p, C->p[4], C->p[5], t, cang, sang);
      const double r = sqrt(r2);
      double out;
      if (r <= rb) {
        const double u = r / rb;
        out = (u * u * u - 2.0) * R2 / r2;
      } else {
        const double u = rb / r;
        out = -u * u * u * R2 / r2;
      }
      pot += out * amp * smooth * cang;
      break;
    }
    case GH_TRIAXHERNQUIST:
    case GH_TRIAXNFW:
    case GH_TRIAXJAFFE:
    case GH_TWOPOWERTRIAX:
    case GH_PERFECTELLIPSOID:
    case GH_TRIAXGAUSSIAN:
    case GH_POWERTRIAX: pot += gh_ellip_potential(T, C, R * cosphi, R * sinphi, z); break;
    case GH_EINASTO: {  // EinastoPotential.c revaluate: -amp 4 pi h^2 n (gamma(3n,x)/s + Gamma(2n,x))
      const double r = sqrt(r2), x = gh_einasto_x(C, r);
      const double g3 = gh_lower_incgamma_gen(C->c[0], x, C->c[1], C->c[2]);
      const double G2 = C->c[4] - gh_lower_incgamma_gen(C->c[3], x, C->c[4], C->c[5]);
      const double h = C->p[1];
      pot += -C->p[0] * C->c[6] / h * (g3 / (r * C->c[7]) + G2);
      break;
    }
    case GH_EXPTRUNCNFW: {  // ExpTruncNFWPotential.c revaluate: -amp (F/r + G)
      const double r = sqrt(r2);
      double ex, e1b;
      const double Fm = gh_exptruncnfw_F(C, r, ex, e1b);
      if (r < C->c[3]) e1b = gh_e1_scaled((C->p[1] + r) * C->c[2]);
      pot += -C->p[0] * (Fm / r + ex / (C->p[1] + r) - ex * e1b * C->c[2]);
      break;
    }
    case GH_TWOPOWERSPH: pot += NAN; break;  // its Eval needs 2F1 outside the family above
    default:
      break;
  }
  return pot;
}

// Potential of a component whose wrappers move the evaluation point (the
// same walk as gh_comp_force_mapped): the leaf's value at the moved point(s)
// times the wrappers' factors; CylindricallySeparable adds its -amp Phi(Rp, 0).
__device__ inline double gh_mapped_potential(const GhPotTable* __restrict__ T, int i, double x0,
                                             double y0, double z0, double t0) {
  const GhComponent* __restrict__ C = &T->comp[i];
  double total = 0.0;
  for (int pass = 0; pass < C->mapwrap; pass++) {
    double x = x0, y = y0, z = z0, t = t0, fac = 1.0, coef = 1.0, add = 0.0;
    double rec[GH_WREC];
    for (int k = 0; k < C->nwrap; k++) {
      const GhWrap* __restrict__ W = &T->wrap[C->wrap[k]];
      const bool os = W->kind == GH_OBLATESTAECKELWRAP;
      gh_wrap_forward(T, W, pass, x, y, z, t, fac, os ? rec : nullptr);
      // the wrappers' constants, once per wrapper (at its first component)
      if (W->kind == GH_CYLSEPWRAP && pass == 1 && W->first == i) total -= fac * W->p[2];
      if (os) {  // innermost: Phi = amp (ch^2 Phi_A + P0 Phi_B - refpot) / pre
        coef = rec[10];
        add = (W->first == i) ? rec[11] : 0.0;
      }
    }
    const double R = sqrt(x * x + y * y);
    const double cp = (R > 0.0) ? x / R : 1.0, sp = (R > 0.0) ? y / R : 0.0;
    total += fac * (coef * gh_comp_potential(T, i, R, z, cp, sp, t) + add);
  }
  return total;
}

// Potential value: the sum over components; wrapped components follow the
// wrappers' Python _evaluate (the component at phi - offset times the factor,
// or at the point the wrappers move it to).
__device__ __forceinline__ double gh_potential(const GhPotTable* __restrict__ T, double R,
                                               double z, double cosphi0, double sinphi0,
                                               double t) {
  double total = 0.0;
  for (int i = 0; i < T->npot; i++) {
    const GhComponent* __restrict__ C = &T->comp[i];
    if (C->densonly) continue;
    if (C->mapwrap) {
      total += gh_mapped_potential(T, i, R * cosphi0, R * sinphi0, z, t);
      continue;
    }
    double fac = 1.0, cosphi = cosphi0, sinphi = sinphi0;
    if (C->nwrap) {
      double off, doff;
      gh_wrap_transform(T, C, R, t, fac, off, doff);
      const GhSinCos sc = gh_sincos(off);
      const double so = sc.s, co = sc.c;
      cosphi = cosphi0 * co + sinphi0 * so;
      sinphi = sinphi0 * co - cosphi0 * so;
    }
    total += fac * gh_comp_potential(T, i, R, z, cosphi, sinphi, t);
  }
  return total;
}
