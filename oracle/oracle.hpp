// CPU oracle for galpy's C orbit-integration path -- TEST INFRASTRUCTURE ONLY.
//
// This header restates, in plain C++ (templated on the scalar type), the
// algorithm galpy runs inside `integrateFullOrbit`: the six potential force
// kernels used by the benchmark configurations, the cylindrical aggregators,
// the rectangular force callbacks, the coordinate transforms, leapfrog,
// symplec4 (+ their step estimators) and DOP853 with dense output -- and the
// planar twins of the forces, callbacks and transforms that
// `integratePlanarOrbit` uses.  Only the
// tests, the smoke check and the CPU-baseline leg of the benchmark may load
// the library built from it; the HIP product never calls it.
//
// Templating on the scalar type R lets the same code run
//   * with R = double   -> the oracle proper (and the CPU baseline), and
//   * with R = CountedD -> an op-counting twin that tallies FP64 operations,
//     which gives the algorithmic FLOP count per orbit-step.
//
// Every function cites the reference file:line it follows.  Operation order
// inside each formula follows the reference so that oracle and reference agree
// to rounding, except where the reference calls GSL (absent here): the
// incomplete gamma function is this file's own (series + continued fraction,
// written from the mathematics) and is pinned against mpmath.
#pragma once
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>
#include <cfloat>
#include <algorithm>

namespace orc {

using std::sqrt; using std::log; using std::exp; using std::pow;
using std::fabs; using std::sin; using std::cos; using std::acos;
using std::atan2; using std::atan; using std::tgamma; using std::cosh; using std::tanh;

inline double dval(double x) { return x; }

// ---------------------------------------------------------------------------
// Op-counting scalar.  Each IEEE add/sub/mul/div/sqrt is one counted op; each
// libm transcendental call (exp, log, pow, sin, cos, cosh, tanh, acos, atan2) is counted
// separately as a "special" call.  Comparisons, fabs, negation and selects are
// free.  Counters are thread-local so OpenMP loops can count in parallel.
// ---------------------------------------------------------------------------
struct OpCount {
  unsigned long long add, mul, div, sqrt, special;
};
extern thread_local OpCount g_ops;

struct CountedD {
  double v;
  CountedD() : v(0.) {}
  CountedD(double x) : v(x) {}
};
inline double dval(CountedD x) { return x.v; }
inline CountedD operator+(CountedD a, CountedD b) { g_ops.add++; return CountedD(a.v + b.v); }
inline CountedD operator-(CountedD a, CountedD b) { g_ops.add++; return CountedD(a.v - b.v); }
inline CountedD operator*(CountedD a, CountedD b) { g_ops.mul++; return CountedD(a.v * b.v); }
inline CountedD operator/(CountedD a, CountedD b) { g_ops.div++; return CountedD(a.v / b.v); }
inline CountedD operator-(CountedD a) { return CountedD(-a.v); }
inline CountedD& operator+=(CountedD& a, CountedD b) { a = a + b; return a; }
inline CountedD& operator-=(CountedD& a, CountedD b) { a = a - b; return a; }
inline CountedD& operator*=(CountedD& a, CountedD b) { a = a * b; return a; }
inline CountedD& operator/=(CountedD& a, CountedD b) { a = a / b; return a; }
inline bool operator<(CountedD a, CountedD b) { return a.v < b.v; }
inline bool operator>(CountedD a, CountedD b) { return a.v > b.v; }
inline bool operator<=(CountedD a, CountedD b) { return a.v <= b.v; }
inline bool operator>=(CountedD a, CountedD b) { return a.v >= b.v; }
inline bool operator==(CountedD a, CountedD b) { return a.v == b.v; }
inline bool operator!=(CountedD a, CountedD b) { return a.v != b.v; }
inline CountedD sqrt(CountedD a) { g_ops.sqrt++; return CountedD(std::sqrt(a.v)); }
inline CountedD fabs(CountedD a) { return CountedD(std::fabs(a.v)); }
inline CountedD log(CountedD a) { g_ops.special++; return CountedD(std::log(a.v)); }
inline CountedD exp(CountedD a) { g_ops.special++; return CountedD(std::exp(a.v)); }
inline CountedD pow(CountedD a, CountedD b) { g_ops.special++; return CountedD(std::pow(a.v, b.v)); }
inline CountedD sin(CountedD a) { g_ops.special++; return CountedD(std::sin(a.v)); }
inline CountedD cos(CountedD a) { g_ops.special++; return CountedD(std::cos(a.v)); }
inline CountedD acos(CountedD a) { g_ops.special++; return CountedD(std::acos(a.v)); }
inline CountedD atan2(CountedD a, CountedD b) { g_ops.special++; return CountedD(std::atan2(a.v, b.v)); }
inline CountedD atan(CountedD a) { g_ops.special++; return CountedD(std::atan(a.v)); }
inline CountedD cosh(CountedD a) { g_ops.special++; return CountedD(std::cosh(a.v)); }
inline CountedD tanh(CountedD a) { g_ops.special++; return CountedD(std::tanh(a.v)); }

template <class R> inline R rmax(R a, R b) { return (a > b) ? a : b; }
template <class R> inline R rmin(R a, R b) { return (a < b) ? a : b; }

// ---------------------------------------------------------------------------
// Incomplete gamma function (replaces GSL's gsl_sf_gamma_inc_P and
// gsl_sf_gamma_inc at PowerSphericalPotentialwCutoff.c:10,23).
//   lower  gamma(a,x) = int_0^x t^(a-1) e^-t dt
//   series (x < a+1): gamma(a,x) = e^-x x^a sum_{n>=0} x^n / (a (a+1)...(a+n))
//   continued fraction (x >= a+1), modified Lentz:
//     Gamma(a,x) = e^-x x^a / (x+1-a - 1(1-a)/(x+3-a - 2(2-a)/(x+5-a - ...)))
// (Abramowitz & Stegun 6.5.29 and 6.5.31; Numerical Recipes 2nd ed. 6.2.)
// Pinned against mpmath at 50 digits (tests/golden/incgamma.json).
// ---------------------------------------------------------------------------
constexpr double kIGEps = DBL_EPSILON / 2.;
constexpr double kIGTiny = 1e-300;

template <class R> R incgamma_series_sum(double a, R x) {
  R ap = a;
  R del = R(1.) / ap;
  R sum = del;
  for (int n = 1; n < 2000; n++) {
    ap += 1.;
    del *= x / ap;
    sum += del;
    if (fabs(del) < fabs(sum) * kIGEps) break;
  }
  return sum;
}

template <class R> R incgamma_cf(double a, R x) {
  R b = x + 1. - a;
  R c = R(1. / kIGTiny);
  R d = R(1.) / b;
  R h = d;
  for (int i = 1; i < 2000; i++) {
    double an = -i * (i - a);
    b += 2.;
    d = R(an) * d + b;
    if (fabs(d) < R(kIGTiny)) d = kIGTiny;
    c = b + R(an) / c;
    if (fabs(c) < R(kIGTiny)) c = kIGTiny;
    d = R(1.) / d;
    R del = d * c;
    h *= del;
    if (fabs(del - 1.) < R(kIGEps)) break;
  }
  return h;
}

// lower incomplete gamma gamma(a,x); gamma_a = Gamma(a)
template <class R> R lower_incgamma(double a, R x, double gamma_a) {
  if (x <= R(0.)) return R(0.);
  R pref = exp(R(a) * log(x) - x);
  if (x < R(a + 1.)) return pref * incgamma_series_sum(a, x);
  return R(gamma_a) - pref * incgamma_cf(a, x);
}
// upper incomplete gamma Gamma(a,x)
template <class R> R upper_incgamma(double a, R x, double gamma_a) {
  if (x <= R(0.)) return R(gamma_a);
  R pref = exp(R(a) * log(x) - x);
  if (x < R(a + 1.)) return R(gamma_a) - pref * incgamma_series_sum(a, x);
  return pref * incgamma_cf(a, x);
}

// ---------------------------------------------------------------------------
// Exponential integral E1(x), x > 0 (replaces GSL's gsl_sf_expint_E1 at
// ExpTruncNFWPotential.c:25,32), from A&S 5.1.11 and 5.1.22:
//   x <= 1: E1 = -gamma_E - ln x - sum_{k>=1} (-x)^k / (k k!)
//   x > 1 : E1 = e^-x / (x + 1 - 1^2/(x + 3 - 2^2/(x + 5 - ...))), modified Lentz.
// e1_scaled returns e^x E1(x).  Pinned against mpmath.
// ---------------------------------------------------------------------------
constexpr double kEulerGamma = 0.57721566490153286060651209008240243;
template <class R> R e1_scaled(R x) {
  if (x <= R(1.)) {
    R sum = 0., term = 1.;
    for (int k = 1; k < 100; k++) {
      term = term * (-x) / R(k);
      R add = term / R(k);
      sum += add;
      if (fabs(add) < fabs(sum) * R(kIGEps)) break;
    }
    return exp(x) * (R(-kEulerGamma) - log(x) - sum);
  }
  R b = x + R(1.);
  R c = R(1. / kIGTiny);
  R d = R(1.) / b;
  R h = d;
  for (int i = 1; i < 1000; i++) {
    R an = R(-(double)i * i);
    b += R(2.);
    d = R(1.) / (an * d + b);
    c = b + an / c;
    R del = c * d;
    h *= del;
    if (fabs(del - R(1.)) < R(kIGEps)) break;
  }
  return h;
}
template <class R> R expint_e1(R x) { return exp(-x) * e1_scaled(x); }

// ---------------------------------------------------------------------------
// Gauss hypergeometric 2F1(A, B; A+1; -x) for x >= 0, the family galpy's
// TwoPowerSphericalPotential.c calls (the reference evaluates it with the
// vendored xsf library's hyp2f1, wrap_xsf.cpp).  Written from the
// mathematics: 2F1(A, B; A+1; -x) = A x^-A I(x), I(x) = int_0^x u^(A-1)
// (1+u)^-B du, and
//   x <= 3: Pfaff's transformation, 2F1 = (1+x)^-B sum_n (B)_n/(A+1)_n w^n,
//           w = x/(1+x) <= 3/4 (positive terms for B > 0);
//   x > 3 : I(x) = I(3) + int_3^x u^(A-B-1) (1+1/u)^-B du with the binomial
//           series of (1+1/u)^-B integrated term by term (a logarithm for
//           the term whose power is u^-1, when 3 - beta = A - B is a
//           nonnegative integer).
// Pinned against mpmath.hyp2f1.
// ---------------------------------------------------------------------------
template <class R> R hyp2f1_ap1_pfaff(double A, double B, R x) {
  R w = x / (R(1.) + x);
  R sum = 1., term = 1.;
  for (int n = 0; n < 5000; n++) {
    term = term * R((B + n) / (A + 1. + n)) * w;
    sum += term;
    if (fabs(term) < fabs(sum) * R(kIGEps)) break;
  }
  return pow(R(1.) + x, R(-B)) * sum;
}
// amb = A - B as the caller knows it exactly (3 - beta)
template <class R> R hyp2f1_ap1(double A, double B, double amb, R x) {
  const double x0 = 3.;
  if (x <= R(x0)) return hyp2f1_ap1_pfaff(A, B, x);
  R I = R(pow(x0, A) / A) * hyp2f1_ap1_pfaff(A, B, R(x0));
  R ck = 1.;  // (-1)^k (B)_k / k!
  R lx = log(x / R(x0));
  for (int k = 0; k < 2000; k++) {
    if (k > 0) ck = ck * R(-(B + k - 1.) / k);
    double e = amb - k;
    R add = (e == 0.) ? ck * lx : ck * (pow(x, R(e)) - R(pow(x0, e))) / R(e);
    I += add;
    if (k > 2 && fabs(add) < fabs(I) * R(kIGEps) && fabs(ck) * pow(R(x0), R(e)) < fabs(I) * R(kIGEps))
      break;
  }
  return R(A) * pow(x, R(-A)) * I;
}

// ---------------------------------------------------------------------------
// Potential components.  Parameters are decoded once from galpy's packed
// (pot_type, pot_args) arrays (integrateFullOrbit.c:69-1045, cases 0,1,5,8,9,15)
// and kept read-only; constants that the reference recomputes on every call
// (Gamma(a), 2 pi rc^(3-alpha)) are precomputed with the same operations.
// ---------------------------------------------------------------------------
enum PotCode { kLogHalo = 0, kDehnenBar = 1, kMiyamotoNagai = 5, kPowerSph = 7, kHernquist = 8,
               kNFW = 9, kJaffe = 10, kIsochrone = 14, kPowerSphCutoff = 15, kPlummer = 17,
               kPseudoIso = 18, kKuzminDisk = 19, kBurkert = 20, kDehnenCoreSph = 33,
               kDehnenSph = 34, kHomogSphere = 35, kFlattenedPower = 12, kNull = 40,
               kKKStaeckel = 16, kSoftNeedleBar = 25, kEllipDisk = 4, kLopsided = 6, kSteadySpiral = 3,  // 3, 4, 6: planar only
               kTransientSpiral = 2, kCosmphiDisk = 28, kHenonHeiles = 29,  // planar only
               kSpiralArms = 27, kDoubleExpDisk = 11,
               // wrappers (integrateFullOrbit.c:821-895)
               kDehnenSmoothWrap = -1, kSolidBodyRotWrap = -2, kCorotatingRotWrap = -4,
               kGaussianAmpWrap = -5, kRotateTiltWrap = -8, kKuzminLikeWrap = -10, kCylSepWrap = -12,
               kOblateStaeckelWrap = -3, kNonInertial = 39, kNonInertialInterp = 45,
               kChandrasekhar = -7, kFDM = -11,
               kTimeDepAmpWrap = -9, kMovingObjectWrap = -6, kInterpSph = 36, kSCF = 24,
               kDiskSCF = 26, kInterpRZ = 13, kMultipole = 44,
               // EllipsoidalPotential family (integrateFullOrbit.c:420-483, 549, 647, 668, 764)
               kTriaxHernquist = 21, kTriaxNFW = 22, kTriaxJaffe = 23, kPerfectEllipsoid = 30,
               kTriaxGaussian = 37, kPowerTriax = 38, kTwoPowerTriax = 43,
               // spherical components with special functions (integrateFullOrbit.c:731-780)
               kEinasto = 41, kTwoPowerSph = 42, kExpTruncNFW = 46 };

inline bool is_ellipsoidal(int code) {
  return code == kTriaxHernquist || code == kTriaxNFW || code == kTriaxJaffe ||
         code == kPerfectEllipsoid || code == kTriaxGaussian || code == kPowerTriax ||
         code == kTwoPowerTriax;
}

// Components of one parse, wrapped ones included (the C-ABI entry points
// accept up to this many).
constexpr int kOracleMaxComp = 16;

// GSL's natural cubic spline (gsl_interp_cspline, the interpolation galpy's C
// builds with gsl_spline_init; GSL is absent here), restated from its
// published definition: with c_i = y''(x_i)/2, c_0 = c_{n-1} = 0 and, for
// 0 < i < n-1, h_{i-1} c_{i-1} + 2 (h_{i-1} + h_i) c_i + h_i c_{i+1} =
// 3 ((y_{i+1} - y_i)/h_i - (y_i - y_{i-1})/h_{i-1}) (solved by Thomas'
// algorithm); on [x_i, x_{i+1}], with d = x - x_i,
//   y = y_i + d (b_i + d (c_i + d d_i)), b_i = (y_{i+1} - y_i)/h_i -
//   h_i (c_{i+1} + 2 c_i)/3, d_i = (c_{i+1} - c_i)/(3 h_i);
// integrals term by term.  Pinned against scipy's CubicSpline(bc_type=
// 'natural').
struct NatSpline {
  std::vector<double> x, y, b, c, d;
  void build(const double* xs, const double* ys, int n) {
    x.assign(xs, xs + n);
    y.assign(ys, ys + n);
    c.assign(n, 0.);
    b.assign(n, 0.);
    d.assign(n, 0.);
    if (n >= 3) {
      const int m = n - 2;
      std::vector<double> diag(m), off(m), rhs(m);
      for (int i = 1; i <= m; i++) {
        const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
        diag[i - 1] = 2. * (h0 + h1);
        off[i - 1] = h1;
        rhs[i - 1] = 3. * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
      }
      for (int i = 1; i < m; i++) {  // forward elimination
        const double w = off[i - 1] / diag[i - 1];
        diag[i] -= w * off[i - 1];
        rhs[i] -= w * rhs[i - 1];
      }
      c[m] = rhs[m - 1] / diag[m - 1];
      for (int i = m - 2; i >= 0; i--) c[i + 1] = (rhs[i] - off[i] * c[i + 2]) / diag[i];
    }
    for (int i = 0; i + 1 < n; i++) {
      const double h = x[i + 1] - x[i];
      b[i] = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2. * c[i]) / 3.;
      d[i] = (c[i + 1] - c[i]) / (3. * h);
    }
  }
  // interval of t: the largest i with x[i] <= t, within [0, n-2]
  int find(double t) const {
    int lo = 0, hi = (int)x.size() - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (x[mid] > t) hi = mid; else lo = mid;
    }
    return lo;
  }
  template <class R> R eval(R t) const {
    const int i = find(dval(t));
    R dx = t - R(x[i]);
    return R(y[i]) + dx * (R(b[i]) + dx * (R(c[i]) + dx * R(d[i])));
  }
  // derivative (gsl_spline_eval_deriv: b_i + d (2 c_i + 3 d_i d))
  double deriv(double t) const {
    const int i = find(t);
    const double dx = t - x[i];
    return b[i] + dx * (2.0 * c[i] + 3.0 * d[i] * dx);
  }
  // integral over [x[0], t]
  template <class R> R integ(R t) const {
    const int k = find(dval(t));
    R sum = 0.;
    for (int i = 0; i <= k; i++) {
      R u = (i < k) ? R(x[i + 1] - x[i]) : t - R(x[i]);
      sum += u * (R(y[i]) + u * (R(b[i]) / R(2.) + u * (R(c[i]) / R(3.) + u * R(d[i]) / R(4.))));
    }
    return sum;
  }
};

struct Component {
  int type;
  double p[16];  // raw arguments in galpy's order
  double c[6];  // precomputed constants
  // a wrapper's wrapped components (potentialArg.nwrapped /
  // wrappedPotentialArg, galpy_potentials.h), stored in the same parse array
  int nwrapped = 0;
  const Component* wrapped = nullptr;
  // a component with a variable number of arguments reads them where they
  // lie in pot_args, as galpy's potentialArg.args does (valid for the call)
  const double* args = nullptr;
  // a wrapper's time function (potentialArg.tfuncs, galpy_potentials.h)
  double (*tfunc)(double) = nullptr;
  // a component with several time functions: the first's index in g_tfuncs
  int tf0 = 0;
  // cubic splines a component interpolates (potentialArg.spline1d): built
  // at parse time, shared by copies of the component
  std::shared_ptr<std::vector<NatSpline>> spl;
};

// The caller's time functions (galpy's pot_tfuncs), consumed in parse order
// by the components that take them (integrateFullOrbit.c:1034-1038):
// wrapped components first, then the wrapper.  Set before a parse.
inline double (**g_tfuncs)(double) = nullptr;
inline int g_ntfuncs = 0;

inline bool is_wrapper(int code) {
  return code == kDehnenSmoothWrap || code == kSolidBodyRotWrap || code == kCorotatingRotWrap ||
         code == kGaussianAmpWrap || code == kRotateTiltWrap || code == kKuzminLikeWrap ||
         code == kCylSepWrap || code == kTimeDepAmpWrap || code == kMovingObjectWrap ||
         code == kOblateStaeckelWrap;
}

// nargs per code, as consumed from pot_args (integrateFullOrbit.c:80-345;
// planar packing integratePlanarOrbit.c:64-223, where
// PowerSphericalPotentialwCutoff has 3 arguments and no cache slots)
inline int component_nargs(int code, bool planar = false) {
  switch (code) {
    case kLogHalo: return 4;
    case kDehnenBar: return 6;
    case kMiyamotoNagai: return 3;
    case kHernquist: return 2;
    case kNFW: return 2;
    case kPowerSphCutoff: return planar ? 3 : 5;
    case kPowerSph: case kJaffe: case kIsochrone: case kPlummer: case kPseudoIso:
    case kKuzminDisk: case kBurkert: case kDehnenCoreSph: return 2;
    case kFlattenedPower: return planar ? 3 : 4;  // planar: amp, alpha, core^2
    case kNull: return 0;
    case kKKStaeckel: return 3;  // amp, ac, Delta
    case kEllipDisk: return planar ? 6 : -1;  // planar only
    case kLopsided: return planar ? 4 : -1;   // planar only
    case kSteadySpiral: return planar ? 8 : -1;  // planar only
    case kTransientSpiral: return planar ? 8 : -1;  // amp, A, to, sigma^2, alpha, m, omegas, gamma
    case kCosmphiDisk: return planar ? 9 : -1;  // amp, mphio, p, mphib, m, rb, rbp, rb2p, r1p
    case kHenonHeiles: return planar ? 1 : -1;  // amp
    case kSoftNeedleBar: return 13;  // amp, a, b, c2, pa, omegab + 7 cache slots
    case kDehnenSph: case kHomogSphere: return 3;
    case kEinasto: case kExpTruncNFW: return 3;  // amp, h, n / amp, a, rc
    case kTwoPowerSph: return 4;                  // amp, a, alpha, beta
    case kNonInertial: return 23;  // amp, 10 cache slots, 5 flags, Omega[3], Omega^2, Omegadot[3]
    // wrappers' own arguments, read after the wrapped components
    case kDehnenSmoothWrap: return 4;   // amp, tform, tsteady, grow
    case kSolidBodyRotWrap: return 3;   // amp, omega, pa
    case kCorotatingRotWrap: return 5;  // amp, vpo, beta, pa, to
    case kGaussianAmpWrap: return 3;    // amp, to, sigma^2
    // amp, 7 cache slots, rot[9], rotSet, offsetSet, offset[3], 10 cache slots
    case kRotateTiltWrap: return planar ? -1 : 32;
    case kKuzminLikeWrap: return 3;     // amp, a, b^2
    case kCylSepWrap: return 3;         // amp, Rp, refpot
    case kOblateStaeckelWrap: return 5; // amp, delta, u0, v0, refpot
    case kTimeDepAmpWrap: return 1;     // amp (and one time function)
    default: return -1;
  }
}

// Arguments of an EllipsoidalPotential: amp, 15 cache slots, npsi, the psi
// arguments, b^2, c^2, aligned, rot[9], glorder, glx[glorder], glw[glorder]
// (integrateFullOrbit.py:221-300; nargs as integrateFullOrbit.c:437).
// SpiralArmsPotential: nCs, amp, N, sin_alpha, tan_alpha, r_ref, phi_ref,
// Rs, H, omega, Cs[nCs] (integrateFullOrbit.py:337-352; nargs as
// integrateFullOrbit.c:545).
inline int spiralarms_nargs(const double* a) { return 10 + (int)a[0]; }
// DoubleExponentialDiskPotential: amp, -4 pi alpha amp, alpha, beta, de_n,
// then the J0 nodes, J1 nodes, J0 weights, J1 weights, de_n each
// (integrateFullOrbit.py:93-107).
inline int dedisk_nargs(const double* a) { return 5 + 4 * (int)a[4]; }
// interpSphericalPotential: nr, rgrid[nr], rforce[nr], amp, rmin, rmax,
// M(<rmax), Phi0, Phimax (integrateFullOrbit.py:366-373)
inline int interpsph_nargs(const double* a) { return 1 + 2 * (int)a[0] + 6; }
// interpRZPotential: nR, nz, three flags for second-derivative grids, the R
// (or log R) grid, the z grid, the B-spline coefficients of the potential,
// R force and z force (and of each flagged second derivative), amp, logR
// (integrateFullOrbit.py:111-173)
inline int interprz_nargs(const double* a) {
  const int nR = (int)a[0], nz = (int)a[1];
  const int nextra = (int)a[2] + (int)a[3] + (int)a[4];
  return 5 + nR + nz + (3 + nextra) * nR * nz + 2;
}
// SCFPotential: a, isNonAxi, N, L, M, Nt, then Acos[N L M] (and Asin when
// non-axisymmetric), or tgrid[Nt] and the time-PPoly blocks [Nt-1][NLM][4]
// of Acos (and Asin), then 11 cache slots (integrateFullOrbit.c nargs of
// case 24; packing integratePlanarOrbit.py:726-762)
inline int scf_nargs(const double* a) {
  const int nonaxi = (int)a[1], NLM = (int)a[2] * (int)a[3] * (int)a[4], Nt = (int)a[5];
  return Nt == 0 ? 6 + (1 + nonaxi) * NLM + 11 : 6 + Nt + (1 + nonaxi) * NLM * 4 * (Nt - 1) + 11;
}
// MultipoleExpansionPotential: Nr, L, M, isNonAxi, rgrid[Nr], amp, Nt, then
// per (l, m < min(l + 1, M)) and trig (cos; sin when m > 0) the I_inner and
// I_outer quintic coefficients [Nr-1][6] and the density at the nodes [Nr]
// (static), or, after tgrid[Nt], their cubic-in-t blocks [Nt-1][6 (Nr-1)][4]
// and [Nt-1][Nr][4] (MultipoleExpansionPotential.c:183-399; packing
// MultipoleExpansionPotential.py:2101-2216)
inline int multipole_nblk(int L, int M) {
  int n = 0;
  for (int l = 0; l < L; l++) n += 2 * std::min(l + 1, M) - 1;
  return n;
}
inline int multipole_nargs(const double* a) {
  const long Nr = (long)a[0], L = (long)a[1], M = (long)a[2];
  const long Nt = (long)a[4 + Nr + 1], nr = 6 * (Nr - 1);
  const long per = Nt == 0 ? 2 * nr + Nr : 4 * (Nt - 1) * (2 * nr + Nr);
  return (int)(4 + Nr + 2 + Nt + multipole_nblk((int)L, (int)M) * per);
}
// MovingObjectPotential's own arguments: nPts, t[nPts], x, y[, z], amp, t0,
// tf (integrateFullOrbit.py:449-463; the planar packing has no z)
inline int movingobject_nargs(const double* a, bool planar) { return 1 + (planar ? 3 : 4) * (int)a[0] + 3; }

inline int ellipsoidal_nargs(const double* a) {
  const int npsi = (int)a[16];
  const int glorder = (int)a[17 + npsi + 12];
  return 30 + npsi + 2 * glorder;
}

// Decode n components into dst, consuming codes and arguments in galpy's
// order (parse_leapFuncArgs_Full, integrateFullOrbit.c:55-1045): a wrapper's
// code is followed by its wrapped components' codes; its arguments are
// nwrapped, the wrapped components' arguments, then its own
// (integrateFullOrbit.c:1010-1018, 1026-1031).  Wrapped components are
// stored in out[next...].
inline int parse_components(int n, const int*& pot_type, const double*& pot_args, Component* out,
                            int& next, Component* dst, bool planar, int& tf_next) {
  for (int i = 0; i < n; i++) {
    int code = *pot_type++;
    int na = is_ellipsoidal(code)    ? ellipsoidal_nargs(pot_args)
             : code == kSpiralArms ? spiralarms_nargs(pot_args)
             : code == kDoubleExpDisk ? dedisk_nargs(pot_args)
             : code == kInterpSph ? interpsph_nargs(pot_args)
             : code == kSCF ? scf_nargs(pot_args)
             : code == kDiskSCF ? (int)pot_args[0] + 3
             : code == kInterpRZ ? interprz_nargs(pot_args)
             : code == kMultipole ? multipole_nargs(pot_args)
             : code == kNonInertialInterp ? 0
             : code == kNonInertial ? 23
             : code == kMovingObjectWrap || code == kChandrasekhar || code == kFDM ? 0  // counted after its wrapped components
                                   : component_nargs(code, planar);
    if (na < 0) return -1;
    Component& C = dst[i];
    C.type = code;
    C.nwrapped = 0;
    C.wrapped = nullptr;
    std::memset(C.p, 0, sizeof(C.p));
    std::memset(C.c, 0, sizeof(C.c));
    if (code < 0) {
      int nw = (int)*pot_args++;
      if (nw < 1 || next + nw > kOracleMaxComp) return -1;
      Component* kids = out + next;
      next += nw;
      if (parse_components(nw, pot_type, pot_args, out, next, kids, planar, tf_next) != 0) return -1;
      C.nwrapped = nw;
      C.wrapped = kids;
    }
    C.args = pot_args;
    C.spl.reset();
    if (code == kInterpSph) {  // interpSphericalPotential's force spline
      const int nr = (int)pot_args[0];
      C.spl = std::make_shared<std::vector<NatSpline>>(1);
      (*C.spl)[0].build(pot_args + 1, pot_args + 1 + nr, nr);
      for (int j = 0; j < 6; j++) C.p[j] = pot_args[1 + 2 * nr + j];  // amp, rmin, rmax, Mmax, Phi0, Phimax
      pot_args += na;
      continue;
    }
    if (code == kNonInertialInterp) {  // initNonInertialFrameForceSplines, integrateFullOrbit.c:1435-1462
      const int ns = (int)pot_args[0], npts = (int)pot_args[1];
      const double* tg = pot_args + 2;
      C.spl = std::make_shared<std::vector<NatSpline>>(ns);
      for (int k = 0; k < ns; k++) (*C.spl)[k].build(tg, tg + (k + 1) * npts, npts);
      C.args = pot_args + 2 + (1 + ns) * npts;  // the 23 arguments of case 39, then tmin, tmax
      C.p[0] = C.args[0];
      pot_args = C.args + 25;
      continue;
    }
    if (code == kNonInertial) {  // its time functions, in parse order (integrateFullOrbit.c:689-704)
      const double* a = pot_args;
      const int ntf = (int)(3 * a[12] * (1 + 2 * a[11]) + (6 - 4 * a[13]) * a[15]);
      if (tf_next + ntf > g_ntfuncs) return -1;
      C.tf0 = tf_next;
      tf_next += ntf;
      C.args = a;
      C.p[0] = a[0];
      pot_args += 23;
      continue;
    }
    if (code == kFDM) {  // its own nPts, r, sigma_r, then 18: Chandrasekhar's 16, m/hbar, const factor
      const int n = (int)pot_args[0];
      C.args = pot_args + 1 + 2 * n;
      C.p[0] = C.args[0];
      pot_args = C.args + 18;
      continue;
    }
    if (code == kChandrasekhar) {  // initChandrasekharDynamicalFrictionSplines, integrateFullOrbit.c:1464-1491
      const int n = (int)pot_args[0];
      const double* ra = pot_args + 1;
      const double* own = ra + 2 * n;  // amp, 8 cache slots, ms, rhm, gamma^2, lnLambda, minr^2, ro, rf
      std::vector<double> rn(n);
      for (int j = 0; j < n; j++) rn[j] = (ra[j] - own[14]) / (own[15] - own[14]);
      C.spl = std::make_shared<std::vector<NatSpline>>(1);
      (*C.spl)[0].build(rn.data(), ra + n, n);
      C.args = own;
      C.p[0] = own[0];
      pot_args = own + 16;
      continue;
    }
    if (code == kMovingObjectWrap) {  // initMovingObjectSplines, integrateFullOrbit.c:1392-1431
      const int n = (int)pot_args[0], nc = planar ? 2 : 3;
      const double* tarr = pot_args + 1;
      const double to = tarr[(nc + 1) * n + 1], tf = tarr[(nc + 1) * n + 2];
      std::vector<double> tt(n);
      for (int j = 0; j < n; j++) tt[j] = (tarr[j] - to) / (tf - to);
      C.spl = std::make_shared<std::vector<NatSpline>>(nc);
      for (int k = 0; k < nc; k++) (*C.spl)[k].build(tt.data(), tarr + (k + 1) * n, n);
      C.p[0] = tarr[(nc + 1) * n];  // amp
      C.p[1] = to;
      C.p[2] = tf;
      pot_args += movingobject_nargs(pot_args, planar);
      continue;
    }
    if (is_ellipsoidal(code) || code == kSpiralArms || code == kDoubleExpDisk || code == kSCF ||
        code == kDiskSCF || code == kInterpRZ || code == kMultipole) {
      C.p[0] = pot_args[0];
      pot_args += na;
      continue;
    }
    for (int j = 0; j < na; j++) {  // the first 16 in p (all of them but RotateAndTilt's)
      const double v = *pot_args++;
      if (j < 16) C.p[j] = v;
    }
    C.tfunc = nullptr;
    if (code == kTimeDepAmpWrap) {
      if (tf_next >= g_ntfuncs) return -1;  // needs a time function
      C.tfunc = g_tfuncs[tf_next++];
    }
    if (code == kFlattenedPower && planar) {
      // integratePlanarOrbit.py:183-186 packs (amp, alpha, core^2); held in
      // the full layout (q^2 = 1): at z = 0 the full functions are the planar
      // ones (FlattenedPowerPotential.c PlanarRforce, PlanarR2deriv)
      C.p[3] = C.p[2];
      C.p[2] = 1.;
    }
    if (code == kPowerSphCutoff) {
      double alpha = C.p[1], rc = C.p[2];
      C.c[0] = 1.5 - 0.5 * alpha;                  // a for the force
      C.c[1] = tgamma(1.5 - 0.5 * alpha);          // Gamma(a)
      C.c[2] = 2. * M_PI * pow(rc, 3. - alpha);    // 2 pi rc^(3-alpha)
      C.c[3] = 1. - 0.5 * alpha;                   // for Eval
      C.c[4] = tgamma(1. - 0.5 * alpha);
    }
  }
  return 0;
}

// Decode galpy's packed potential; out needs room for kOracleMaxComp
// components (npot top-level ones first).  Returns 0 on success, -1 on an
// unsupported code or too many components.
inline int parse_potential(int npot, const int* pot_type, const double* pot_args,
                           Component* out, bool planar = false) {
  if (npot < 0 || npot > kOracleMaxComp) return -1;
  int next = npot, tf_next = 0;
  return parse_components(npot, pot_type, pot_args, out, next, out, planar, tf_next);
}

// DehnenSmoothWrapperPotential.c:4-16 (dehnenSmooth) and
// GaussianAmplitudeWrapperPotential.c:3-5 (gaussSmooth)
template <class R> R dehnen_smooth(R t, double tform, double tsteady, bool grow) {
  R smooth;
  if (t < R(tform))
    smooth = 0.;
  else if (t < R(tsteady)) {
    R xi = R(2.) * (t - R(tform)) / R(tsteady - tform) - R(1.);
    smooth = R(3. / 16.) * pow(xi, R(5.)) - R(5. / 8.) * pow(xi, R(3.)) + R(15. / 16.) * xi + R(.5);
  } else
    smooth = 1.;
  return grow ? smooth : R(1.) - smooth;
}
template <class R> R gauss_smooth(R t, double to, double sigma2) {
  return exp(R(-0.5) * (t - R(to)) * (t - R(to)) / R(sigma2));
}

// The amplitude factor and the angle offset of a wrapper at (R, t), and
// dOffset/dR (CorotatingRotationWrapperPotential.c:3-15): the wrapped
// components are evaluated at phi - offset and their forces scaled by the
// factor; the corotating wrapper's radial force gains -phitorque dOffset/dR.
template <class R> void wrapper_transform(const Component& C, R Rc, R t, R& fac, R& off, R& doff) {
  fac = R(C.p[0]);
  off = 0.;
  doff = 0.;
  switch (C.type) {
    case kDehnenSmoothWrap: fac = fac * dehnen_smooth(t, C.p[1], C.p[2], C.p[3] != 0.); break;
    case kGaussianAmpWrap: fac = fac * gauss_smooth(t, C.p[1], C.p[2]); break;
    // TimeDependentAmplitudeWrapperPotential.c:3-40: amp A(t)
    case kTimeDepAmpWrap: fac = fac * R(C.tfunc(dval(t))); break;
    case kSolidBodyRotWrap: off = R(C.p[1]) * t + R(C.p[2]); break;  // SolidBodyRotationWrapperPotential.c:4-9
    case kCorotatingRotWrap: {
      double vpo = C.p[1], beta = C.p[2], pa = C.p[3], to = C.p[4];
      off = R(vpo) * pow(Rc, R(beta - 1.)) * (t - R(to)) + R(pa);
      doff = R(vpo * (beta - 1.)) * pow(Rc, R(beta - 2.)) * (t - R(to));
      break;
    }
    default: break;
  }
}

// PowerSphericalPotentialwCutoff.c:9-11 (mass), :25-46 (Rforce), :60-81 (zforce)
// Returns the radial factor force/r-vector: F_R = f*R, F_z = f*z.
template <class R> R psc_force_factor(const Component& C, R r2) {
  double amp = C.p[0];
  double rc = C.p[2];
  R x = r2 / R(rc) / R(rc);
  // mass = 2 pi rc^(3-alpha) * (Gamma(a) * P(a, x)), P = gamma(a,x)/Gamma(a)
  R P = lower_incgamma(C.c[0], x, C.c[1]) / R(C.c[1]);
  R mass = R(C.c[2]) * (R(C.c[1]) * P);
  return -R(amp) * mass / pow(r2, R(1.5));
}
// PowerSphericalPotentialwCutoff.c:12-24
template <class R> R psc_eval(const Component& C, R Rc, R Z) {
  double amp = C.p[0], rc = C.p[2];
  R r2 = Rc * Rc + Z * Z;
  R r = sqrt(r2);
  R x = r2 / R(rc) / R(rc);
  R g1 = lower_incgamma(C.c[3], x, C.c[4]);   // Gamma(1-a/2) - Gamma(1-a/2, x)
  R g2 = lower_incgamma(C.c[0], x, C.c[1]);   // Gamma(1.5-a/2) - Gamma(1.5-a/2, x)
  return R(amp) * R(C.c[2]) / r * (r / R(rc) * g1 - g2);
}

// MiyamotoNagaiPotential.c:5-15 (Eval), :16-26 (Rforce), :38-53 (zforce)
template <class R> void mn_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  double amp = C.p[0], a = C.p[1], b = C.p[2];
  FR = -R(amp) * Rc * pow(Rc * Rc + pow(R(a) + pow(z * z + R(b * b), R(0.5)), R(2.)), R(-1.5));
  R sqrtbz = pow(R(b * b) + z * z, R(.5));
  R asqrtbz = R(a) + sqrtbz;
  if (a == 0.)
    Fz = R(amp) * (-z * pow(Rc * Rc + asqrtbz * asqrtbz, R(-1.5)));
  else
    Fz = R(amp) * (-z * asqrtbz / sqrtbz * pow(Rc * Rc + asqrtbz * asqrtbz, R(-1.5)));
}
template <class R> R mn_eval(const Component& C, R Rc, R z) {
  double amp = C.p[0], a = C.p[1], b = C.p[2];
  return -R(amp) * pow(Rc * Rc + pow(R(a) + pow(z * z + R(b * b), R(0.5)), R(2.)), R(-0.5));
}

// NFWPotential.c:5-15 (Eval), :16-27 (Rforce), :38-49 (zforce)
template <class R> void nfw_forces(const Component& C, R Rc, R Z, R& FR, R& Fz) {
  double amp = C.p[0], a = C.p[1];
  R Rz = Rc * Rc + Z * Z;
  R sqrtRz = pow(Rz, R(0.5));
  // the reference evaluates the bracket twice (once per component); the
  // value is identical, so it is formed once here
  R br = R(1.) / Rz / (R(a) + sqrtRz) - log(R(1.) + sqrtRz / R(a)) / sqrtRz / Rz;
  FR = R(amp) * Rc * br;
  Fz = R(amp) * Z * br;
}
template <class R> R nfw_eval(const Component& C, R Rc, R Z) {
  double amp = C.p[0], a = C.p[1];
  R sqrtRz = pow(Rc * Rc + Z * Z, R(0.5));
  return -R(amp) * log(R(1.) + sqrtRz / R(a)) / sqrtRz;
}

// HernquistPotential.c:5-14 (Eval), :15-25 (Rforce), :36-46 (zforce)
template <class R> void hq_forces(const Component& C, R Rc, R Z, R& FR, R& Fz) {
  double amp = C.p[0], a = C.p[1];
  R sqrtRz = pow(Rc * Rc + Z * Z, R(0.5));
  R f = pow(R(1.) + sqrtRz / R(a), R(-2.));
  FR = -R(amp) * Rc / R(a) / sqrtRz * f / R(2.) / R(a);
  Fz = -R(amp) * Z / R(a) / sqrtRz * f / R(2.) / R(a);
}
template <class R> R hq_eval(const Component& C, R Rc, R Z) {
  double amp = C.p[0], a = C.p[1];
  R sqrtRz = pow(Rc * Rc + Z * Z, R(0.5));
  return -R(amp) / (R(1.) + sqrtRz / R(a)) / R(2.) / R(a);
}

// LogarithmicHaloPotential.c:5-20 (Eval), :21-38 (Rforce), :55-72 (zforce),
// :73-91 (phitorque).  args: amp, q, core^2, 1-1/b^2 (>=1 means axisymmetric)
template <class R> void lh_forces(const Component& C, R Rc, R Z, R phi, R& FR, R& Fz, R& Fphi) {
  double amp = C.p[0], q = C.p[1], c = C.p[2], onem1overb2 = C.p[3];
  R zq = Z / R(q);
  if (onem1overb2 < 1) {
    R s = sin(phi);
    R Rt2 = Rc * Rc * (R(1.) - R(onem1overb2) * pow(s, R(2)));
    R den = Rt2 + zq * zq + R(c);
    FR = -R(amp) * Rt2 / Rc / den;
    Fz = -R(amp) * zq / R(q) / den;
    Fphi = R(amp) * Rc * Rc / den * sin(R(2.) * phi) * R(onem1overb2) / R(2.);
  } else {
    FR = -R(amp) * Rc / (Rc * Rc + zq * zq + R(c));
    Fz = -R(amp) * zq / R(q) / (Rc * Rc + zq * zq + R(c));
    Fphi = R(0.);
  }
}
template <class R> R lh_eval(const Component& C, R Rc, R Z, R phi) {
  double amp = C.p[0], q = C.p[1], c = C.p[2], onem1overb2 = C.p[3];
  R zq = Z / R(q);
  if (onem1overb2 < 1)
    return R(0.5) * R(amp) * log(Rc * Rc * (R(1.) - R(onem1overb2) * pow(sin(phi), R(2))) + zq * zq + R(c));
  return R(0.5) * R(amp) * log(Rc * Rc + zq * zq + R(c));
}

// DehnenBarPotential.c:5-17 (smooth), :18-40 (Rforce), :60-82 (phitorque),
// :102-124 (zforce).  args: amp*Af, tform, tsteady, rb, omegab, barphi
template <class R> R dehnen_smooth(R t, double tform, double tsteady) {
  if (t < R(tform)) return R(0.);
  if (t < R(tsteady)) {
    R deltat = t - R(tform);
    R xi = R(2.) * deltat / R(tsteady - tform) - R(1.);
    return R(3. / 16.) * pow(xi, R(5.)) - R(5. / 8.) * pow(xi, R(3.)) + R(15. / 16.) * xi + R(.5);
  }
  return R(1.);
}
template <class R> void bar_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fphi) {
  double amp = C.p[0], tform = C.p[1], tsteady = C.p[2], rb = C.p[3], omegab = C.p[4], barphi = C.p[5];
  R smooth = dehnen_smooth(t, tform, tsteady);
  R r2 = Rc * Rc + z * z;
  R r = sqrt(r2);
  R ang = R(2.) * (phi - R(omegab) * t - R(barphi));
  R cang = cos(ang);
  R sang = sin(ang);
  if (r <= R(rb)) {
    FR = -R(amp) * smooth * cang * (pow(r / R(rb), R(3.)) * Rc * (R(3.) * Rc * Rc + R(2.) * z * z) - R(4.) * Rc * z * z) / pow(r, R(4.));
    Fphi = R(2.) * R(amp) * smooth * sang * (pow(r / R(rb), R(3.)) - R(2.)) * Rc * Rc / r2;
    Fz = -R(amp) * smooth * cang * (pow(r / R(rb), R(3.)) + R(4.)) * Rc * Rc * z / pow(r, R(4.));
  } else {
    FR = -R(amp) * smooth * cang * pow(R(rb) / r, R(3.)) * Rc / pow(r, R(4)) * (R(3.) * Rc * Rc - R(2.) * z * z);
    Fphi = -R(2.) * R(amp) * smooth * sang * pow(R(rb) / r, R(3.)) * Rc * Rc / r2;
    Fz = -R(5.) * R(amp) * smooth * cang * pow(R(rb) / r, R(3.)) * Rc * Rc * z / pow(r, R(4));
  }
}
// DehnenBarPotential.py:173-200 (the C library has no Eval for the bar)
template <class R> R bar_eval(const Component& C, R Rc, R z, R phi, R t) {
  double amp = C.p[0], tform = C.p[1], tsteady = C.p[2], rb = C.p[3], omegab = C.p[4], barphi = C.p[5];
  R smooth = dehnen_smooth(t, tform, tsteady);
  R r2 = Rc * Rc + z * z;
  R r = sqrt(r2);
  R out;
  if (r <= R(rb))
    out = (pow(r / R(rb), R(3.)) - R(2.)) * (Rc * Rc / r2);
  else
    out = -pow(R(rb) / r, R(3.)) * R(1.) / (R(1.) + z * z / (Rc * Rc));
  return out * R(amp) * smooth * cos(R(2.) * (phi - R(omegab) * t - R(barphi)));
}

// ---------------------------------------------------------------------------
// Spherical components (codes 7, 10, 14, 17, 18, 20, 33, 34, 35): each
// reference file writes Rforce = f(r) R, zforce = f(r) z, planarRforce =
// -Phi'(r) and planarR2deriv = Phi''(r), and builds R2deriv / z2deriv /
// Rzderiv from Phi' and Phi'' (e.g. IsochronePotential.c:60-92).  sph_dphi
// restates planarRforce (negated) and planarR2deriv at radius r; the 3D
// force is -Phi'(r) x / r.
// ---------------------------------------------------------------------------
// ExpTruncNFWPotential.c:11-33: F(r) = M(<r)/amp (a series below 1e-3
// min(a, rc)) and G(r) of the potential -(F/r + G)
template <class R> R exptruncnfw_F(R r, double a, double rc) {
  double small_r_thresh = 1e-3 * (a < rc ? a : rc);
  if (r < R(small_r_thresh)) {
    double c2 = 0.5;
    double c3 = -(1.0 / rc + 2.0 / a) / 3.0;
    double c4 = (0.5 / (rc * rc) + 2.0 / (rc * a) + 3.0 / (a * a)) / 4.0;
    double c5 = -(1.0 / (6.0 * rc * rc * rc) + 1.0 / (rc * rc * a) + 3.0 / (rc * a * a) + 4.0 / (a * a * a)) / 5.0;
    return (r * r / R(a * a)) * (R(c2) + r * (R(c3) + r * (R(c4) + r * R(c5))));
  }
  double alpha = a / rc;
  R beta = (R(a) + r) / R(rc);
  // exp(alpha) (1 + alpha) (E1(alpha) - E1(beta)), with exp(alpha) E1(beta)
  // = exp(-r/rc) (e^beta E1(beta)) so that nothing overflows at large r
  R ex = exp(-r / R(rc));
  return R(1. + alpha) * (e1_scaled(R(alpha)) - ex * e1_scaled(beta)) - R(1.) + R(a) * ex / (R(a) + r);
}
template <class R> R exptruncnfw_G(R r, double a, double rc) {
  R beta = (R(a) + r) / R(rc);
  return exp(-r / R(rc)) / (R(a) + r) - exp(-r / R(rc)) * e1_scaled(beta) / R(rc);
}

template <class R> void sph_dphi(const Component& C, R r, R& Phip, R& Phipp) {
  double amp = C.p[0];
  switch (C.type) {
    case kPowerSph: {  // PowerSphericalPotential.c: planarRforce, planarR2deriv
      double alpha = C.p[1];
      Phip = R(amp) * pow(r, R(-alpha + 1.));
      Phipp = R(amp) * R(1. - alpha) * pow(r, R(-alpha));
      break;
    }
    case kJaffe: {  // JaffePotential.c
      double a = C.p[1];
      Phip = R(amp) * pow(r, R(-2.)) / (R(1.) + R(a) / r);
      Phipp = -R(amp) * (R(a) + R(2.) * r) * pow(r, R(-4.)) * pow(R(1.) + R(a) / r, R(-2.));
      break;
    }
    case kIsochrone: {  // IsochronePotential.c
      double b = C.p[1];
      R r2 = r * r;
      R rb = sqrt(r2 + R(b * b));
      Phip = R(amp) * r / rb * pow(R(b) + rb, R(-2.));
      Phipp = -R(amp) * (-R(b * b * b) - R(b * b) * rb + R(2.) * r2 * rb) * pow(rb * (R(b) + rb), R(-3.));
      break;
    }
    case kPlummer: {  // PlummerPotential.c
      double b2 = C.p[1] * C.p[1];
      Phip = R(amp) * r * pow(r * r + R(b2), R(-1.5));
      Phipp = R(amp) * (R(b2) - R(2.) * r * r) * pow(r * r + R(b2), R(-2.5));
      break;
    }
    case kPseudoIso: {  // PseudoIsothermalPotential.c
      double a = C.p[1], a2 = a * a;
      R at = atan(r / R(a));
      Phip = R(amp) * (R(1.) / r - R(a) / r / r * at) / R(a);
      Phipp = R(amp) / (r * r) * (R(2. * a) / r * at - (R(2. * a2) + r * r) / (R(a2) + r * r)) / R(a);
      break;
    }
    case kBurkert: {  // BurkertPotential.c
      double a = C.p[1];
      R x = r / R(a);
      R B = R(M_PI) - R(2.) * atan(R(1.) / x) - R(2.) * log(R(1.) + x) - log(R(1.) + x * x);
      Phip = -R(amp) * R(a) * R(M_PI) / x / x * B;
      R r5 = pow(r, R(5));
      Phipp = -R(amp) * R(M_PI) * R(a * a * a) / r5 *
              (R(-4.) * r5 / (R(a * a) + r * r) / (R(a) + r) - R(2.) * r * r * B);
      break;
    }
    case kDehnenCoreSph: {  // DehnenCorePotential.c
      double a = C.p[1];
      Phip = R(amp) * (r / pow(R(a) + r, R(3.)) / R(3.));
      Phipp = -R(amp) * (pow(R(a) + r, R(-4.)) * (R(2.) * r - R(a))) / R(3.);
      break;
    }
    case kDehnenSph: {  // DehnenPotential.c
      double a = C.p[1], alpha = C.p[2];
      Phip = R(amp) * (r * pow(R(1.) + R(a) / r, R(alpha)) / pow(R(a) + r, R(3.)) / R(3. - alpha));
      Phipp = -R(amp) * (pow(R(1.) + R(a) / r, R(alpha)) * pow(R(a) + r, R(-4.)) *
                         (R(2.) * r + R(a * (alpha - 1.)))) / R(3. - alpha);
      break;
    }
    case kHomogSphere: {  // HomogeneousSpherePotential.c
      double R2 = C.p[1], R3 = C.p[2];
      if (r * r < R(R2)) {
        Phip = R(2.) * R(amp) * r;
        Phipp = R(2.) * R(amp);
      } else {
        Phip = R(2.) * R(amp) * R(R3) / (r * r);
        Phipp = R(-4.) * R(amp) * R(R3) / pow(r * r, R(1.5));
      }
      break;
    }
    case kEinasto: {
      // EinastoPotential.c rforce (:27-47), r2deriv (:49-70).  galpy forms
      // Q(3n, x) - 1, which cancels to nothing for x = s^(1/n) << 3n (a
      // relative error of 1e-3 at r = 1e-3 h for n = 2.5); restated with
      // -P(3n, x) = -gamma(3n, x) / Gamma(3n), the same function without the
      // cancellation.
      double h = C.p[1], n = C.p[2];
      R s = r / R(h);
      double g3n = tgamma(3. * n);
      R xs = pow(s, R(1. / n));
      R Qm1 = -lower_incgamma(3. * n, xs, g3n) / R(g3n);
      Phip = -R(amp) * (R(4. * M_PI * h * n * g3n) * pow(s, R(-2.)) * Qm1);
      Phipp = R(amp) * -(R(4. * M_PI * n * g3n) *
                         ((R(-2.) * pow(s, R(-3.))) * Qm1 - (R(1. / n) * exp(-xs) / R(g3n))));
      break;
    }
    case kExpTruncNFW: {  // ExpTruncNFWPotential.c rforce, r2deriv and F (:11-28)
      double a = C.p[1], rc = C.p[2];
      R F = exptruncnfw_F(r, a, rc);
      Phip = R(amp) * F / (r * r);
      R dens = exp(-r / R(rc)) / (R(4. * M_PI * a * a) * r * pow(R(1.) + r / R(a), R(2.)));
      Phipp = R(amp) * (R(4. * M_PI) * dens - R(2.) * F / (r * r * r));
      break;
    }
    case kTwoPowerSph: {  // TwoPowerSphericalPotential.c PlanarRforce, PlanarR2deriv
      double a = C.p[1], alpha = C.p[2], beta = C.p[3];
      R Aa = R(pow(a, alpha - 3.) / (3. - alpha));
      R hyper = hyp2f1_ap1(3. - alpha, beta - alpha, 3. - beta, r / R(a));
      R hyper_deriv = R((3. - alpha) * (beta - alpha) / (4. - alpha)) *
                      hyp2f1_ap1(4. - alpha, 1. + beta - alpha, 3. - beta, r / R(a));
      Phip = R(amp) * pow(r, R(1. - alpha)) * Aa * hyper;
      R term1 = Aa * pow(r, R(-alpha)) * hyper;
      R term2 = R(-alpha) * Aa * pow(r, R(-alpha)) * hyper;
      R term3 = -Aa * pow(r, R(1. - alpha)) * R(1. / a) * hyper_deriv;
      Phipp = R(amp) * (term1 + term2 + term3);
      break;
    }
    default:
      Phip = 0.;
      Phipp = 0.;
  }
}

inline bool is_spherical_new(int code) {
  return code == kPowerSph || code == kJaffe || code == kIsochrone || code == kPlummer ||
         code == kPseudoIso || code == kBurkert || code == kDehnenCoreSph || code == kDehnenSph ||
         code == kHomogSphere || code == kEinasto || code == kExpTruncNFW || code == kTwoPowerSph;
}

// 3D forces of the spherical components as their files write them
// (Rforce = -amp R ..., zforce = -amp z ...): F = -Phi'(r) (R, z) / r.
template <class R> void sph_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  R r = sqrt(Rc * Rc + z * z);
  R Phip, Phipp;
  sph_dphi(C, r, Phip, Phipp);
  FR = -Phip * Rc / r;
  Fz = -Phip * z / r;
}

// Potential values (the files' Eval functions).
template <class R> R sph_eval(const Component& C, R r) {
  double amp = C.p[0];
  switch (C.type) {
    case kPowerSph: {
      double alpha = C.p[1];
      if (alpha == 2.) return R(0.5) * R(amp) * log(r * r);
      return -R(amp) * pow(r * r, R(1. - 0.5 * alpha)) / R(alpha - 2.);
    }
    case kJaffe: return -R(amp) * log(R(1.) + R(C.p[1]) / r) / R(C.p[1]);
    case kIsochrone: return -R(amp) / (R(C.p[1]) + sqrt(r * r + R(C.p[1] * C.p[1])));
    case kPlummer: return -R(amp) / sqrt(r * r + R(C.p[1] * C.p[1]));
    case kPseudoIso: {
      double a = C.p[1];
      return R(amp) * (R(0.5) * log(R(1.) + r * r / R(a * a)) + R(a) / r * atan(r / R(a))) / R(a);
    }
    case kBurkert: {
      double a = C.p[1];
      R x = r / R(a);
      return -R(amp) * R(a * a * M_PI) / x *
             (R(-M_PI) + R(2.) * (R(1.) + x) * atan(R(1.) / x) + R(2.) * (R(1.) + x) * log(R(1.) + x) +
              (R(1.) - x) * log(R(1.) + x * x));
    }
    case kDehnenCoreSph: {
      double a = C.p[1];
      return -R(amp) * (R(1.) - pow(r / (r + R(a)), R(2.))) / R(6. * a);
    }
    case kDehnenSph: {
      double a = C.p[1], alpha = C.p[2];
      return -R(amp) * (R(1.) - pow(r / (r + R(a)), R(2. - alpha))) / R(a * (2. - alpha) * (3. - alpha));
    }
    case kHomogSphere: {
      double R2 = C.p[1], R3 = C.p[2];
      if (r * r < R(R2)) return R(amp) * (r * r - R(3. * R2));
      return R(-2.) * R(amp) * R(R3) / r;
    }
    case kEinasto: {  // EinastoPotential.c revaluate (:9-26; GSL's gamma_inc is the upper Gamma(a,x));
      double h = C.p[1], n = C.p[2];
      R s = r / R(h);
      R xs = pow(s, R(1. / n));
      double g3n = tgamma(3. * n);
      // 1 - Q(3n, x) as P(3n, x) (see sph_dphi)
      return R(amp) * (-(R(4. * M_PI * h * h * n * g3n)) * pow(s, R(-1.)) *
                       (lower_incgamma(3. * n, xs, g3n) / R(g3n) +
                        s * upper_incgamma(2. * n, xs, tgamma(2. * n)) / R(g3n)));
    }
    case kExpTruncNFW:  // ExpTruncNFWPotential.c revaluate
      return R(amp) * -(exptruncnfw_F(r, C.p[1], C.p[2]) / r + exptruncnfw_G(r, C.p[1], C.p[2]));
    case kTwoPowerSph: return R(NAN);  // its Eval needs 2F1 outside the restated family
    default: return R(0.);
  }
}

// KuzminDiskPotential.c: Eval, Rforce, zforce, planarRforce and the second
// derivatives (R2deriv, z2deriv, Rzderiv)
template <class R> void kuzmin_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  double amp = C.p[0], a = C.p[1];
  R az = R(a) + fabs(z);
  R dm = pow(Rc * Rc + pow(az, R(2.)), R(-1.5));
  FR = -R(amp) * Rc * dm;
  R zsign = (z > R(0.)) ? R(1.) : ((z < R(0.)) ? R(-1.) : R(0.));
  Fz = -zsign * R(amp) * dm * az;
}
template <class R> R kuzmin_eval(const Component& C, R Rc, R z) {
  return -R(C.p[0]) * pow(Rc * Rc + pow(R(C.p[1]) + fabs(z), R(2.)), R(-0.5));
}

// KuzminKutuzovStaeckelPotential.c: Eval (:5-23, n clamped at 0), Rforce
// (:24-45), zforce (:64-85) and planarRforce (:46-63, in the plane n =
// -gamma and l = R^2 - alpha), with the reference's coordinate set-up:
// gamma = Delta^2 / (1 - ac^2), alpha = gamma - Delta^2 and (l, n) the roots
// of the spheroidal-coordinate quadratic of (R, z).
template <class R> void kks_coords(const Component& C, R Rc, R z, R& l, R& n, R& sd) {
  const double ac = C.p[1], Delta = C.p[2];
  const double gamma = Delta * Delta / (1. - ac * ac);
  const double alpha = gamma - Delta * Delta;
  R term = Rc * Rc + z * z - R(alpha) - R(gamma);
  R discr = pow(Rc * Rc + z * z - R(Delta * Delta), R(2.)) + (R(4.) * R(Delta * Delta) * Rc * Rc);
  sd = sqrt(discr);
  l = R(0.5) * (term + sd);
  n = R(0.5) * (term - sd);
}
template <class R> void kks_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  const double amp = C.p[0], Delta = C.p[2];
  R l, n, sd;
  kks_coords(C, Rc, z, l, n, sd);
  R dVdl = R(0.5) / sqrt(l) / pow(sqrt(l) + sqrt(n), R(2.));
  R dVdn = R(0.5) / sqrt(n) / pow(sqrt(l) + sqrt(n), R(2.));
  R dldR = Rc * (R(1.) + (Rc * Rc + z * z + R(Delta * Delta)) / sd);
  R dndR = Rc * (R(1.) - (Rc * Rc + z * z + R(Delta * Delta)) / sd);
  R dldz = z * (R(1.) + (Rc * Rc + z * z - R(Delta * Delta)) / sd);
  R dndz = z * (R(1.) - (Rc * Rc + z * z - R(Delta * Delta)) / sd);
  FR = -R(amp) * (dldR * dVdl + dndR * dVdn);
  Fz = -R(amp) * (dldz * dVdl + dndz * dVdn);
}
template <class R> R kks_eval(const Component& C, R Rc, R z) {
  R l, n, sd;
  kks_coords(C, Rc, z, l, n, sd);
  n = (n > R(0.)) ? n : R(0.);
  return -R(C.p[0]) / (sqrt(l) + sqrt(n));
}
template <class R> R kks_planar_rforce(const Component& C, R Rc) {
  const double amp = C.p[0], ac = C.p[1], Delta = C.p[2];
  const double gamma = Delta * Delta / (1. - ac * ac);
  const double alpha = gamma - Delta * Delta;
  R l = Rc * Rc - R(alpha);
  R n = R(-gamma);
  R dVdl = R(0.5) / sqrt(l) / pow(sqrt(l) + sqrt(n), R(2.));
  return -R(amp) * (R(2.) * Rc) * dVdl;
}

// SoftenedNeedleBarPotential.c: compute_TpTm (:5-12), Eval (:13-29), the
// aligned-frame forces rotated back (:30-63), Rforce (:64-85), phitorque
// (:90-...) and zforce; the planar versions are these at z = 0.
template <class R> void needle_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz,
                                      R& Fphi) {
  const double amp = C.p[0], a = C.p[1], b = C.p[2], c2 = C.p[3], pa = C.p[4], om = C.p[5];
  R x = Rc * cos(phi - R(pa) - R(om) * t), y = Rc * sin(phi - R(pa) - R(om) * t);
  R secondpart = y * y + pow(R(b) + sqrt(z * z + R(c2)), R(2.));
  R Tp = sqrt(pow(R(a) + x, R(2.)) + secondpart);
  R Tm = sqrt(pow(R(a) - x, R(2.)) + secondpart);
  R zc = sqrt(z * z + R(c2));
  R Fx = R(-2.) * x / Tp / Tm / (Tp + Tm);
  R Fy = -y / R(2.) / Tp / Tm * (Tp + Tm - R(4.) * x * x / (Tp + Tm)) / (y * y + pow(R(b) + zc, R(2.)));
  R Fzz = -z / R(2.) / Tp / Tm * (Tp + Tm - R(4.) * x * x / (Tp + Tm)) /
          (y * y + pow(R(b) + zc, R(2.))) * (R(b) + zc) / zc;
  R cp = cos(R(pa) + R(om) * t), sp = sin(R(pa) + R(om) * t);
  R Fxi = cp * Fx - sp * Fy, Fyi = sp * Fx + cp * Fy;
  FR = R(amp) * (cos(phi) * Fxi + sin(phi) * Fyi);
  Fphi = R(amp) * Rc * (-sin(phi) * Fxi + cos(phi) * Fyi);
  Fz = R(amp) * Fzz;
}
template <class R> R needle_eval(const Component& C, R Rc, R z, R phi, R t) {
  const double amp = C.p[0], a = C.p[1], b = C.p[2], c2 = C.p[3], pa = C.p[4], om = C.p[5];
  R x = Rc * cos(phi - R(pa) - R(om) * t), y = Rc * sin(phi - R(pa) - R(om) * t);
  R secondpart = y * y + pow(R(b) + sqrt(z * z + R(c2)), R(2.));
  R Tp = sqrt(pow(R(a) + x, R(2.)) + secondpart);
  R Tm = sqrt(pow(R(a) - x, R(2.)) + secondpart);
  return R(0.5) * R(amp) * log((x - R(a) + Tm) / (x + R(a) + Tp)) / R(a);
}

// EllipticalDiskPotential.c (planar only): EllipticalDiskSmooth (:4-20),
// Rforce (:21-35), phitorque (:36-50).
template <class R> R ellipdisk_smooth(R t, double tform, double tsteady) {
  R smooth;
  if (!std::isnan(tform)) {
    if (t < R(tform)) smooth = 0.;
    else if (t < R(tsteady)) {
      R deltat = t - R(tform);
      R xi = R(2.) * deltat / R(tsteady - tform) - R(1.);
      smooth = (R(3. / 16.) * pow(xi, R(5.)) - R(5. / 8.) * pow(xi, R(3.)) + R(15. / 16.) * xi + R(.5));
    } else smooth = 1.;
  } else smooth = 1.;
  return smooth;
}
template <class R> void ellipdisk_planar_forces(const Component& C, R Rc, R phi, R t, R& FR, R& Fphi) {
  const double amp = C.p[0], twophio = C.p[3], p = C.p[4], phib = C.p[5];
  R smooth = ellipdisk_smooth(t, C.p[1], C.p[2]);
  FR = -R(amp) * smooth * R(p) * R(twophio) / R(2.) * pow(Rc, R(p - 1.)) * cos(R(2.) * (phi - R(phib)));
  Fphi = R(amp) * smooth * R(twophio) * pow(Rc, R(p)) * sin(R(2.) * (phi - R(phib)));
}

// ---------------------------------------------------------------------------
// EllipsoidalPotential family.  The density rho(m) and psi(m) =
// -int_m^inf dm'^2 rho of each member (TwoPowerTriaxialPotential.c:7-118,
// PerfectEllipsoidPotential.c:4-20, TriaxialGaussianPotential.c:4-16,
// PowerTriaxialPotential.c:4-20; psi arguments start at args + 17), the
// Gauss-Legendre force sum EllipsoidalPotentialxyzforces_xyz
// (EllipsoidalPotential.c:40-77), its cylindrical projections (:97-114) and
// the potential EllipsoidalPotentialEval (:6-39).  TwoPowerTriaxial's psi
// needs the Gauss hypergeometric function and is not restated (NaN).
// ---------------------------------------------------------------------------
template <class R> R ellip_mdens(int type, const double* pa, R m) {
  switch (type) {
    case kTriaxHernquist: { R mpa = m + R(pa[0]); return R(pa[1]) * pow(mpa, R(-3.)) / m; }
    case kTriaxJaffe: { R mpa = m + R(pa[0]); return R(pa[1]) * R(pa[1]) * pow(m * mpa, R(-2.)); }
    case kTriaxNFW: { R mpa = m + R(pa[0]); return R(pa[1]) / m / mpa / mpa; }
    case kTwoPowerTriax: {
      R mpa = m / R(pa[0]) + R(1.);
      return pow(R(pa[0]) / m, R(pa[1])) / pow(mpa, R(pa[3]));
    }
    case kPerfectEllipsoid: { R m2 = m * m; return R(1.) / (R(pa[0]) + m2) / (R(pa[0]) + m2); }
    case kTriaxGaussian: return exp(m * m / R(pa[0]));
    case kPowerTriax: return pow(m * m, R(-pa[0] / 2.));
    default: return R(NAN);
  }
}
template <class R> R ellip_psi(int type, const double* pa, R m) {
  switch (type) {
    case kTriaxHernquist: { R mpa = m + R(pa[0]); return -R(pa[1]) / mpa / mpa; }
    case kTriaxJaffe: { R mpa = m + R(pa[0]); return R(2.) * R(pa[1]) * (R(pa[0]) / mpa + log(m / mpa)); }
    case kTriaxNFW: return R(-2.) * R(pa[1]) / (m + R(pa[0]));
    case kPerfectEllipsoid: return R(-1.) / (R(pa[0]) + m * m);
    case kTriaxGaussian: return R(pa[0]) * exp(m * m / R(pa[0]));
    case kPowerTriax: return R(2.) / R(2. - pa[0]) * pow(m * m, R(1. - pa[0] / 2.));
    default: return R(NAN);
  }
}
// the layout behind the psi arguments
struct EllipLayout {
  const double* psi;
  double b2, c2;
  bool aligned;
  const double* rot;
  int glorder;
  const double* glx;
  const double* glw;
};
inline EllipLayout ellip_layout(const Component& C) {
  EllipLayout L;
  const double* a = C.args;
  L.psi = a + 17;
  const double* e = a + 17 + (int)a[16];
  L.b2 = e[0];
  L.c2 = e[1];
  L.aligned = e[2] != 0.;
  L.rot = e + 3;
  L.glorder = (int)e[12];
  L.glx = e + 13;
  L.glw = e + 13 + L.glorder;
  return L;
}
// galpy_potentials.c:391-410
template <class R> void rotate_xyz(R& x, R& y, R& z, const double* rot) {
  R xp = R(rot[0]) * x + R(rot[1]) * y + R(rot[2]) * z;
  R yp = R(rot[3]) * x + R(rot[4]) * y + R(rot[5]) * z;
  R zp = R(rot[6]) * x + R(rot[7]) * y + R(rot[8]) * z;
  x = xp; y = yp; z = zp;
}
template <class R> void rotate_force_xyz(R& Fx, R& Fy, R& Fz, const double* rot) {
  R Fxp = R(rot[0]) * Fx + R(rot[3]) * Fy + R(rot[6]) * Fz;
  R Fyp = R(rot[1]) * Fx + R(rot[4]) * Fy + R(rot[7]) * Fz;
  R Fzp = R(rot[2]) * Fx + R(rot[5]) * Fy + R(rot[8]) * Fz;
  Fx = Fxp; Fy = Fyp; Fz = Fzp;
}
// Cartesian forces / amp at (x, y, z) in the inertial frame
template <class R> void ellip_forces_xyz(const Component& C, R x, R y, R z, R& Fx, R& Fy, R& Fz) {
  const EllipLayout L = ellip_layout(C);
  if (!L.aligned) rotate_xyz(x, y, z, L.rot);
  Fx = 0.; Fy = 0.; Fz = 0.;
  for (int ii = 0; ii < L.glorder; ii++) {
    R t = R(1.) / R(L.glx[ii]) / R(L.glx[ii]) - R(1.);
    R td = R(L.glw[ii]) * ellip_mdens(C.type, L.psi,
                                      sqrt(x * x / (R(1.) + t) + y * y / (R(L.b2) + t) + z * z / (R(L.c2) + t)));
    Fx += td * x / (R(1.) + t);
    Fy += td * y / (R(L.b2) + t);
    Fz += td * z / (R(L.c2) + t);
  }
  if (!L.aligned) rotate_force_xyz(Fx, Fy, Fz, L.rot);
}
template <class R> void ellip_forces(const Component& C, R Rc, R z, R phi, R& FR, R& Fz, R& Fphi) {
  R Fx, Fy, Fzz;
  ellip_forces_xyz(C, Rc * cos(phi), Rc * sin(phi), z, Fx, Fy, Fzz);
  const R amp = R(C.p[0]);
  FR = amp * (cos(phi) * Fx + sin(phi) * Fy);
  Fphi = amp * Rc * (-sin(phi) * Fx + cos(phi) * Fy);
  Fz = amp * Fzz;
}
template <class R> R ellip_eval(const Component& C, R Rc, R z, R phi) {
  const EllipLayout L = ellip_layout(C);
  R x = Rc * cos(phi), y = Rc * sin(phi);
  if (!L.aligned) rotate_xyz(x, y, z, L.rot);
  R out = 0.;
  for (int ii = 0; ii < L.glorder; ii++) {
    R s = R(1.) / R(L.glx[ii]) / R(L.glx[ii]) - R(1.);
    out += R(L.glw[ii]) *
           ellip_psi(C.type, L.psi, sqrt(x * x / (R(1.) + s) + y * y / (R(L.b2) + s) + z * z / (R(L.c2) + s)));
  }
  return R(-0.5) * R(C.p[0]) * out;
}

// Aggregated cylindrical forces: galpy_potentials.c:120-179 sums each of
// Rforce, phitorque, zforce over components in index order starting from 0.
// Summing the three per component gives the identical three sums.
// interpSphericalPotential.c:5-37 with SphericalPotential.c:4-38: the radial
// force is the natural spline of the tabulated force on [rmin, rmax), 0
// inside rmin and -M/r^2 outside rmax; the potential is minus the spline's
// integral from rmin plus Phi0, and -M/r + Phimax outside rmax.
// p: amp, rmin, rmax, M(<rmax), Phi0, Phimax.
template <class R> R interpsph_rforce(const Component& C, R r) {
  if (r >= R(C.p[2])) return -R(C.p[3]) / r / r;
  return r < R(C.p[1]) ? R(0.) : (*C.spl)[0].eval(r);
}
template <class R> R interpsph_reval(const Component& C, R r) {
  if (r >= R(C.p[2])) return -R(C.p[3]) / r + R(C.p[5]);
  return r < R(C.p[1]) ? R(0.) : -(*C.spl)[0].integ(r) + R(C.p[4]);
}
// MovingObjectPotential.c:12-120: the object's position from natural splines
// of its orbit over the normalised time (t - t0)/(tf - t0), clamped to
// [0, 1]; the wrapped (axisymmetric) potential at the cylindrical distance
// and the height difference.  p: amp, t0, tf.
template <class R> void movingobject_pos(const Component& C, R t, R& ox, R& oy, R& oz) {
  R d = (t - R(C.p[1])) / R(C.p[2] - C.p[1]);
  if (d < R(0.)) d = 0.;
  if (d > R(1.)) d = 1.;
  ox = (*C.spl)[0].eval(d);
  oy = (*C.spl)[1].eval(d);
  oz = C.spl->size() > 2 ? (*C.spl)[2].eval(d) : R(0.);
}

// SCFPotential.c (Hernquist & Ostriker 1992 basis), with GSL's Gegenbauer
// and Legendre arrays (absent here) restated from their recurrences:
//   C_0^a = 1, C_1^a = 2 a x, C_n^a = (2 (n - 1 + a) x C_{n-1}^a -
//     (n - 2 + 2a) C_{n-2}^a) / n  (gsl_sf_gegenpoly_array);
//   P_m^m = (-1)^m (2m-1)!! (1 - x^2)^(m/2), P_{m+1}^m = (2m+1) x P_m^m,
//   P_l^m = ((2l-1) x P_{l-1}^m - (l+m-1) P_{l-2}^m) / (l - m)
//     (no normalisation, Condon-Shortley phase: GSL_SF_LEGENDRE_NONE,
//     csphase -1), dP_l^m/dx = (l x P_l^m - (l+m) P_{l-1}^m) / (x^2 - 1);
// the sums as SCFPotential.c:366-438 (sum_expansion, sum_spher_forces) and
// the forces :865-895, coefficients interpolated in time (:121-172) when
// the potential is time-dependent.  Pinned by mpmath/scipy (Gegenbauer,
// Legendre), the Hernquist case, forces = -grad Phi and Poisson's equation.
inline void scf_gegenpoly(int nmax, double alpha, double x, double* c) {
  c[0] = 1.;
  if (nmax >= 1) c[1] = 2. * alpha * x;
  for (int k = 2; k <= nmax; k++)
    c[k] = (2. * (k - 1. + alpha) * x * c[k - 1] - (k - 2. + 2. * alpha) * c[k - 2]) / k;
}
// P_l^m(x) and dP_l^m/dx for 0 <= m < Meff, m <= l < L, stored [m][l]
inline void scf_legendre(double x, int L, int Meff, std::vector<double>& P, std::vector<double>& dP) {
  const double s = std::sqrt((1. - x) * (1. + x));
  double pmm = 1.;
  for (int m = 0; m < Meff && m < L; m++) {
    if (m > 0) pmm *= -(2. * m - 1.) * s;
    double pm2 = 0., pm1 = pmm;
    for (int l = m; l < L; l++) {
      double pl;
      if (l == m) pl = pmm;
      else if (l == m + 1) pl = (2. * m + 1.) * x * pmm;
      else pl = ((2. * l - 1.) * x * pm1 - (l + m - 1.) * pm2) / (l - m);
      if (l > m) { pm2 = pm1; pm1 = pl; }
      const double plm1 = (l == m) ? 0. : pm2;
      P[m * L + l] = pl;
      dP[m * L + l] = (l == 0) ? 0. : (l * x * pl - (l + m) * plm1) / (x * x - 1.);
    }
  }
}
// the (possibly time-interpolated) coefficient arrays of an SCF component
inline void scf_coeffs(const Component& C, double t, std::vector<double>& Ac, std::vector<double>& As) {
  const double* a = C.args;
  const int nonaxi = (int)a[1], NLM = (int)a[2] * (int)a[3] * (int)a[4], Nt = (int)a[5];
  Ac.assign(NLM, 0.);
  As.assign(NLM, 0.);
  if (Nt == 0) {
    for (int i = 0; i < NLM; i++) Ac[i] = a[6 + i];
    if (nonaxi) for (int i = 0; i < NLM; i++) As[i] = a[6 + NLM + i];
    return;
  }
  const double* tg = a + 6;
  int it;  // scf_find_time_interval: clamped to the boundary intervals
  if (t <= tg[0]) it = 0;
  else if (t >= tg[Nt - 1]) it = Nt - 2;
  else {
    int lo = 0, hi = Nt - 2;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tg[mid] <= t) lo = mid; else hi = mid - 1;
    }
    it = lo;
  }
  const double dt = t - tg[it];
  const double* ppc = tg + Nt;
  const double* pps = ppc + (size_t)NLM * 4 * (Nt - 1);
  for (int ci = 0; ci < NLM; ci++) {
    const double* cp = ppc + ((size_t)it * NLM + ci) * 4;
    Ac[ci] = ((cp[0] * dt + cp[1]) * dt + cp[2]) * dt + cp[3];
    if (nonaxi) {
      const double* sp = pps + ((size_t)it * NLM + ci) * 4;
      As[ci] = ((sp[0] * dt + sp[1]) * dt + sp[2]) * dt + sp[3];
    }
  }
}
// spherical derivatives F = (dPhi/dr, dPhi/dtheta, dPhi/dphi) with the
// reference's signs (F[0], F[1] minus the derivatives, F[2] the torque),
// and the potential; evaluated in double (not op-counted)
inline void scf_spher(const Component& C, double Rc, double z, double phi, double t, double* F, double& pot) {
  const double* a = C.args;
  const double sa = a[0];
  const int nonaxi = (int)a[1], N = (int)a[2], L = (int)a[3], M = (int)a[4];
  const int Meff = nonaxi ? M : 1;
  const double r = std::sqrt(Rc * Rc + z * z), theta = std::atan2(Rc, z);
  const double xi = (r - sa) / (r + sa);
  std::vector<double> Ac, As, P(Meff * L, 0.), dP(Meff * L, 0.), Cg(N * L), dC(N * L, 0.), phiT(N * L), dphiT(N * L);
  scf_coeffs(C, t, Ac, As);
  for (int l = 0; l < L; l++) {
    scf_gegenpoly(N - 1, 1.5 + 2 * l, xi, &Cg[l * N]);
    if (N > 1) scf_gegenpoly(N - 2, 2.5 + 2 * l, xi, &dC[l * N + 1]);
    for (int n = 0; n < N; n++) dC[l * N + n] *= 2.0 * (2 * l + 1.5);
  }
  double rterms = -1.0 / (r + sa);  // compute_phiTilde
  for (int l = 0; l < L; l++) {
    if (l > 0) rterms *= r * sa / ((sa + r) * (sa + r));
    for (int n = 0; n < N; n++) phiT[l * N + n] = rterms * Cg[l * N + n];
  }
  const double ar = sa + r;  // compute_dphiTilde
  double rterm = 1.0 / (r * ar * ar * ar);
  for (int l = 0; l < L; l++) {
    if (l > 0) rterm *= sa * r / (ar * ar);
    for (int n = 0; n < N; n++) {
      const int i = l * N + n;
      dphiT[i] = rterm * (((2 * l + 1) * r * ar - l * ar * ar) * Cg[i] - 2.0 * sa * r * dC[i]);
    }
  }
  scf_legendre(std::cos(theta), L, Meff, P, dP);
  F[0] = F[1] = F[2] = 0.;
  pot = 0.;
  for (int l = 0; l < L; l++) {
    for (int m = 0; m <= (nonaxi ? l : 0); m++) {
      const double mc = std::cos(m * phi), ms = std::sin(m * phi);
      const double Pl = P[m * L + l], dPl = dP[m * L + l];
      for (int n = 0; n < N; n++) {
        const int ci = n * L * M + l * M + m, ri = l * N + n;
        const double cs = nonaxi ? Ac[ci] * mc + As[ci] * ms : Ac[ci];
        const double sd = nonaxi ? Ac[ci] * ms - As[ci] * mc : 0.;
        F[0] -= cs * Pl * dphiT[ri];
        F[1] -= cs * dPl * phiT[ri];
        F[2] += m * sd * Pl * phiT[ri];
        pot += cs * Pl * phiT[ri];
      }
    }
  }
  const double s4 = std::sqrt(4.0 * M_PI);
  F[0] *= s4;
  F[1] *= s4 * (-std::sin(theta));
  F[2] *= s4;
  pot *= s4;
}
template <class R> void scf_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fp) {
  double F[3], pot;
  scf_spher(C, dval(Rc), dval(z), dval(phi), dval(t), F, pot);
  const double Rd = dval(Rc), zd = dval(z), r = std::sqrt(Rd * Rd + zd * zd);
  FR = R(F[0] * (Rd / r) + F[1] * (zd / (r * r)));
  Fz = R(F[0] * (zd / r) + F[1] * (-Rd / (r * r)));
  Fp = R(F[2]);
}
template <class R> R scf_eval(const Component& C, R Rc, R z, R phi, R t) {
  double F[3], pot;
  scf_spher(C, dval(Rc), dval(z), dval(phi), dval(t), F, pot);
  return R(pot);
}

// MultipoleExpansionPotential.c:411-507 (eval_radial_lm, below_grid_integrals),
// :526-600 (the fused time and radius evaluation), :719-1029 (potential and
// forces): R_lm = r^-(l+1) I_inner + r^l I_outer from the quintic
// coefficients of the interval holding r (Horner), the density held at its
// innermost value below rmin, only I_inner(rmax) r^-(l+1) above rmax; summed
// over l, m with GSL's unnormalised P_l^m(cos theta) (Condon-Shortley phase)
// and cos/sin m phi from the Chebyshev recurrence.  In double (not
// op-counted).
struct MpCoef {
  const double* a;
  int Nr, L, M, nonaxi, Nt;
  const double *rg, *tg, *blocks;
  double amp;
  long per;
};
inline MpCoef mp_view(const Component& C) {
  MpCoef v;
  v.a = C.args;
  v.Nr = (int)v.a[0];
  v.L = (int)v.a[1];
  v.M = (int)v.a[2];
  v.nonaxi = (int)v.a[3];
  v.rg = v.a + 4;
  v.amp = v.rg[v.Nr];
  v.Nt = (int)v.rg[v.Nr + 1];
  v.tg = v.rg + v.Nr + 2;
  v.blocks = v.tg + v.Nt;
  const long nr = 6L * (v.Nr - 1);
  v.per = v.Nt == 0 ? 2 * nr + v.Nr : 4L * (v.Nt - 1) * (2 * nr + v.Nr);
  return v;
}
// ppoly_find_interval (:97-124) without its cache: the i with x[i] <= v <
// x[i+1], clamped to [0, n-2]
inline int mp_find(const double* x, int n, double v) {
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (x[mid] <= v) lo = mid; else hi = mid - 1;
  }
  return lo;
}
// the six coefficients of interval ir of one PPoly (inner: half 0, outer: 1)
// of block k, interpolated in time when time-dependent
inline void mp_coeffs(const MpCoef& v, int k, int half, int ir, int it, double dt, double* c) {
  const long nr = 6L * (v.Nr - 1);
  const double* b = v.blocks + k * v.per;
  if (v.Nt == 0) {
    const double* q = b + half * nr + 6L * ir;
    for (int p = 0; p < 6; p++) c[p] = q[p];
    return;
  }
  const double* q = b + half * (long)(v.Nt - 1) * nr * 4 + (long)it * nr * 4 + 6L * ir * 4;
  for (int p = 0; p < 6; p++) {
    const double* cp = q + p * 4;
    c[p] = ((cp[0] * dt + cp[1]) * dt + cp[2]) * dt + cp[3];
  }
}
inline double mp_horner(const double* c, double d) {
  return ((((c[0] * d + c[1]) * d + c[2]) * d + c[3]) * d + c[4]) * d + c[5];
}
inline double mp_horner_d(const double* c, double d) {
  return (((5 * c[0] * d + 4 * c[1]) * d + 3 * c[2]) * d + 2 * c[3]) * d + c[4];
}
// R_lm and dR_lm/dr of block k (eval_radial_lm / eval_radial_lm_timedep)
inline double mp_radial(const MpCoef& v, int k, double r, int l, double r_l, double r_nl, int it,
                        double dt, double& dR) {
  const double rmin = v.rg[0], rmax = v.rg[v.Nr - 1];
  double ci[6], co[6];
  if (r < rmin) {
    mp_coeffs(v, k, 0, 0, it, dt, ci);
    mp_coeffs(v, k, 1, 0, it, dt, co);
    const double P0 = mp_horner_d(ci, 0.0) / std::pow(rmin, l + 2);
    const double Iie = P0 / (l + 3) * std::pow(r, l + 3);
    const double extra = (l == 2) ? P0 * std::log(rmin / r)
                                  : P0 / (2 - l) * (std::pow(rmin, 2 - l) - std::pow(r, 2 - l));
    const double Ioe = mp_horner(co, 0.0) + extra;
    const double a = std::pow(r, -(l + 1)), b = std::pow(r, l);
    dR = -(l + 1) * a / r * Iie + l * b / r * Ioe;
    return a * Iie + b * Ioe;
  }
  if (r <= rmax) {
    const int ir = mp_find(v.rg, v.Nr, r);
    const double d = r - v.rg[ir];
    mp_coeffs(v, k, 0, ir, it, dt, ci);
    mp_coeffs(v, k, 1, ir, it, dt, co);
    const double Ii = mp_horner(ci, d), Io = mp_horner(co, d);
    dR = -(l + 1) * r_nl / r * Ii + r_nl * mp_horner_d(ci, d) + l * r_l / r * Io + r_l * mp_horner_d(co, d);
    return r_nl * Ii + r_l * Io;
  }
  const int ir = v.Nr - 2;
  mp_coeffs(v, k, 0, ir, it, dt, ci);
  const double Ii = mp_horner(ci, rmax - v.rg[ir]);
  dR = (-(l + 1)) * Ii * r_nl / r;
  return Ii * r_nl;
}
// F = -amp (dPhi/dr, dPhi/dtheta, dPhi/dphi) (compute_multipole_spher_forces)
// and the potential (MultipoleExpansionPotentialEval)
inline void mp_spher(const Component& C, double Rc, double z, double phi, double t, double* F, double& pot) {
  const MpCoef v = mp_view(C);
  F[0] = F[1] = F[2] = 0.;
  pot = 0.;
  const double r = std::sqrt(Rc * Rc + z * z), theta = std::atan2(Rc, z);
  if (r == 0.0 || !std::isfinite(r)) return;
  const double ct = std::cos(theta), st = std::sin(theta);
  const int L = v.L, M = v.M, Meff = v.nonaxi ? M : 1;
  std::vector<double> P(std::max(Meff, 1) * L, 0.), dP(std::max(Meff, 1) * L, 0.);
  scf_legendre(ct, L, Meff, P, dP);
  std::vector<double> rl(L), rnl(L), cm(M), sm(M);
  double rp = 1.0, rnp = 1.0 / r;
  for (int l = 0; l < L; l++) {
    rl[l] = rp;
    rnl[l] = rnp;
    rp *= r;
    rnp /= r;
  }
  cm[0] = 1.0;
  sm[0] = 0.0;
  if (M > 1) {
    cm[1] = std::cos(phi);
    sm[1] = std::sin(phi);
    for (int m = 2; m < M; m++) {
      cm[m] = 2.0 * cm[1] * cm[m - 1] - cm[m - 2];
      sm[m] = 2.0 * cm[1] * sm[m - 1] - sm[m - 2];
    }
  }
  int it = 0;
  double dt = 0.;
  if (v.Nt > 0) {
    it = mp_find(v.tg, v.Nt, t);
    dt = t - v.tg[it];
  }
  double dr = 0., dth = 0., dph = 0.;
  int k = 0;
  for (int l = 0; l < L; l++) {
    for (int m = 0; m < std::min(l + 1, M); m++) {
      const double Pl = P[m * L + l], dPl = dP[m * L + l];
      double dRc, dRs = 0.;
      const double Rcv = mp_radial(v, k++, r, l, rl[l], rnl[l], it, dt, dRc);
      dr += Pl * cm[m] * dRc;
      dth += dPl * (-st) * cm[m] * Rcv;
      dph += Pl * (-m * sm[m]) * Rcv;
      double rad = Rcv * cm[m];
      if (m > 0) {
        const double Rsv = mp_radial(v, k++, r, l, rl[l], rnl[l], it, dt, dRs);
        dr += Pl * sm[m] * dRs;
        dth += dPl * (-st) * sm[m] * Rsv;
        dph += Pl * (m * cm[m]) * Rsv;
        rad += Rsv * sm[m];
      }
      pot += Pl * rad;
    }
  }
  F[0] = -v.amp * dr;
  F[1] = -v.amp * dth;
  F[2] = -v.amp * dph;
  pot *= v.amp;
}
template <class R> void mp_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fp) {
  double F[3], pot;
  mp_spher(C, dval(Rc), dval(z), dval(phi), dval(t), F, pot);
  const double Rd = dval(Rc), zd = dval(z), r = std::sqrt(Rd * Rd + zd * zd);
  if (r == 0.0 || !std::isfinite(r)) {
    FR = R(0.);
    Fz = R(0.);
    Fp = R(F[2]);
    return;
  }
  FR = R(F[0] * (Rd / r) + F[1] * (zd / (r * r)));
  Fz = R(F[0] * (zd / r) + F[1] * (-Rd / (r * r)));
  Fp = R(F[2]);
}
template <class R> R mp_eval(const Component& C, R Rc, R z, R phi, R t) {
  double F[3], pot;
  mp_spher(C, dval(Rc), dval(z), dval(phi), dval(t), F, pot);
  return R(pot);
}

// DiskSCFPotential.c:33-120 (one [Sigma_i, h_i] approximation pair of a
// DiskSCFPotential): Phi = Sigma(r) H(z), Sigma exponential (type 0: amp
// exp(-r/h)) or with a central hole (type 1: amp exp(-Rhole/r - r/h)), H
// the integral twice over of the vertical profile (type 0 exponential,
// 1 sech^2).  args: nsigma_args, the Sigma arguments, hz type, hz h.
template <class R> void dscf_sigma(const double* sa, R r, R& S, R& dS) {
  if ((int)sa[0] == 0) {
    S = R(sa[1]) * exp(-r / R(sa[2]));
    dS = -R(sa[1]) * exp(-r / R(sa[2])) / R(sa[2]);
  } else {
    S = R(sa[1]) * exp(-R(sa[3]) / r - r / R(sa[2]));
    dS = R(sa[1]) * (R(sa[3]) / r / r - R(1.) / R(sa[2])) * exp(-R(sa[3]) / r - r / R(sa[2]));
  }
}
template <class R> void dscf_hz(const double* ha, R z, R& H, R& dH) {
  R fz = fabs(z);
  const double h = ha[1];
  if ((int)ha[0] == 0) {
    H = R(0.5) * (exp(-fz / R(h)) - R(1.) + fz / R(h)) * R(h);
    dH = (z >= R(0.) ? R(1.) : R(-1.)) * R(0.5) * (R(1.) - exp(-fz / R(h)));  // copysign(., z)
  } else {
    H = R(h) * (log(R(1.) + exp(-fz / R(h))) + R(0.5) * fz / R(h) - R(M_LN2));
    dH = R(0.5) * tanh(R(0.5) * z / R(h));
  }
}
template <class R> void dscf_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  const double* a = C.args;
  const int ns = (int)a[0];
  R r = sqrt(Rc * Rc + z * z), S, dS, H, dH;
  dscf_sigma(a + 1, r, S, dS);
  dscf_hz(a + 1 + ns, z, H, dH);
  FR = -dS * H * Rc / r;
  Fz = -dS * H * z / r - S * dH;
}

// interpRZPotential.c:6-70 with interp_2d_eval_cubic_bspline
// (util/interp_2d/interp_2d.c:92-114) and Thevenaz's cubic B-spline
// interpolation with mirror boundaries (cubic_bspline_2d_interpol.c): the
// point clamped to the grid, its fractional grid index from the bracketing
// interval (binary search), the 4 x 4 weighted coefficients.
inline int interprz_find(const double* x, int n, double v) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (x[mid] > v) hi = mid; else lo = mid;
  }
  return lo;
}
inline double bspline2d(const double* coeffs, long width, long height, double x, double y) {
  long xi[4], yi[4];
  double xw[4], yw[4];
  const long width2 = 2L * width - 2L, height2 = 2L * height - 2L;
  long i = (long)std::floor(x) - 1L, j = (long)std::floor(y) - 1L;
  for (int k = 0; k < 4; k++) { xi[k] = i++; yi[k] = j++; }
  double w = x - (double)xi[1];
  xw[3] = (1.0 / 6.0) * w * w * w;
  xw[0] = (1.0 / 6.0) + (1.0 / 2.0) * w * (w - 1.0) - xw[3];
  xw[2] = w + xw[0] - 2.0 * xw[3];
  xw[1] = 1.0 - xw[0] - xw[2] - xw[3];
  w = y - (double)yi[1];
  yw[3] = (1.0 / 6.0) * w * w * w;
  yw[0] = (1.0 / 6.0) + (1.0 / 2.0) * w * (w - 1.0) - yw[3];
  yw[2] = w + yw[0] - 2.0 * yw[3];
  yw[1] = 1.0 - yw[0] - yw[2] - yw[3];
  for (int k = 0; k < 4; k++) {
    xi[k] = (width == 1L) ? 0L : ((xi[k] < 0L) ? (-xi[k] - width2 * ((-xi[k]) / width2)) : (xi[k] - width2 * (xi[k] / width2)));
    if (width <= xi[k]) xi[k] = width2 - xi[k];
    yi[k] = (height == 1L) ? 0L : ((yi[k] < 0L) ? (-yi[k] - height2 * ((-yi[k]) / height2)) : (yi[k] - height2 * (yi[k] / height2)));
    if (height <= yi[k]) yi[k] = height2 - yi[k];
  }
  double out = 0.0;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) out += coeffs[xi[a] * height + yi[b]] * xw[a] * yw[b];
  return out;
}
// grid: 0 potential, 1 R force, 2 z force
inline double interprz_eval(const Component& C, int grid, double R, double z) {
  const double* a = C.args;
  const int nR = (int)a[0], nz = (int)a[1];
  const double* Rg = a + 5;
  const double* zg = Rg + nR;
  const double* co = zg + nz + (size_t)grid * nR * nz;
  const double* tail = zg + nz + (size_t)(3 + (int)a[2] + (int)a[3] + (int)a[4]) * nR * nz;
  const int logR = (int)tail[1];
  double x = logR == 1 ? (R > 0. ? std::log(R) : -20.72326583694641) : R;
  double y = std::fabs(z);
  x = x > Rg[nR - 1] ? Rg[nR - 1] : x;
  x = x < Rg[0] ? Rg[0] : x;
  y = y > zg[nz - 1] ? zg[nz - 1] : y;
  y = y < zg[0] ? zg[0] : y;
  const int ix = interprz_find(Rg, nR, x), iy = interprz_find(zg, nz, y);
  const double xn = ix + (x - Rg[ix]) / (Rg[ix + 1] - Rg[ix]);
  const double yn = iy + (y - zg[iy]) / (zg[iy + 1] - zg[iy]);
  return tail[0] * bspline2d(co, nR, nz, xn, yn);
}

// KuzminLikeWrapperPotential.c:5-8: xi = sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)
template <class R> R kuzminlike_xi(R Rc, R z, double a, double b2) {
  R asqrtbz = R(a) + sqrt(z * z + R(b2));
  return sqrt(Rc * Rc + asqrtbz * asqrtbz);
}
// RotateAndTiltWrapperPotential.c:21-43: the point in the aligned frame,
// rot . x + offset (own arguments: amp, 7 cache slots, rot[9] at 8,
// rotSet 17, offsetSet 18, offset 19-21)
template <class R> void rottilt_point(const Component& C, R Rc, R z, R phi, R& x, R& y, R& zz) {
  const double* a = C.args;
  const double* rot = a + 8;
  x = Rc * cos(phi);
  y = Rc * sin(phi);
  zz = z;
  if (a[17] != 0.) {  // rotate, galpy_potentials.c:391-399
    R xp = R(rot[0]) * x + R(rot[1]) * y + R(rot[2]) * zz;
    R yp = R(rot[3]) * x + R(rot[4]) * y + R(rot[5]) * zz;
    R zp = R(rot[6]) * x + R(rot[7]) * y + R(rot[8]) * zz;
    x = xp; y = yp; zz = zp;
  }
  if (a[18] != 0.) {
    x += R(a[19]); y += R(a[20]); zz += R(a[21]);
  }
}
template <class R>
void cyl_forces(int npot, const Component* P, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fphi);
// RotateAndTiltWrapperPotentialxyzforces (:21-62) and its Rforce, phitorque,
// zforce (:80-106): the wrapped forces at the aligned-frame point, to
// Cartesian there, rotated back by rot^T (rotate_force, galpy_potentials.c:400-409)
template <class R> void rottilt_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fp) {
  const double* a = C.args;
  const double* rot = a + 8;
  R x, y, zz;
  rottilt_point(C, Rc, z, phi, x, y, zz);
  R Rp = sqrt(x * x + y * y), php = atan2(y, x);
  R wr, wz, wp;
  cyl_forces(C.nwrapped, C.wrapped, Rp, zz, php, t, wr, wz, wp);
  R Fx = cos(php) * wr - sin(php) * wp / Rp;
  R Fy = sin(php) * wr + cos(php) * wp / Rp;
  R Fzz = wz;
  if (a[17] != 0.) {
    R Fxp = R(rot[0]) * Fx + R(rot[3]) * Fy + R(rot[6]) * Fzz;
    R Fyp = R(rot[1]) * Fx + R(rot[4]) * Fy + R(rot[7]) * Fzz;
    R Fzp = R(rot[2]) * Fx + R(rot[5]) * Fy + R(rot[8]) * Fzz;
    Fx = Fxp; Fy = Fyp; Fzz = Fzp;
  }
  FR = R(a[0]) * (cos(phi) * Fx + sin(phi) * Fy);
  Fp = R(a[0]) * Rc * (-sin(phi) * Fx + cos(phi) * Fy);
  Fz = R(a[0]) * Fzz;
}

template <class R>
void cyl_forces(int npot, const Component* P, R Rc, R z, R phi, R t,
                R& FR, R& Fz, R& Fphi) {
  FR = 0.; Fz = 0.; Fphi = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    R fr = 0., fz = 0., fp = 0.;
    switch (C.type) {
      case kPowerSphCutoff: {
        R r2 = Rc * Rc + z * z;
        R f = psc_force_factor(C, r2);
        fr = f * Rc; fz = f * z;
        break;
      }
      case kMiyamotoNagai: mn_forces(C, Rc, z, fr, fz); break;
      case kNFW: nfw_forces(C, Rc, z, fr, fz); break;
      case kHernquist: hq_forces(C, Rc, z, fr, fz); break;
      case kLogHalo: lh_forces(C, Rc, z, phi, fr, fz, fp); break;
      case kDehnenBar: bar_forces(C, Rc, z, phi, t, fr, fz, fp); break;
      case kKuzminDisk: kuzmin_forces(C, Rc, z, fr, fz); break;
      case kKKStaeckel: kks_forces(C, Rc, z, fr, fz); break;
      case kSoftNeedleBar: needle_forces(C, Rc, z, phi, t, fr, fz, fp); break;
      case kSpiralArms: spiral_forces(C, Rc, z, phi, t, fr, fz, fp); break;
      case kDoubleExpDisk: dedisk_forces(C, Rc, z, fr, fz); break;
      case kSCF: scf_forces(C, Rc, z, phi, t, fr, fz, fp); break;
      case kMultipole: mp_forces(C, Rc, z, phi, t, fr, fz, fp); break;
      case kDiskSCF: dscf_forces(C, Rc, z, fr, fz); break;
      case kInterpRZ: {  // interpRZPotential.c:22-70
        fr = R(interprz_eval(C, 1, dval(Rc), dval(z)));
        const double fzv = interprz_eval(C, 2, dval(Rc), dval(z));
        fz = R(dval(z) < 0. ? -fzv : fzv);
        break;
      }
      case kInterpSph: {
        R r = sqrt(Rc * Rc + z * z);
        R f = R(C.p[0]) * interpsph_rforce(C, r);
        fr = f * Rc / r;
        fz = f * z / r;
        break;
      }
      case kFlattenedPower: {  // FlattenedPowerPotential.c Rforce, zforce
        double amp = C.p[0], alpha = C.p[1], q2 = C.p[2], core2 = C.p[3];
        if (alpha == 0.) {
          fr = -R(amp) * Rc / (Rc * Rc + z * z / R(q2) + R(core2));
          fz = -R(amp) * z / R(q2) / (Rc * Rc + z * z / R(q2) + R(core2));
        } else {
          R m2 = R(core2) + Rc * Rc + z * z / R(q2);
          fr = -R(amp) * pow(m2, R(-0.5 * alpha - 1.)) * Rc;
          fz = -R(amp) * pow(m2, R(-0.5 * alpha - 1.)) * z / R(q2);
        }
        break;
      }
      case kNull: break;
      default:
        if (is_ellipsoidal(C.type)) {
          ellip_forces(C, Rc, z, phi, fr, fz, fp);
        } else if (is_spherical_new(C.type)) {
          sph_forces(C, Rc, z, fr, fz);
        } else if (C.type == kMovingObjectWrap) {
          R ox, oy, oz;
          movingobject_pos(C, t, ox, oy, oz);
          R x = Rc * cos(phi), y = Rc * sin(phi);
          R Rdist = pow(pow(x - ox, R(2)) + pow(y - oy, R(2)), R(0.5));
          R wr, wz, wp;
          cyl_forces(C.nwrapped, C.wrapped, Rdist, oz - z, phi, t, wr, wz, wp);
          fr = -R(C.p[0]) * wr * (cos(phi) * (ox - x) + sin(phi) * (oy - y)) / Rdist;
          fz = -R(C.p[0]) * wz;
          fp = -R(C.p[0]) * wr * Rc * (cos(phi) * (oy - y) - sin(phi) * (ox - x)) / Rdist;
        } else if (C.type == kRotateTiltWrap) {
          rottilt_forces(C, Rc, z, phi, t, fr, fz, fp);
        } else if (C.type == kKuzminLikeWrap) {
          // KuzminLikeWrapperPotential.c:58-92: the wrapped radial force at
          // (xi, 0), phi = 0, times dxi/dR and dxi/dz
          const double amp = C.p[0], a = C.p[1], b2 = C.p[2];
          R xi = kuzminlike_xi(Rc, z, a, b2);
          R wr, wz, wp;
          cyl_forces(C.nwrapped, C.wrapped, xi, R(0.), R(0.), t, wr, wz, wp);
          R sqrtbz = sqrt(z * z + R(b2));
          fr = R(amp) * wr * (Rc / xi);
          fz = R(amp) * wr * ((R(a) + sqrtbz) * z / xi / sqrtbz);
        } else if (C.type == kOblateStaeckelWrap) {
          double fr_, fz_;
          os_forces(C, dval(Rc), dval(z), fr_, fz_);
          fr = R(fr_);
          fz = R(fz_);
        } else if (C.type == kCylSepWrap) {
          // CylindricallySeparablePotentialWrapper.c:14-27: the wrapped
          // radial force at (R, 0) and vertical force at (Rp, z), phi = t = 0
          const double amp = C.p[0], Rp = C.p[1];
          R wr, wz, wp;
          cyl_forces(C.nwrapped, C.wrapped, Rc, R(0.), R(0.), R(0.), wr, wz, wp);
          fr = R(amp) * wr;
          cyl_forces(C.nwrapped, C.wrapped, R(Rp), z, R(0.), R(0.), wr, wz, wp);
          fz = R(amp) * wz;
        } else if (is_wrapper(C.type)) {
          // calcRforce/calczforce/calcphitorque of the wrapped components at
          // phi - offset (the wrappers' Rforce/zforce/phitorque functions)
          R fac, off, doff, wr, wz, wp;
          wrapper_transform(C, Rc, t, fac, off, doff);
          cyl_forces(C.nwrapped, C.wrapped, Rc, z, phi - off, t, wr, wz, wp);
          fr = fac * (wr - wp * doff);
          fz = fac * wz;
          fp = fac * wp;
        }
    }
    FR += fr; Fz += fz; Fphi += fp;
  }
}

// OblateStaeckelWrapperPotential.c:4-229: the wrapped potential put in
// Staeckel form, Phi = amp (U(u) - V(v)) / (sinh^2 u + sin^2 v), in prolate
// spheroidal coordinates of focal length delta, with U(u) = cosh^2 u
// Phi_w(u, v0) and V(v) = refpot - (sinh^2 u0 + sin^2 v) Phi_w(u0, v) from the
// wrapped potential and forces at phi = t = 0 (U, dUdu, V, dVdv at :31-53,
// :114-137; the planar force at :215-229 uses the wrapped planar force and
// no V).  args: amp, delta, u0, v0, refpot.  In double (not op-counted).
template <class R> void cyl_forces(int npot, const Component* P, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fphi);
template <class R> void planar_forces(int npot, const Component* P, R Rc, R phi, R t, R& FR, R& Fphi);
template <class R> R potential_value(int npot, const Component* P, R Rc, R z, R phi, R t);
inline void os_Rz_to_uv(double R, double z, double delta, double& u, double& v) {
  const double d12 = std::pow(z + delta, 2) + std::pow(R, 2);
  const double d22 = std::pow(z - delta, 2) + std::pow(R, 2);
  const double coshu = 0.5 / delta * (std::sqrt(d12) + std::sqrt(d22));
  const double cosv = 0.5 / delta * (std::sqrt(d12) - std::sqrt(d22));
  u = std::acosh(coshu);
  v = std::acos(cosv);
}
inline void os_UV(const Component& C, double u, double v, bool planar, double& U, double& dU, double& V,
                  double& dV) {
  const double delta = C.p[1], u0 = C.p[2], v0 = C.p[3], refpot = C.p[4];
  const int n = C.nwrapped;
  const Component* W = C.wrapped;
  double fr, fz, fp;
  // U and dUdu on the curve v = v0
  const double Ra = delta * std::sinh(u) * std::sin(v0), za = delta * std::cosh(u) * std::cos(v0);
  const double pa = potential_value(n, W, Ra, za, 0., 0.);
  U = std::pow(std::cosh(u), 2) * pa;
  if (planar) {
    planar_forces(n, W, Ra, 0., 0., fr, fp);
    dU = 2 * std::cosh(u) * std::sinh(u) * pa - std::pow(std::cosh(u), 2) * fr * Ra / (std::tanh(u) + 1e-12);
    V = dV = 0.;
    return;
  }
  cyl_forces(n, W, Ra, za, 0., 0., fr, fz, fp);
  dU = 2 * std::cosh(u) * std::sinh(u) * pa -
       std::pow(std::cosh(u), 2) * (fr * Ra / (std::tanh(u) + 1e-12) + fz * za * std::tanh(u));
  // V and dVdv on the curve u = u0
  const double Rb = delta * std::sinh(u0) * std::sin(v), zb = delta * std::cosh(u0) * std::cos(v);
  const double pb = potential_value(n, W, Rb, zb, 0., 0.);
  const double pre = std::pow(std::sinh(u0), 2) + std::pow(std::sin(v), 2);
  V = refpot - pre * pb;
  cyl_forces(n, W, Rb, zb, 0., 0., fr, fz, fp);
  dV = -2 * std::sin(v) * std::cos(v) * pb + pre * (fr * Rb / std::tan(v) - fz * zb * std::tan(v));
}
inline void os_forces(const Component& C, double Rc, double z, double& FR, double& Fz) {
  const double amp = C.p[0], delta = C.p[1];
  double u, v, U, dU, V, dV;
  os_Rz_to_uv(Rc, z, delta, u, v);
  os_UV(C, u, v, false, U, dU, V, dV);
  const double pre = std::pow(std::sinh(u), 2) + std::pow(std::sin(v), 2);
  const double dpu = 2 * std::sinh(u) * std::cosh(u), dpv = 2 * std::sin(v) * std::cos(v);
  FR = amp * ((-dU * delta * std::sin(v) * std::cosh(u) + dV * std::tanh(u) * z +
               (U - V) * (dpu * delta * std::sin(v) * std::cosh(u) + dpv * std::tanh(u) * z) / pre) /
              std::pow(delta * pre, 2));
  Fz = amp * ((-dU * Rc / std::tan(v) - dV * delta * std::sin(v) * std::cosh(u) +
               (U - V) * (dpu / std::tan(v) * Rc - dpv * delta * std::sin(v) * std::cosh(u)) / pre) /
              std::pow(delta * pre, 2));
}
inline double os_planar_Rforce(const Component& C, double Rc) {
  const double amp = C.p[0], delta = C.p[1];
  double u, v, U, dU, V, dV;
  os_Rz_to_uv(Rc, 0., delta, u, v);
  os_UV(C, u, v, true, U, dU, V, dV);
  const double pre = std::pow(std::sinh(u), 2) + std::pow(std::sin(v), 2);
  const double dpu = 2 * std::sinh(u) * std::cosh(u);
  return amp * ((-dU * delta * std::sin(v) * std::cosh(u) + U * dpu * delta * std::sin(v) * std::cosh(u) / pre) /
                std::pow(delta * pre, 2));
}
inline double os_eval(const Component& C, double Rc, double z) {
  double u, v, U, dU, V, dV;
  os_Rz_to_uv(Rc, z, C.p[1], u, v);
  os_UV(C, u, v, false, U, dU, V, dV);
  return C.p[0] * (U - V) / (std::pow(std::sinh(u), 2) + std::pow(std::sin(v), 2));
}

// DoubleExponentialDiskPotential.c: Ogata's double-exponential quadrature of
// the disk's Hankel integrals over packed J0/J1 nodes and weights, each sum
// stopped once a term is below 1e-15 of the running sum or the nodes run
// out: Eval (:6-34, J0), Rforce (:35-62, J1), PlanarRforce (:63-86, J1,
// the z = 0 form 1/(beta + x)), zforce (:87-120, J0).  args[1] is
// -4 pi alpha amp.
template <class R> void dedisk_forces(const Component& C, R Rc, R z, R& FR, R& Fz) {
  const double* a = C.args;
  const double amp = a[1], alpha = a[2], beta = a[3];
  const int n = (int)a[4];
  const double *j0x = a + 5, *j1x = a + 5 + n, *j0w = a + 5 + 2 * n, *j1w = a + 5 + 3 * n;
  R alpha2 = R(alpha) * R(alpha), beta2 = R(beta) * R(beta);
  R fz = fabs(z);
  R ebetafz = exp(-R(beta) * fz);
  R out = 0., prev = 1.;
  for (int ii = 0; fabs(prev) > R(1e-15) && ii < n; ii++) {
    R x = R(j1x[ii]) / Rc;
    prev = R(j1w[ii]) * x * pow(alpha2 + x * x, R(-1.5)) * (R(beta) * exp(-x * fz) - x * ebetafz) / (beta2 - x * x);
    out += prev;
    prev /= out;
  }
  FR = R(amp) * out / Rc;
  out = 0.;
  prev = 1.;
  for (int ii = 0; fabs(prev) > R(1e-15) && ii < n; ii++) {
    R x = R(j0x[ii]) / Rc;
    prev = R(j0w[ii]) * pow(alpha2 + x * x, R(-1.5)) * x * (exp(-x * fz) - ebetafz) / (beta2 - x * x);
    out += prev;
    prev /= out;
  }
  Fz = (z > R(0.)) ? R(amp) * out * R(beta) / Rc : -R(amp) * out * R(beta) / Rc;
}
template <class R> R dedisk_planar_Rforce(const Component& C, R Rc) {
  const double* a = C.args;
  const double amp = a[1], alpha = a[2], beta = a[3];
  const int n = (int)a[4];
  const double *j1x = a + 5 + n, *j1w = a + 5 + 3 * n;
  R alpha2 = R(alpha) * R(alpha);
  R out = 0., prev = 1.;
  for (int ii = 0; fabs(prev) > R(1e-15) && ii < n; ii++) {
    R x = R(j1x[ii]) / Rc;
    prev = R(j1w[ii]) * x * pow(alpha2 + x * x, R(-1.5)) / (R(beta) + x);
    out += prev;
    prev /= out;
  }
  return R(amp) * out / Rc;
}
template <class R> R dedisk_eval(const Component& C, R Rc, R z) {
  const double* a = C.args;
  const double amp = a[1], alpha = a[2], beta = a[3];
  const int n = (int)a[4];
  const double *j0x = a + 5, *j0w = a + 5 + 2 * n;
  R alpha2 = R(alpha) * R(alpha), beta2 = R(beta) * R(beta);
  R fz = fabs(z);
  R ebetafz = exp(-R(beta) * fz);
  R out = 0., prev = 1.;
  for (int ii = 0; fabs(prev) > R(1e-15) && ii < n; ii++) {
    R x = R(j0x[ii]) / Rc;
    prev = R(j0w[ii]) * pow(alpha2 + x * x, R(-1.5)) * (R(beta) * exp(-x * fz) - x * ebetafz) / (beta2 - x * x);
    out += prev;
    prev /= out;
  }
  return R(amp) * out / Rc;
}

// SpiralArmsPotential.c (Cox & Gomez 2002 arms): the helpers gam, dgam_dR,
// K, B, D, dK_dR, dB_dR, dD_dR (:904-940) and the forces Rforce (:61-126),
// zforce (:128-170), phitorque (:172-210), Eval (:21-58), in the reference's
// order.  The planar functions (:625-716) are these at z = 0.
template <class R> R spa_K(R Rc, double n, double N, double sa) { return R(n * N) / Rc / R(sa); }
template <class R> R spa_B(R Rc, double H, double n, double N, double sa) {
  R HNn_R = R(H * N * n) / Rc;
  return HNn_R / R(sa) * (R(0.4) * HNn_R / R(sa) + R(1.));
}
template <class R> R spa_D(R Rc, double H, double n, double N, double sa) {
  double HNn = H * N * n;
  return (R(0.3 * HNn * HNn / sa) / Rc + R(HNn) + Rc * R(sa)) / (R(0.3 * HNn) + Rc * R(sa));
}
template <class R> R spa_dK(R Rc, double n, double N, double sa) { return R(-n * N) / (Rc * Rc) / R(sa); }
template <class R> R spa_dB(R Rc, double H, double n, double N, double sa) {
  double HNn = H * N * n;
  return R(-HNn) / Rc / Rc / Rc / R(sa) / R(sa) * (R(0.8 * HNn) + Rc * R(sa));
}
template <class R> R spa_dD(R Rc, double H, double n, double N, double sa) {
  R k = R(H * N * n) / Rc / R(sa);
  return k * (R(0.3) * (k + R(0.3) * k * k + R(1.)) / Rc / pow(R(0.3) * k + R(1.), R(2.)) -
              (R(1.) / Rc * (R(1.) + R(0.6) * k) / (R(0.3) * k + R(1.))));
}
template <class R> void spiral_forces(const Component& C, R Rc, R z, R phi, R t, R& FR, R& Fz, R& Fphi) {
  const double* a = C.args;
  const int nCs = (int)a[0];
  const double amp = a[1], N = a[2], sa = a[3], ta = a[4], r_ref = a[5], phi_ref = a[6], Rs = a[7], H = a[8],
               omega = a[9];
  R g = R(N) * ((phi - R(omega) * t) - R(phi_ref) - log(Rc / R(r_ref)) / R(ta));
  R dg = R(-N) / Rc / R(ta);
  R sr = 0., sz = 0., sp = 0.;
  for (int k = 1; k <= nCs; k++) {
    const double n = k, Cn = a[9 + k];
    R Kn = spa_K(Rc, n, N, sa), Bn = spa_B(Rc, H, n, N, sa), Dn = spa_D(Rc, H, n, N, sa);
    R dKn = spa_dK(Rc, n, N, sa), dBn = spa_dB(Rc, H, n, N, sa), dDn = spa_dD(Rc, H, n, N, sa);
    R cng = cos(R(n) * g), sng = sin(R(n) * g);
    R zKB = z * Kn / Bn;
    R sech = R(1.) / cosh(zKB);
    sr += R(Cn) * pow(sech, Bn) / Dn *
          ((R(n) * dg / Kn * sng + cng * (z * tanh(zKB) * (dKn / Kn - dBn / Bn) - dBn / Kn * log(sech) +
                                          dKn / Kn / Kn + dDn / Dn / Kn)) +
           cng / Kn / R(Rs));
    sz += R(Cn) / Dn * cng * tanh(zKB) / pow(cosh(zKB), Bn);
    sp += R(N * n * Cn) / Dn / Kn / pow(cosh(z * Kn / Bn), Bn) * sng;
  }
  R pre = R(-amp * H) * exp(-(Rc - R(r_ref)) / R(Rs));
  FR = pre * sr;
  Fz = pre * sz;
  Fphi = pre * sp;
}
template <class R> R spiral_eval(const Component& C, R Rc, R z, R phi, R t) {
  const double* a = C.args;
  const int nCs = (int)a[0];
  const double amp = a[1], N = a[2], sa = a[3], ta = a[4], r_ref = a[5], phi_ref = a[6], Rs = a[7], H = a[8],
               omega = a[9];
  R g = R(N) * ((phi - R(omega) * t) - R(phi_ref) - log(Rc / R(r_ref)) / R(ta));
  R sum = 0.;
  for (int k = 1; k <= nCs; k++) {
    const double n = k, Cn = a[9 + k];
    R Kn = spa_K(Rc, n, N, sa), Bn = spa_B(Rc, H, n, N, sa), Dn = spa_D(Rc, H, n, N, sa);
    sum += R(Cn) / Kn / Dn * cos(R(n) * g) / pow(cosh(Kn * z / Bn), Bn);
  }
  return R(-amp * H) * exp(-(Rc - R(r_ref)) / R(Rs)) * sum;
}

// Second derivatives d2Phi/dR2, d2Phi/dz2, d2Phi/dRdz of the axisymmetric
// components (the phi derivatives vanish): PowerSphericalPotentialwCutoff.c
// R2deriv/z2deriv/Rzderiv (Phi'' and Phi' of r), MiyamotoNagaiPotential.c:66-119,
// NFWPotential.c:63-107, HernquistPotential.c (spherical Phi', Phi''),
// LogarithmicHaloPotential.c (axisymmetric branch).  Returns false for a
// component without a restated Hessian (the bar, triaxial haloes).
template <class R>
bool cyl_hessian(int npot, const Component* P, R Rc, R z, R& R2, R& z2, R& Rz) {
  R2 = 0.; z2 = 0.; Rz = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    R r2 = Rc * Rc + z * z;
    R r = sqrt(r2);
    switch (C.type) {
      case kPowerSphCutoff: {
        double amp = C.p[0], alpha = C.p[1], rc = C.p[2];
        R x = r2 / R(rc) / R(rc);
        R mass = R(C.c[2]) * lower_incgamma(C.c[0], x, C.c[1]);
        R Phipp = R(amp) * (R(4. * M_PI) * pow(r2, R(-0.5 * alpha)) * exp(-r2 / R(rc) / R(rc)) -
                            R(2.) * mass / pow(r2, R(1.5)));
        R Phip = R(amp) * mass / r2;
        R ir2 = R(1.) / r2, ir3 = ir2 / r;
        R2 += Phipp * Rc * Rc * ir2 + Phip * z * z * ir3;
        z2 += Phipp * z * z * ir2 + Phip * Rc * Rc * ir3;
        Rz += Rc * z * (Phipp * ir2 - Phip * ir3);
        break;
      }
      case kMiyamotoNagai: {
        double amp = C.p[0], a = C.p[1], b = C.p[2];
        R b2 = R(b * b);
        R sqrtbz = pow(b2 + z * z, R(0.5));
        R asqrtbz = R(a) + sqrtbz;
        R denom = Rc * Rc + asqrtbz * asqrtbz;
        R2 += R(amp) * (pow(denom, R(-1.5)) - R(3.) * Rc * Rc * pow(denom, R(-2.5)));
        if (a == 0.) {
          z2 += R(amp) * (b2 + Rc * Rc - R(2.) * z * z) * pow(b2 + Rc * Rc + z * z, R(-2.5));
          Rz += R(amp) * (R(-3.) * Rc * z * pow(denom, R(-2.5)));
        } else {
          R z4 = z * z * z * z;
          z2 += R(amp) * (R(a * a * a) * b2 + R(a * a) * (R(3.) * b2 - R(2.) * z * z) * sqrtbz +
                          (b2 + Rc * Rc - R(2.) * z * z) * pow(b2 + z * z, R(1.5)) +
                          R(a) * (R(3.) * b2 * b2 - R(4.) * z4 + b2 * (Rc * Rc - z * z))) /
                (pow(b2 + z * z, R(1.5)) * pow(denom, R(2.5)));
          Rz += R(amp) * (R(-3.) * Rc * z * asqrtbz / sqrtbz * pow(denom, R(-2.5)));
        }
        break;
      }
      case kNFW: {
        double amp = C.p[0], a = C.p[1];
        R dphi = log(R(1.) + r / R(a)) / r / r - R(1.) / r / (R(a) + r);
        R ar = R(a) + r;
        R d2phi = (r * (R(2. * a) + R(3.) * r) - R(2.) * ar * ar * log(R(1.) + r / R(a))) / r / r / r / ar / ar;
        R2 += R(amp) * (d2phi * Rc * Rc / r / r + dphi * z * z / r / r / r);
        z2 += R(amp) * (d2phi * z * z / r / r + dphi * Rc * Rc / r / r / r);
        Rz += R(amp) * Rc * z * (d2phi / r / r - dphi / r / r / r);
        break;
      }
      case kHernquist: {
        double amp = C.p[0], a = C.p[1];
        R ar = R(a) + r;
        R dphi = R(amp) / (R(2.) * ar * ar);
        R d2phi = -R(amp) / (ar * ar * ar);
        R2 += d2phi * Rc * Rc / r2 + dphi * z * z / (r2 * r);
        z2 += d2phi * z * z / r2 + dphi * Rc * Rc / (r2 * r);
        Rz += Rc * z * (d2phi / r2 - dphi / (r2 * r));
        break;
      }
      case kFlattenedPower: {  // FlattenedPowerPotential.c R2deriv, z2deriv, Rzderiv
        double amp = C.p[0], alpha = C.p[1], q2 = C.p[2], core2 = C.p[3];
        if (alpha == 0.) {
          R denom = R(1.) / (Rc * Rc + z * z / R(q2) + R(core2));
          R2 += R(amp) * (denom - R(2.) * Rc * Rc * denom * denom);
          z2 += R(amp) / R(q2) * (denom - R(2.) * z * z * denom * denom / R(q2));
          Rz += R(-2.) * R(amp) * Rc * z / R(q2) * denom * denom;
        } else {
          R m2 = R(core2) + Rc * Rc + z * z / R(q2);
          R2 += -R(amp) * pow(m2, R(-0.5 * alpha - 1.)) * (R(alpha + 2.) * Rc * Rc / m2 - R(1.));
          z2 += -R(amp) / R(q2) * pow(m2, R(-0.5 * alpha - 1.)) *
                (R(alpha + 2.) * z * z / m2 / R(q2) - R(1.));
          Rz += -R(amp) * R(alpha + 2.) * Rc * z / R(q2) * pow(m2, R(-0.5 * alpha - 2.));
        }
        break;
      }
      case kNull: break;
      case kKuzminDisk: {  // KuzminDiskPotential.c R2deriv, z2deriv, Rzderiv
        double amp = C.p[0], a = C.p[1];
        R az = R(a) + fabs(z);
        R denom = Rc * Rc + az * az;
        R dm15 = R(1.) / (denom * sqrt(denom));
        R zsign = (z > R(0.)) ? R(1.) : ((z < R(0.)) ? R(-1.) : R(0.));
        R2 += R(amp) * (dm15 - R(3.) * Rc * Rc * dm15 / denom);
        z2 += R(amp) * (dm15 - R(3.) * az * az * dm15 / denom);
        Rz += R(amp) * (R(-3.) * zsign * Rc * az * R(1.) / (denom * denom * sqrt(denom)));
        break;
      }
      case kLogHalo: {
        double amp = C.p[0], q = C.p[1], c = C.p[2], onem1overb2 = C.p[3];
        if (onem1overb2 < 1) return false;
        R D = Rc * Rc + z * z / R(q * q) + R(c);
        R2 += R(amp) * (D - R(2.) * Rc * Rc) / (D * D);
        z2 += R(amp) * (D / R(q * q) - R(2.) * z * z / R(q * q * q * q)) / (D * D);
        Rz += -R(amp) * R(2.) * Rc * z / R(q * q) / (D * D);
        break;
      }
      default: {
        if (!is_spherical_new(C.type)) return false;
        // R2deriv = Phi'' R^2/r^2 + Phi' z^2/r^3, z2deriv = Phi'' z^2/r^2 +
        // Phi' R^2/r^3, Rzderiv = R z (Phi''/r^2 - Phi'/r^3)
        R Phip, Phipp;
        sph_dphi(C, r, Phip, Phipp);
        R ir2 = R(1.) / r2, ir3 = ir2 / r;
        R2 += Phipp * Rc * Rc * ir2 + Phip * z * z * ir3;
        z2 += Phipp * z * z * ir2 + Phip * Rc * Rc * ir3;
        Rz += Rc * z * (Phipp * ir2 - Phip * ir3);
      }
    }
  }
  return true;
}

// All six cylindrical second derivatives of DehnenBarPotential,
// DehnenBarPotential.c:180-316 (R2deriv, phi2deriv, Rphideriv, z2deriv,
// Rzderiv, zphideriv), in the reference's operation order.
template <class R>
void bar_hessian(const Component& C, R Rc, R z, R phi, R t, R& R2d, R& phi2d, R& Rphid, R& z2d,
                 R& Rzd, R& zphid) {
  double amp = C.p[0], tform = C.p[1], tsteady = C.p[2], rb = C.p[3], omegab = C.p[4], barphi = C.p[5];
  R smooth = dehnen_smooth(t, tform, tsteady);
  R R2 = Rc * Rc, z2 = z * z;
  R r2 = R2 + z2, r4 = r2 * r2, r6 = r4 * r2;
  R ang = R(2.) * (phi - R(omegab) * t - R(barphi));
  R cprefac = R(amp) * smooth * cos(ang);
  R sprefac = R(-2.) * R(amp) * smooth * sin(ang);
  R rb3 = R(rb * rb * rb);
  if (r2 <= R(rb * rb)) {
    R r3rb3 = pow(r2, R(1.5)) / rb3;
    R2d = cprefac * (r3rb3 * ((R(9.) * R2 + R(2.) * z2) / r4 - R2 / r6 * (R(3.) * R2 + R(2.) * z2)) +
                     R(4.) * z2 / r6 * (R(4.) * R2 - r2));
    phi2d = -R(4.) * cprefac * (r3rb3 - R(2.)) * R2 / r2;
    Rphid = sprefac * (r3rb3 * Rc * (R(3.) * R2 + R(2.) * z2) - R(4.) * Rc * z2) / r4;
    z2d = cprefac * R2 / r6 * (r3rb3 * (r2 - z2) + R(4.) * (r2 - R(4.) * z2));
    Rzd = cprefac * Rc * z / r6 * (r3rb3 * (R(2.) * r2 - R2) + R(8.) * (r2 - R(2.) * R2));
    zphid = sprefac * (r3rb3 + R(4.)) * R2 * z / r4;
  } else {
    R rbr3 = pow(R(rb), R(3.)) * pow(r2, R(-1.5));
    R2d = cprefac * rbr3 / r6 * ((r2 - R(7.) * R2) * (R(3.) * R2 - R(2.) * z2) + R(6.) * R2 * r2);
    phi2d = R(4.) * cprefac * rbr3 * R2 / r2;
    Rphid = sprefac * rbr3 * Rc / r4 * (R(3.) * R2 - R(2.) * z2);
    z2d = cprefac * R(5.) * rbr3 * R2 / r6 * (r2 - R(7.) * z2);
    Rzd = cprefac * R(5.) * rbr3 * Rc * z / r6 * (R(2.) * r2 - R(7.) * R2);
    zphid = sprefac * R(5.) * rbr3 * R2 * z / r4;
  }
}

// The same six for a triaxial LogarithmicHaloPotential (1 - 1/b^2 < 1),
// LogarithmicHaloPotential.c:157-283.
template <class R>
void loghalo_triax_hessian(const Component& C, R Rc, R z, R phi, R& R2d, R& phi2d, R& Rphid,
                           R& z2d, R& Rzd, R& zphid) {
  double amp = C.p[0], q = C.p[1], c = C.p[2], onem1overb2 = C.p[3];
  R zq = z / R(q);
  R q2 = R(q * q);
  R R2 = Rc * Rc;
  R Rt2 = R2 * (R(1.) - R(onem1overb2) * pow(sin(phi), R(2)));
  R denom = R(1.) / (Rt2 + zq * zq + R(c));
  R2d = R(amp) * (denom - R(2.) * Rt2 * denom * denom) * Rt2 / (Rc * Rc);
  z2d = R(amp) * (denom / q2 - R(2.) * z * z * denom * denom / (q2 * q2));
  Rzd = -R(amp) * R(2.) * Rt2 / Rc * z / q2 * denom * denom;
  R s2phi = sin(R(2.) * phi);
  phi2d = -R(amp) * R(onem1overb2) *
          (R(0.5) * R2 * R2 * s2phi * s2phi * R(onem1overb2) * denom * denom + R2 * denom * cos(R(2.) * phi));
  Rphid = -R(amp) * (denom - Rt2 * denom * denom) * Rc * sin(R(2.) * phi) * R(onem1overb2);
  zphid = R(amp) * R(2.) * Rc * Rc * z * sin(phi) * cos(phi) * R(onem1overb2) * denom * denom / q2;
}

// calcR2deriv .. calczphideriv (galpy_potentials.c), summed over the
// components in index order: the axisymmetric components through
// cyl_hessian (their phi derivatives vanish), the bar and the triaxial halo
// through the functions above.  Returns false for a component without a
// restated Hessian.
template <class R>
bool cyl_hessian_full(int npot, const Component* P, R Rc, R z, R phi, R t, R& R2, R& z2, R& Rz,
                      R& phi2, R& Rphi, R& zphi) {
  R2 = 0.; z2 = 0.; Rz = 0.; phi2 = 0.; Rphi = 0.; zphi = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    R a = 0., b = 0., cc = 0., d = 0., e = 0., f = 0.;
    if (C.type == kDehnenBar) {
      bar_hessian(C, Rc, z, phi, t, a, d, e, b, cc, f);
    } else if (C.type == kLogHalo && C.p[3] < 1) {
      loghalo_triax_hessian(C, Rc, z, phi, a, d, e, b, cc, f);
    } else if (!cyl_hessian(1, &C, Rc, z, a, b, cc)) {
      return false;
    }
    R2 += a; z2 += b; Rz += cc; phi2 += d; Rphi += e; zphi += f;
  }
  return true;
}

// calcPlanarR2deriv / phi2deriv / Rphideriv: the planar functions of the bar
// (DehnenBarPotential.c:125-179) and the triaxial halo
// (LogarithmicHaloPotential.c:108-156); the other components' planar
// second derivatives are their full ones in the plane.
template <class R>
bool planar_hessian(int npot, const Component* P, R Rc, R phi, R t, R& R2, R& phi2, R& Rphi) {
  R2 = 0.; phi2 = 0.; Rphi = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    R a = 0., d = 0., e = 0.;
    if (C.type == kDehnenBar) {
      double amp = C.p[0], tform = C.p[1], tsteady = C.p[2], rb = C.p[3], omegab = C.p[4], barphi = C.p[5];
      R smooth = dehnen_smooth(t, tform, tsteady);
      R ang = R(2.) * (phi - R(omegab) * t - R(barphi));
      if (Rc <= R(rb)) {
        a = R(6.) * R(amp) * smooth * cos(ang) * pow(Rc / R(rb), R(3.)) / Rc / Rc;
        d = -R(4.) * R(amp) * smooth * cos(ang) * (pow(Rc / R(rb), R(3.)) - R(2.));
        e = -R(6.) * R(amp) * smooth * sin(ang) * pow(Rc / R(rb), R(3.)) / Rc;
      } else {
        a = -R(12.) * R(amp) * smooth * cos(ang) * pow(R(rb) / Rc, R(3.)) / Rc / Rc;
        d = R(4.) * R(amp) * smooth * cos(ang) * pow(R(rb) / Rc, R(3.));
        e = -R(6.) * R(amp) * smooth * sin(ang) * pow(R(rb) / Rc, R(3.)) / Rc;
      }
    } else if (C.type == kLogHalo && C.p[3] < 1) {
      double amp = C.p[0], c = C.p[2], onem1overb2 = C.p[3];
      R Rt2 = Rc * Rc * (R(1.) - R(onem1overb2) * pow(sin(phi), R(2)));
      a = R(amp) * (R(1.) - R(2.) * Rt2 / (Rt2 + R(c))) / (Rt2 + R(c)) * Rt2 / Rc / Rc;
      d = -R(amp) * R(onem1overb2) *
          (R(0.5) * pow(Rc * Rc * sin(R(2.) * phi), R(2.)) * R(onem1overb2) / (Rt2 + R(c)) / (Rt2 + R(c)) +
           Rc * Rc / (Rt2 + R(c)) * cos(R(2.) * phi));
      e = -R(amp) * R(c) / (Rt2 + R(c)) / (Rt2 + R(c)) * Rc * sin(R(2.) * phi) * R(onem1overb2);
    } else {
      R z2d, Rzd;
      if (!cyl_hessian(1, &C, Rc, R(0.), a, z2d, Rzd)) return false;
    }
    R2 += a; phi2 += d; Rphi += e;
  }
  return true;
}

// evalRectForce_dxdv, integrateFullOrbit.c:1619-1681: base acceleration in
// a[0..2] (as rect_force) and, per deviation column j, K dq_j with K the
// Cartesian Hessian of -Phi assembled from the six cylindrical second
// derivatives and the (conservative) forces, in the reference's form.
template <class R>
bool rect_force_dxdv(int npot, const Component* P, R t, const R* q, R* a, int nde) {
  rect_force(npot, P, t, q, a);
  if (nde == 0) return true;
  R x = q[0], y = q[1], z = q[2];
  R Rc = sqrt(x * x + y * y);
  R sinphi = y / Rc, cosphi = x / Rc;
  R phi = acos(x / Rc);
  if (y < R(0.)) phi = R(2. * M_PI) - phi;
  R R2deriv, z2deriv, Rzderiv, phi2deriv, Rphideriv, zphideriv;
  if (!cyl_hessian_full(npot, P, Rc, z, phi, t, R2deriv, z2deriv, Rzderiv, phi2deriv, Rphideriv,
                        zphideriv))
    return false;
  R RforceK, zforce, phitorqueK;
  cyl_forces(npot, P, Rc, z, phi, t, RforceK, zforce, phitorqueK);
  R dFxdx = -cosphi * cosphi * R2deriv + R(2.) * cosphi * sinphi / Rc / Rc * phitorqueK +
            sinphi * sinphi / Rc * RforceK + R(2.) * sinphi * cosphi / Rc * Rphideriv -
            sinphi * sinphi / Rc / Rc * phi2deriv;
  R dFxdy = -sinphi * cosphi * R2deriv + (sinphi * sinphi - cosphi * cosphi) / Rc / Rc * phitorqueK -
            cosphi * sinphi / Rc * RforceK - (cosphi * cosphi - sinphi * sinphi) / Rc * Rphideriv +
            cosphi * sinphi / Rc / Rc * phi2deriv;
  R dFydy = -sinphi * sinphi * R2deriv - R(2.) * sinphi * cosphi / Rc / Rc * phitorqueK -
            R(2.) * sinphi * cosphi / Rc * Rphideriv + cosphi * cosphi / Rc * RforceK -
            cosphi * cosphi / Rc / Rc * phi2deriv;
  R dFxdz = -cosphi * Rzderiv + sinphi / Rc * zphideriv;
  R dFydz = -sinphi * Rzderiv - cosphi / Rc * zphideriv;
  R dFzdz = -z2deriv;
  for (int j = 1; j <= nde; j++) {
    R dx = q[3 * j], dy = q[3 * j + 1], dz = q[3 * j + 2];
    a[3 * j] = dFxdx * dx + dFxdy * dy + dFxdz * dz;
    a[3 * j + 1] = dFxdy * dx + dFydy * dy + dFydz * dz;
    a[3 * j + 2] = dFxdz * dx + dFydz * dy + dFzdz * dz;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Planar forces (integratePlanarOrbit.c:55-739 wires each code to these
// functions): galpy's separate planarRforce / planarphitorque kernels, as
// written in the component files.
// ---------------------------------------------------------------------------
// PowerSphericalPotentialwCutoff.c:47-58
template <class R> R psc_planar_Rforce(const Component& C, R Rc) {
  double amp = C.p[0], rc = C.p[2];
  R r2 = Rc * Rc;
  R x = r2 / R(rc) / R(rc);
  R P = lower_incgamma(C.c[0], x, C.c[1]) / R(C.c[1]);
  R mass = R(C.c[2]) * (R(C.c[1]) * P);
  return -R(amp) * mass / r2;
}
// MiyamotoNagaiPotential.c:27-37
template <class R> R mn_planar_Rforce(const Component& C, R Rc) {
  double amp = C.p[0], a = C.p[1], b = C.p[2];
  return -R(amp) * Rc * pow(Rc * Rc + pow(R(a + b), R(2.)), R(-1.5));
}
// NFWPotential.c:28-37
template <class R> R nfw_planar_Rforce(const Component& C, R Rc) {
  double amp = C.p[0], a = C.p[1];
  return R(amp) / Rc * (R(1.) / (R(a) + Rc) - log(R(1.) + Rc / R(a)) / Rc);
}
// HernquistPotential.c:26-35
template <class R> R hq_planar_Rforce(const Component& C, R Rc) {
  double amp = C.p[0], a = C.p[1];
  return -R(amp) / R(a) * pow(R(1.) + Rc / R(a), R(-2.)) / R(2.) / R(a);
}
// LogarithmicHaloPotential.c:39-54 (planarRforce), :92-107 (planarphitorque)
template <class R> void lh_planar_forces(const Component& C, R Rc, R phi, R& FR, R& Fphi) {
  double amp = C.p[0], c = C.p[2], onem1overb2 = C.p[3];
  if (onem1overb2 < 1) {
    R Rt2 = Rc * Rc * (R(1.) - R(onem1overb2) * pow(sin(phi), R(2)));
    FR = -R(amp) * Rt2 / Rc / (Rt2 + R(c));
    Fphi = R(amp) * Rc * Rc / (Rt2 + R(c)) * sin(R(2) * phi) * R(onem1overb2) / R(2.);
  } else {
    FR = -R(amp) * Rc / (Rc * Rc + R(c));
    Fphi = R(0.);
  }
}
// DehnenBarPotential.c:41-59 (planarRforce), :83-101 (planarphitorque)
template <class R> void bar_planar_forces(const Component& C, R Rc, R phi, R t, R& FR, R& Fphi) {
  double amp = C.p[0], tform = C.p[1], tsteady = C.p[2], rb = C.p[3], omegab = C.p[4], barphi = C.p[5];
  R smooth = dehnen_smooth(t, tform, tsteady);
  R ang = R(2.) * (phi - R(omegab) * t - R(barphi));
  if (Rc <= R(rb)) {
    FR = -R(3.) * R(amp) * smooth * cos(ang) * pow(Rc / R(rb), R(3.)) / Rc;
    Fphi = R(2.) * R(amp) * smooth * sin(ang) * (pow(Rc / R(rb), R(3.)) - R(2.));
  } else {
    FR = -R(3.) * R(amp) * smooth * cos(ang) * pow(R(rb) / Rc, R(3.)) / Rc;
    Fphi = -R(2.) * R(amp) * smooth * sin(ang) * pow(R(rb) / Rc, R(3.));
  }
}

// calcPlanarRforce / calcPlanarphitorque, galpy_potentials.c:180-209: sums
// over components in index order from 0.
template <class R>
void planar_forces(int npot, const Component* P, R Rc, R phi, R t, R& FR, R& Fphi) {
  FR = 0.; Fphi = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    R fr = 0., fp = 0.;
    switch (C.type) {
      case kPowerSphCutoff: fr = psc_planar_Rforce(C, Rc); break;
      case kMiyamotoNagai: fr = mn_planar_Rforce(C, Rc); break;
      case kNFW: fr = nfw_planar_Rforce(C, Rc); break;
      case kHernquist: fr = hq_planar_Rforce(C, Rc); break;
      case kLogHalo: lh_planar_forces(C, Rc, phi, fr, fp); break;
      case kDehnenBar: bar_planar_forces(C, Rc, phi, t, fr, fp); break;
      case kFlattenedPower: {  // FlattenedPowerPotential.c PlanarRforce
        double amp = C.p[0], alpha = C.p[1], core2 = C.p[3];
        if (alpha == 0.)
          fr = -R(amp) * Rc / (Rc * Rc + R(core2));
        else
          fr = -R(amp) * pow(R(core2) + Rc * Rc, R(-0.5 * alpha - 1.)) * Rc;
        break;
      }
      case kNull: break;
      case kKKStaeckel: fr = kks_planar_rforce(C, Rc); break;
      case kEllipDisk: ellipdisk_planar_forces(C, Rc, phi, t, fr, fp); break;
      case kSteadySpiral: {  // SteadyLogSpiralPotential.c Rforce, phitorque
        const double amp = C.p[0], A = C.p[3], alpha = C.p[4], m = C.p[5], om = C.p[6], gam = C.p[7];
        R smooth = ellipdisk_smooth(t, C.p[1], C.p[2]);  // dehnenSpiralSmooth: the same polynomial
        fr = R(amp) * smooth * R(A) / Rc * sin(R(alpha) * log(Rc) - R(m) * (phi - R(om) * t - R(gam)));
        fp = -R(amp) * smooth * R(A) / R(alpha) * R(m) *
             sin(R(alpha) * log(Rc) - R(m) * (phi - R(om) * t - R(gam)));
        break;
      }
      case kTransientSpiral: {  // TransientLogSpiralPotential.c:5-42
        const double amp = C.p[0], A = C.p[1], to = C.p[2], sigma2 = C.p[3], alpha = C.p[4], m = C.p[5],
                     om = C.p[6], gam = C.p[7];
        R g = exp(-pow(t - R(to), R(2.)) / R(2.) / R(sigma2));
        R sn = sin(R(alpha) * log(Rc) - R(m) * (phi - R(om) * t - R(gam)));
        fr = R(amp) * R(A) * g / Rc * sn;
        fp = -R(amp) * R(A) * g / R(alpha) * R(m) * sn;
        break;
      }
      case kCosmphiDisk: {  // CosmphiDiskPotential.c:4-37 (m read as an int)
        const double amp = C.p[0], mphio = C.p[1], pp = C.p[2], mphib = C.p[3], rb = C.p[5], rbp = C.p[6],
                     rb2p = C.p[7], r1p = C.p[8];
        const int m = (int)C.p[4];
        if (Rc <= R(rb)) {
          fr = -R(amp) * R(pp) * R(mphio) / R(m) * R(rb2p) / pow(Rc, R(pp + 1.)) * cos(R(m) * phi - R(mphib));
          fp = R(amp) * R(mphio) * R(rbp) * (R(2.) * R(r1p) - R(rbp) / pow(Rc, R(pp))) * sin(R(m) * phi - R(mphib));
        } else {
          fr = -R(amp) * R(pp) * R(mphio) / R(m) * pow(Rc, R(pp - 1.)) * cos(R(m) * phi - R(mphib));
          fp = R(amp) * R(mphio) * pow(Rc, R(pp)) * sin(R(m) * phi - R(mphib));
        }
        break;
      }
      case kHenonHeiles:  // HenonHeilesPotential.c:4-14
        fr = -R(C.p[0]) * Rc * (R(1.) + Rc * sin(R(3.) * phi));
        fp = -R(C.p[0]) * pow(Rc, R(3.)) * cos(R(3.) * phi);
        break;
      case kLopsided: {  // LopsidedDiskPotential.c Rforce, phitorque (amp, phio, p, phib)
        const double amp = C.p[0], phio = C.p[1], pp = C.p[2], phib = C.p[3];
        fr = -R(amp) * R(pp) * R(phio) * pow(Rc, R(pp - 1.)) * cos(phi - R(phib));
        fp = R(amp) * R(phio) * pow(Rc, R(pp)) * sin(phi - R(phib));
        break;
      }
      case kSoftNeedleBar: {
        R fz0;
        needle_forces(C, Rc, R(0.), phi, t, fr, fz0, fp);
        break;
      }
      case kDoubleExpDisk: fr = dedisk_planar_Rforce(C, Rc); break;
      case kDiskSCF: {  // DiskSCFPotential.c:133-144: -Sigma'(R) H(0)
        const double* a = C.args;
        R S, dS, H, dH;
        dscf_sigma(a + 1, Rc, S, dS);
        dscf_hz(a + 1 + (int)a[0], R(0.), H, dH);
        fr = -dS * H;
        break;
      }
      case kSCF: {  // SCFPotential.c:896-906: the full forces at z = 0
        R fz0;
        scf_forces(C, Rc, R(0.), phi, t, fr, fz0, fp);
        break;
      }
      case kMultipole: {  // MultipoleExpansionPotential.c:1030-1045: the full forces at z = 0
        R fz0;
        mp_forces(C, Rc, R(0.), phi, t, fr, fz0, fp);
        break;
      }
      case kInterpSph: fr = R(C.p[0]) * interpsph_rforce(C, Rc); break;  // SphericalPotential.c:22-29
      case kSpiralArms: {  // SpiralArmsPotential.c:625-716: the full forces at z = 0
        R fz0;
        spiral_forces(C, Rc, R(0.), phi, t, fr, fz0, fp);
        break;
      }
      case kKuzminDisk:  // KuzminDiskPotential.c planarRforce
        fr = -R(C.p[0]) * Rc * pow(Rc * Rc + R(C.p[1] * C.p[1]), R(-1.5));
        break;
      default:
        if (is_ellipsoidal(C.type)) {  // EllipsoidalPotential.c:116-123: the full forces at z = 0
          R fz0;
          ellip_forces(C, Rc, R(0.), phi, fr, fz0, fp);
        } else if (is_spherical_new(C.type)) {  // planarRforce = -Phi'(R)
          R Phip, Phipp;
          sph_dphi(C, Rc, Phip, Phipp);
          fr = -Phip;
        } else if (C.type == kMovingObjectWrap) {  // MovingObjectPotential.c:89-130
          R ox, oy, oz;
          movingobject_pos(C, t, ox, oy, oz);
          R x = Rc * cos(phi), y = Rc * sin(phi);
          R Rdist = pow(pow(x - ox, R(2)) + pow(y - oy, R(2)), R(0.5));
          R wr, wp;
          planar_forces(C.nwrapped, C.wrapped, Rdist, phi, t, wr, wp);
          fr = -R(C.p[0]) * wr * (cos(phi) * (ox - x) + sin(phi) * (oy - y)) / Rdist;
          fp = -R(C.p[0]) * wr * Rc * (cos(phi) * (oy - y) - sin(phi) * (ox - x)) / Rdist;
        } else if (C.type == kKuzminLikeWrap) {  // KuzminLikeWrapperPotential.c:93-107
          const double amp = C.p[0], a = C.p[1], b2 = C.p[2];
          R xi = kuzminlike_xi(Rc, R(0.), a, b2);
          R wr, wp;
          planar_forces(C.nwrapped, C.wrapped, xi, R(0.), t, wr, wp);
          fr = R(amp) * wr * (Rc / xi);
        } else if (C.type == kOblateStaeckelWrap) {  // OblateStaeckelWrapperPotential.c:215-229
          fr = R(os_planar_Rforce(C, dval(Rc)));
        } else if (C.type == kCylSepWrap) {  // CylindricallySeparablePotentialWrapper.c:28-34
          R wr, wp;
          planar_forces(C.nwrapped, C.wrapped, Rc, R(0.), R(0.), wr, wp);
          fr = R(C.p[0]) * wr;
        } else if (is_wrapper(C.type)) {  // the wrappers' Planar functions
          R fac, off, doff, wr, wp;
          wrapper_transform(C, Rc, t, fac, off, doff);
          planar_forces(C.nwrapped, C.wrapped, Rc, phi - off, t, wr, wp);
          fr = fac * (wr - wp * doff);
          fp = fac * wp;
        }
    }
    FR += fr; Fphi += fp;
  }
}

// evalPlanarRectForce: integratePlanarOrbit.c:1015-1031
template <class R>
void planar_rect_force(int npot, const Component* P, R t, const R* q, R* a) {
  R x = q[0], y = q[1];
  R Rc = sqrt(x * x + y * y);
  R phi = acos(x / Rc);
  R sinphi = y / Rc;
  R cosphi = x / Rc;
  if (y < R(0.)) phi = R(2. * M_PI) - phi;
  R FR, Fphi;
  planar_forces(npot, P, Rc, phi, t, FR, Fphi);
  a[0] = cosphi * FR - R(1.) / Rc * sinphi * Fphi;
  a[1] = sinphi * FR + R(1.) / Rc * cosphi * Fphi;
}

// evalPlanarRectDeriv: integratePlanarOrbit.c:1032-1055 (vR, vT feed only
// dissipative forces, none in scope)
inline void diss_forces(int npot, const Component* P, double Rc, double z, double phi, double t, double vR,
                        double vT, double vz, double& FR, double& Fz, double& Fp);
inline bool has_dissipative(int npot, const Component* P);
// the velocity-dependent forces of planar orbits (evalPlanarRectDeriv with
// PlanarRforce(R, phi, t, vR, vT) = Rforce(R, 0, phi, t, vR, vT, 0))
inline void planar_diss_add(int npot, const Component* P, double t, const double* q, double* a) {
  const double x = q[0], y = q[1], Rc = std::sqrt(x * x + y * y);
  double phi = std::acos(x / Rc);
  const double sinphi = y / Rc, cosphi = x / Rc;
  if (y < 0.) phi = 2. * M_PI - phi;
  const double vR = q[2] * cosphi + q[3] * sinphi, vT = -q[2] * sinphi + q[3] * cosphi;
  double FR, Fz, Fp;
  diss_forces(npot, P, Rc, 0., phi, t, vR, vT, 0., FR, Fz, Fp);
  a[2] += cosphi * FR - 1. / Rc * sinphi * Fp;
  a[3] += sinphi * FR + 1. / Rc * cosphi * Fp;
}
template <class R>
void planar_rect_deriv(int npot, const Component* P, R t, const R* q, R* a) {
  a[0] = q[2]; a[1] = q[3];
  planar_rect_force(npot, P, t, q, a + 2);
  if (has_dissipative(npot, P)) {
    const double qd[4] = {dval(q[0]), dval(q[1]), dval(q[2]), dval(q[3])};
    double ad[4] = {0., 0., 0., 0.};
    planar_diss_add(npot, P, dval(t), qd, ad);
    a[2] += R(ad[2]);
    a[3] += R(ad[3]);
  }
}

// Components with a restated Hessian (cyl_hessian_full, planar_hessian):
// the axisymmetric ones, the bar and the triaxial halo.
inline bool has_hessian(int npot, const Component* P) {
  for (int i = 0; i < npot; i++) {
    const int c = P[i].type;
    if (c == kKKStaeckel || c == kSoftNeedleBar || is_wrapper(c) || c == kEllipDisk || c == kLopsided ||
        c == kSteadySpiral || c == kTransientSpiral || c == kCosmphiDisk || c == kHenonHeiles ||
        c == kSpiralArms || c == kDoubleExpDisk || c == kInterpSph || c == kSCF || c == kDiskSCF || c == kInterpRZ || c == kMultipole || is_ellipsoidal(c))
      return false;
  }
  return true;
}

// evalRectDeriv_dxdv, integrateFullOrbit.c:1496-1600, for conservative
// components (the dissipative Jacobian blocks vanish): the 12-dim state
// (x,y,z,vx,vy,vz | dx,dy,dz,dvx,dvy,dvz) of integrate_dxdv's DOP853 path;
// the deviation's acceleration is K (dx,dy,dz) with K assembled as in
// rect_force_dxdv.  Returns false for a component without a Hessian.
template <class R>
bool rect_deriv_dxdv(int npot, const Component* P, R t, const R* q, R* a) {
  R qq[6] = {q[0], q[1], q[2], q[6], q[7], q[8]};
  R aa[6] = {0., 0., 0., 0., 0., 0.};
  const bool ok = rect_force_dxdv(npot, P, t, qq, aa, 1);
  a[0] = q[3]; a[1] = q[4]; a[2] = q[5];
  a[3] = aa[0]; a[4] = aa[1]; a[5] = aa[2];
  a[6] = q[9]; a[7] = q[10]; a[8] = q[11];
  a[9] = aa[3]; a[10] = aa[4]; a[11] = aa[5];
  return ok;
}

// evalPlanarRectForce_dxdv, integratePlanarOrbit.c:1228-1275: planar
// acceleration and, per deviation column, the kick tangent K (dx, dy).  galpy
// wires calcPlanarR2deriv / phi2deriv / Rphideriv to each component's planar
// second derivatives (planar_hessian).
template <class R>
bool planar_rect_force_dxdv(int npot, const Component* P, R t, const R* q, R* a, int nde) {
  R x = q[0], y = q[1];
  R Rc = sqrt(x * x + y * y);
  R phi = acos(x / Rc);
  R sinphi = y / Rc;
  R cosphi = x / Rc;
  if (y < R(0.)) phi = R(2. * M_PI) - phi;
  R Rforce, phitorque;
  planar_forces(npot, P, Rc, phi, t, Rforce, phitorque);
  a[0] = cosphi * Rforce - R(1.) / Rc * sinphi * phitorque;
  a[1] = sinphi * Rforce + R(1.) / Rc * cosphi * phitorque;
  if (nde == 0) return true;
  R R2deriv, phi2deriv, Rphideriv;
  if (!planar_hessian(npot, P, Rc, phi, t, R2deriv, phi2deriv, Rphideriv)) return false;
  R dFxdx = -cosphi * cosphi * R2deriv + R(2.) * cosphi * sinphi / Rc / Rc * phitorque +
            sinphi * sinphi / Rc * Rforce + R(2.) * sinphi * cosphi / Rc * Rphideriv -
            sinphi * sinphi / Rc / Rc * phi2deriv;
  R dFxdy = -sinphi * cosphi * R2deriv + (sinphi * sinphi - cosphi * cosphi) / Rc / Rc * phitorque -
            cosphi * sinphi / Rc * Rforce - (cosphi * cosphi - sinphi * sinphi) / Rc * Rphideriv +
            cosphi * sinphi / Rc / Rc * phi2deriv;
  R dFydy = -sinphi * sinphi * R2deriv - R(2.) * sinphi * cosphi / Rc / Rc * phitorque -
            R(2.) * sinphi * cosphi / Rc * Rphideriv + cosphi * cosphi / Rc * Rforce -
            cosphi * cosphi / Rc / Rc * phi2deriv;
  for (int j = 1; j <= nde; j++) {
    R dx = q[2 * j], dy = q[2 * j + 1];
    a[2 * j] = dFxdx * dx + dFxdy * dy;
    a[2 * j + 1] = dFxdy * dx + dFydy * dy;
  }
  return true;
}

// evalPlanarRectDeriv_dxdv, integratePlanarOrbit.c:1124-1175: the 8-dim state
// (x,y,vx,vy | dx,dy,dvx,dvy)
template <class R>
bool planar_rect_deriv_dxdv(int npot, const Component* P, R t, const R* q, R* a) {
  R qq[4] = {q[0], q[1], q[4], q[5]};
  R aa[4] = {0., 0., 0., 0.};
  const bool ok = planar_rect_force_dxdv(npot, P, t, qq, aa, 1);
  a[0] = q[2]; a[1] = q[3];
  a[2] = aa[0]; a[3] = aa[1];
  a[4] = q[6]; a[5] = q[7];
  a[6] = aa[2]; a[7] = aa[3];
  return ok;
}

// Potential value: galpy_potentials.c evaluatePotentials analog (sum in
// component order); the bar uses the Python expression.
template <class R>
R potential_value(int npot, const Component* P, R Rc, R z, R phi, R t) {
  R pot = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    switch (C.type) {
      case kPowerSphCutoff: pot += psc_eval(C, Rc, z); break;
      case kMiyamotoNagai: pot += mn_eval(C, Rc, z); break;
      case kNFW: pot += nfw_eval(C, Rc, z); break;
      case kHernquist: pot += hq_eval(C, Rc, z); break;
      case kLogHalo: pot += lh_eval(C, Rc, z, phi); break;
      case kDehnenBar: pot += bar_eval(C, Rc, z, phi, t); break;
      case kKuzminDisk: pot += kuzmin_eval(C, Rc, z); break;
      case kKKStaeckel: pot += kks_eval(C, Rc, z); break;
      case kSoftNeedleBar: pot += needle_eval(C, Rc, z, phi, t); break;
      case kSpiralArms: pot += spiral_eval(C, Rc, z, phi, t); break;
      case kDoubleExpDisk: pot += dedisk_eval(C, Rc, z); break;
      case kSCF: pot += scf_eval(C, Rc, z, phi, t); break;
      case kMultipole: pot += mp_eval(C, Rc, z, phi, t); break;
      case kInterpRZ: pot += R(interprz_eval(C, 0, dval(Rc), dval(z))); break;  // interpRZPotential.c:6-20
      case kDiskSCF: {  // DiskSCFPotential.c:108-120 (Eval)
        const double* a = C.args;
        R S, dS, H, dH;
        dscf_sigma(a + 1, sqrt(Rc * Rc + z * z), S, dS);
        dscf_hz(a + 1 + (int)a[0], z, H, dH);
        pot += S * H;
        break;
      }
      case kInterpSph: pot += R(C.p[0]) * interpsph_reval(C, sqrt(Rc * Rc + z * z)); break;
      case kFlattenedPower: {  // FlattenedPowerPotential.c Eval
        double amp = C.p[0], alpha = C.p[1], q2 = C.p[2], core2 = C.p[3];
        if (alpha == 0.)
          pot += R(0.5) * R(amp) * log(Rc * Rc + z * z / R(q2) + R(core2));
        else
          pot += -R(amp) * pow(R(core2) + Rc * Rc + z * z / R(q2), R(-0.5 * alpha)) / R(alpha);
        break;
      }
      case kNull: break;
      default:
        if (is_ellipsoidal(C.type)) {
          pot += ellip_eval(C, Rc, z, phi);
        } else if (is_spherical_new(C.type)) {
          pot += sph_eval(C, sqrt(Rc * Rc + z * z));
        } else if (C.type == kMovingObjectWrap) {  // MovingObjectPotential.py _evaluate
          R ox, oy, oz;
          movingobject_pos(C, t, ox, oy, oz);
          R x = Rc * cos(phi), y = Rc * sin(phi);
          pot += R(C.p[0]) * potential_value(C.nwrapped, C.wrapped, sqrt((x - ox) * (x - ox) + (y - oy) * (y - oy)),
                                             oz - z, phi, t);
        } else if (C.type == kRotateTiltWrap) {  // RotateAndTiltWrapperPotential.py _evaluate
          R xp, yp, zp;
          rottilt_point(C, Rc, z, phi, xp, yp, zp);
          pot += R(C.p[0]) * potential_value(C.nwrapped, C.wrapped, sqrt(xp * xp + yp * yp), zp,
                                             atan2(yp, xp), t);
        } else if (C.type == kKuzminLikeWrap) {  // KuzminLikeWrapperPotential.py _evaluate
          pot += R(C.p[0]) * potential_value(C.nwrapped, C.wrapped, kuzminlike_xi(Rc, z, C.p[1], C.p[2]),
                                             R(0.), phi, t);
        } else if (C.type == kOblateStaeckelWrap) {  // OblateStaeckelWrapperPotential.c:174-182
          pot += R(os_eval(C, dval(Rc), dval(z)));
        } else if (C.type == kCylSepWrap) {  // CylindricallySeparablePotentialWrapper.py _evaluate
          pot += R(C.p[0]) * (potential_value(C.nwrapped, C.wrapped, Rc, R(0.), R(0.), R(0.)) +
                              potential_value(C.nwrapped, C.wrapped, R(C.p[1]), z, R(0.), R(0.)) -
                              R(C.p[2]));
        } else if (is_wrapper(C.type)) {
          // the wrappers' Python _evaluate (the wrapped potential at
          // phi - offset and the same t, times the factor)
          R fac, off, doff;
          wrapper_transform(C, Rc, t, fac, off, doff);
          pot += fac * potential_value(C.nwrapped, C.wrapped, Rc, z, phi - off, t);
        }
    }
  }
  return pot;
}

// NonInertialFrameForce.c:27-170: the fictitious forces of a frame whose
// origin accelerates by a0(t) and that rotates at Omega(t): -a0 +
// 2 v x Omega... written out as galpy does (centrifugal, Coriolis, Euler
// terms, with the frame origin's x0, v0 when both accelerations are on).
// Time inputs from the caller's functions (pot_type 39; Omegadot its own
// function) or from natural splines of them (45; Omegadot the Omega
// spline's derivative, times clamped to [tmin, tmax]).  args: amp, 10 cache
// slots, rot_acc, lin_acc, omegaz_only, const_freq, Omega_as_func,
// Omega[3], Omega^2, Omegadot[3] (then tmin, tmax for 45).
inline void noninertial_xyz(const Component& C, double x, double y, double z, double vx, double vy, double vz,
                            double t, double& Fx, double& Fy, double& Fz) {
  const double* a = C.args;
  const bool interp = C.type == kNonInertialInterp;
  const double tq = interp ? (t < a[23] ? a[23] : (t > a[24] ? a[24] : t)) : t;
  auto tdep = [&](int idx) { return interp ? (*C.spl)[idx].eval(tq) : g_tfuncs[C.tf0 + idx](t); };
  auto tdepd = [&](int sidx, int tidx) { return interp ? (*C.spl)[sidx].deriv(tq) : g_tfuncs[C.tf0 + tidx](t); };
  const bool rot_acc = a[11] != 0., lin_acc = a[12] != 0., Omega_as_func = a[15] != 0.;
  const int la = lin_acc ? 1 : 0;
  Fx = Fy = Fz = 0.;
  if (rot_acc) {
    const bool omegaz_only = a[13] != 0., const_freq = a[14] != 0.;
    if (omegaz_only) {
      double Oz, O2, Odz = 0.;
      if (Omega_as_func) {
        Oz = tdep(9 * la);
        O2 = Oz * Oz;
      } else {
        Oz = a[18];
        if (!const_freq) {
          Oz += a[22] * t;
          O2 = Oz * Oz;
        } else {
          O2 = a[19];
        }
      }
      Fx += 2. * Oz * vy + O2 * x;
      Fy += -2. * Oz * vx + O2 * y;
      if (!const_freq) {
        Odz = Omega_as_func ? tdepd(9 * la, 9 * la + 1) : a[22];
        Fx += Odz * y;
        Fy -= Odz * x;
      }
      if (lin_acc) {
        const double x0x = tdep(3), x0y = tdep(4), v0x = tdep(6), v0y = tdep(7);
        Fx += 2. * Oz * v0y + O2 * x0x;
        Fy += -2. * Oz * v0x + O2 * x0y;
        if (!const_freq) {
          Fx += Odz * x0y;
          Fy -= Odz * x0x;
        }
      }
    } else {
      double Ox, Oy, Oz, O2, Odx = 0., Ody = 0., Odz = 0.;
      if (Omega_as_func) {
        Ox = tdep(9 * la);
        Oy = tdep(9 * la + 1);
        Oz = tdep(9 * la + 2);
        O2 = Ox * Ox + Oy * Oy + Oz * Oz;
      } else {
        Ox = a[16];
        Oy = a[17];
        Oz = a[18];
        if (!const_freq) {
          Ox += a[20] * t;
          Oy += a[21] * t;
          Oz += a[22] * t;
          O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        } else {
          O2 = a[19];
        }
      }
      double Ov = Ox * x + Oy * y + Oz * z;
      Fx += 2. * (Oz * vy - Oy * vz) + O2 * x - Ox * Ov;
      Fy += -2. * (Oz * vx - Ox * vz) + O2 * y - Oy * Ov;
      Fz += 2. * (Oy * vx - Ox * vy) + O2 * z - Oz * Ov;
      if (!const_freq) {
        if (Omega_as_func) {
          Odx = tdepd(9 * la, 9 * la + 3);
          Ody = tdepd(9 * la + 1, 9 * la + 4);
          Odz = tdepd(9 * la + 2, 9 * la + 5);
        } else {
          Odx = a[20];
          Ody = a[21];
          Odz = a[22];
        }
        Fx -= -Odz * y + Ody * z;
        Fy -= Odz * x - Odx * z;
        Fz -= -Ody * x + Odx * y;
      }
      if (lin_acc) {
        const double x0x = tdep(3), x0y = tdep(4), x0z = tdep(5), v0x = tdep(6), v0y = tdep(7), v0z = tdep(8);
        Ov = Ox * x0x + Oy * x0y + Oz * x0z;
        Fx += 2. * (Oz * v0y - Oy * v0z) + O2 * x0x - Ox * Ov;
        Fy += -2. * (Oz * v0x - Ox * v0z) + O2 * x0y - Oy * Ov;
        Fz += 2. * (Oy * v0x - Ox * v0y) + O2 * x0z - Oz * Ov;
        if (!const_freq) {
          Fx -= -Odz * x0y + Ody * x0z;
          Fy -= Odz * x0x - Odx * x0z;
          Fz -= -Ody * x0x + Odx * x0y;
        }
      }
    }
  }
  if (lin_acc) {
    Fx -= tdep(0);
    Fy -= tdep(1);
    Fz -= tdep(2);
  }
}
// The densities the friction force needs (the Dens functions of
// HernquistPotential.c, NFWPotential.c, PlummerPotential.c,
// MiyamotoNagaiPotential.c, PowerSphericalPotentialwCutoff.c,
// LogarithmicHaloPotential.c, JaffePotential.c, IsochronePotential.c,
// PowerSphericalPotential.c, BurkertPotential.c); calcDensity
// (galpy_potentials.c:380-390) sums them.  Pinned by Poisson's equation
// against each potential's forces.
inline double comp_density(const Component& C, double R, double z, double phi) {
  const double* p = C.p;
  const double r2 = R * R + z * z, r = std::sqrt(r2);
  switch (C.type) {
    case kHernquist: return p[0] * M_1_PI / 4. / p[1] / p[1] / r * std::pow(1. + r / p[1], -3.);
    case kNFW: return p[0] * M_1_PI / 4. / p[1] / p[1] / (1. + r / p[1]) / (1. + r / p[1]) / r;
    case kPlummer: {
      const double b2 = p[1] * p[1];
      return 3. * p[0] * M_1_PI / 4. * b2 * std::pow(r2 + b2, -2.5);
    }
    case kMiyamotoNagai: {
      const double a = p[1], b2 = p[2] * p[2], sqrtbz = std::sqrt(b2 + z * z);
      double asqrtbz = a + sqrtbz;
      asqrtbz *= asqrtbz;
      if (a == 0.) return 3. * p[0] * M_1_PI / 4. * b2 * std::pow(R * R + sqrtbz * sqrtbz, -2.5);
      return p[0] * M_1_PI / 4. * b2 * (a * R * R + (a + 3. * sqrtbz) * asqrtbz) * std::pow(R * R + asqrtbz, -2.5) *
             std::pow(sqrtbz, -3.);
    }
    case kPowerSphCutoff: return p[0] * std::pow(r, -p[1]) * std::exp(-r2 / p[2] / p[2]);
    case kLogHalo: {
      const double amp = p[0], q = p[1], c = p[2], om = p[3], zq = z / q, q2 = q * q;
      if (om < 1) {
        const double R2 = R * R, Rt2 = R2 * (1. - om * std::pow(std::sin(phi), 2));
        const double denom = 1. / (Rt2 + zq * zq + c), denom2 = denom * denom;
        return amp * M_1_PI / 4. *
               (2. * Rt2 / R2 * (denom - Rt2 * denom2) + denom / q2 - 2. * zq * zq * denom2 / q2 -
                om * (2. * R2 * std::pow(std::sin(2. * phi), 2) / 4. * om * denom2 + denom * std::cos(2. * phi)));
      }
      return amp * M_1_PI / 4. / q2 * ((2. * q2 + 1.) * c + R * R + (2. - 1. / q2) * z * z) /
             std::pow(R * R + zq * zq + c, 2.);
    }
    case kJaffe: return p[0] * M_1_PI / 4. / p[1] * std::pow(r * (1. + r / p[1]), -2.);
    case kIsochrone: {
      const double b = p[1], rb = std::sqrt(r2 + b * b), brbrb = (b + rb) * rb;
      return p[0] * M_1_PI / 4. * (3. * brbrb * rb - r2 * (b + 3. * rb)) * std::pow(brbrb, -3.);
    }
    case kPowerSph: return p[0] * M_1_PI / 4. * (3. - p[1]) * std::pow(r2, -0.5 * p[1]);
    case kBurkert: {
      const double x = r / p[1];
      return p[0] / (1. + x) / (1. + x * x);
    }
    default: return std::nan("");
  }
}
inline bool is_dissipative(int code) {
  return code == kNonInertial || code == kNonInertialInterp || code == kChandrasekhar || code == kFDM;
}
// The cosine integral Ci(x), x > 0 (gsl_sf_Ci, which FDMDynamicalFrictionForce.c
// calls): the power series gamma + ln x + sum_k (-x^2)^k / (2k (2k)!) for
// x <= 4, else Ci(x) = -Re E1(i x) with E1's continued fraction (A&S 5.1.22)
// by the modified Lentz method.  Pinned against scipy's sici.
inline double cos_integral(double x) {
  const double euler = 0.57721566490153286061;
  if (x <= 4.0) {
    double term = 1.0, sum = 0.0;
    for (int k = 1; k < 60; k++) {
      term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
      const double add = term / (2.0 * k);
      sum += add;
      if (std::fabs(add) < 1e-17 * std::fabs(sum)) break;
    }
    return euler + std::log(x) + sum;
  }
  // E1(z) = e^-z / (z + 1 - 1/(z + 3 - 4/(z + 5 - ...))), z = i x
  const std::complex<double> z(0.0, x);
  const double tiny = 1e-300;
  std::complex<double> b = z + 1.0, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int n = 1; n < 1000; n++) {
    const double an = -double(n) * n;
    b += 2.0;
    d = 1.0 / (an * d + b);
    c = b + an / c;
    const std::complex<double> del = c * d;
    h *= del;
    if (std::abs(del - 1.0) < 1e-16) break;
  }
  return -(h * std::exp(-z)).real();
}
// FDMDynamicalFrictionForce.c:14-99: the Chandrasekhar amplitude times
// FDMfactor / CDMfactor (when below 1, or always for a constant factor):
// CDMfactor = lnLambda Xfactor; FDMfactor the zero-dispersion form gamma +
// ln kr - Ci(kr) + sin(kr)/kr - 1 for kr = 2 (m/hbar) v r < M = v/sigma_r,
// ln(kr/M) Xfactor past 4M, a linear blend between.  args as
// Chandrasekhar's, then m/hbar, the constant factor (< 0: none).
inline double fdm_ratio(const Component& C, double R, double z, double vR, double vT, double vz) {
  const double* a = C.args;
  const double ms = a[9], rhm = a[10], gamma2 = a[11], ro = a[14], rf = a[15], mhbar = a[16], cf = a[17];
  double lnLambda = a[12];
  const double r2 = R * R + z * z, r = std::sqrt(r2), v2 = vR * vR + vT * vT + vz * vz, v = std::sqrt(v2);
  const double GMvs = ms / v / v;
  if (lnLambda < 0) {
    if (GMvs < rhm)
      lnLambda = 0.5 * std::log(1. + r2 / gamma2 / rhm / rhm);
    else
      lnLambda = 0.5 * std::log(1. + r2 / gamma2 / GMvs / GMvs);
  }
  double d_ind = (r - ro) / (rf - ro);
  d_ind = d_ind < 0 ? 0. : (d_ind > 1 ? 1. : d_ind);
  const double sr = (*C.wrapped[0].spl)[0].eval(d_ind);
  const double X = M_SQRT1_2 * v / sr;
  const double Xfactor = std::erf(X) - M_2_SQRTPI * X * std::exp(-X * X);
  const double CDMfactor = lnLambda * Xfactor;
  const double euler = 0.57721566490153286061;
  double FDMfactor;
  if (cf < 0) {
    const double kr = 2.0 * mhbar * v * r, Ms = v / sr;
    if (kr < Ms) {
      FDMfactor = euler + std::log(kr) - cos_integral(kr) + std::sin(kr) / kr - 1.0;
    } else if (kr > 4 * Ms) {
      FDMfactor = std::log(kr / Ms) * Xfactor;
    } else {
      const double C_fdm = euler + std::log(Ms) - cos_integral(Ms) + std::sin(Ms) / Ms - 1.0;
      const double C_disp = std::log(4.0) * Xfactor;
      const double mu = (2 * Ms - kr / 2) / (2 * Ms - Ms / 2);
      FDMfactor = mu * C_fdm + (1 - mu) * C_disp;
    }
  } else {
    FDMfactor = cf;
  }
  return FDMfactor / CDMfactor;
}
// ChandrasekharDynamicalFrictionForce.c:13-140: F = A v with A = -amp
// Xfactor(X) lnLambda rho / v^3, X = v / (sqrt 2 sigma_r(r)) (sigma_r a
// natural spline over the normalised radius, clamped), Xfactor = erf X -
// (2/sqrt pi) X e^-X^2, lnLambda constant or 0.5 ln(1 + r^2 / (gamma^2
// max(rhm, GM/v^2)^2)); zero inside minr.  args: amp, 8 cache slots, ms, rhm,
// gamma^2, lnLambda (< 0: variable), minr^2, ro, rf.
inline double chandrasekhar_amplitude(const Component& C, double R, double z, double phi, double vR, double vT,
                                      double vz) {
  const double* a = C.args;
  const double r2 = R * R + z * z;
  if (r2 < a[13]) return 0.;
  const double amp = a[0], ms = a[9], rhm = a[10], gamma2 = a[11], ro = a[14], rf = a[15];
  double lnLambda = a[12];
  const double r = std::sqrt(r2), v2 = vR * vR + vT * vT + vz * vz, v = std::sqrt(v2);
  if (lnLambda < 0) {
    const double GMvs = ms / v / v;
    if (GMvs < rhm)
      lnLambda = 0.5 * std::log(1. + r2 / gamma2 / rhm / rhm);
    else
      lnLambda = 0.5 * std::log(1. + r2 / gamma2 / GMvs / GMvs);
  }
  double d_ind = (r - ro) / (rf - ro);
  d_ind = d_ind < 0 ? 0. : (d_ind > 1 ? 1. : d_ind);
  const double sr = (*C.spl)[0].eval(d_ind);
  const double X = M_SQRT1_2 * v / sr;
  const double Xfactor = std::erf(X) - M_2_SQRTPI * X * std::exp(-X * X);
  double dens = 0.;
  for (int k = 0; k < C.nwrapped; k++) dens += comp_density(C.wrapped[k], R, z, phi);
  return -amp * Xfactor * lnLambda / v2 / v * dens;
}
// calcRforce / calczforce / calcphitorque with velocities
// (galpy_potentials.c, the RforceVelocity branch): the dissipative
// components' cylindrical forces at (R, z, phi, t; vR, vT, vz)
inline void diss_forces(int npot, const Component* P, double Rc, double z, double phi, double t, double vR,
                        double vT, double vz, double& FR, double& Fz, double& Fp) {
  FR = Fz = Fp = 0.;
  for (int i = 0; i < npot; i++) {
    const Component& C = P[i];
    if (!is_dissipative(C.type)) continue;
    if (C.type == kChandrasekhar || C.type == kFDM) {  // Rforce, zforce, phitorque: A vR, A vz, A vT R
      double A;
      if (C.type == kFDM) {  // FDMDynamicalFrictionForceAmplitude (:82-99)
        if (Rc * Rc + z * z < C.args[13]) continue;
        A = chandrasekhar_amplitude(C.wrapped[0], Rc, z, phi, vR, vT, vz);
        const double ratio = fdm_ratio(C, Rc, z, vR, vT, vz);
        if (ratio < 1. || C.args[17] >= 0) A *= ratio;
      } else {
        A = chandrasekhar_amplitude(C, Rc, z, phi, vR, vT, vz);
      }
      FR += A * vR;
      Fz += A * vz;
      Fp += A * vT * Rc;
      continue;
    }
    // cyl_to_rect, cyl_to_rect_vec (bovy_coords.c)
    const double cp = std::cos(phi), sp = std::sin(phi);
    const double x = Rc * cp, y = Rc * sp, vx = vR * cp - vT * sp, vy = vR * sp + vT * cp;
    double Fx, Fy, Fzz;
    noninertial_xyz(C, x, y, z, vx, vy, vz, t, Fx, Fy, Fzz);
    const double amp = C.p[0];
    FR += amp * (cp * Fx + sp * Fy);
    Fp += amp * Rc * (-sp * Fx + cp * Fy);
    Fz += amp * Fzz;
  }
}
inline bool has_dissipative(int npot, const Component* P) {
  for (int i = 0; i < npot; i++)
    if (is_dissipative(P[i].type)) return true;
  return false;
}

// evalRectForce: integrateFullOrbit.c:1307-1325
template <class R>
void rect_force(int npot, const Component* P, R t, const R* q, R* a) {
  R x = q[0], y = q[1], z = q[2];
  R Rc = sqrt(x * x + y * y);
  R phi = acos(x / Rc);
  R sinphi = y / Rc;
  R cosphi = x / Rc;
  if (y < R(0.)) phi = R(2. * M_PI) - phi;
  R FR, Fz, Fphi;
  cyl_forces(npot, P, Rc, z, phi, t, FR, Fz, Fphi);
  a[0] = cosphi * FR - R(1.) / Rc * sinphi * Fphi;
  a[1] = sinphi * FR + R(1.) / Rc * cosphi * Fphi;
  a[2] = Fz;
}

// evalRectDeriv: integrateFullOrbit.c:1326-1351 (vR, vT only feed dissipative
// forces, none of which are in scope, so they are not formed)
template <class R>
void rect_deriv(int npot, const Component* P, R t, const R* q, R* a) {
  a[0] = q[3]; a[1] = q[4]; a[2] = q[5];
  rect_force(npot, P, t, q, a + 3);
  if (has_dissipative(npot, P)) {  // the velocity-dependent forces (vR, vT as evalRectDeriv forms them)
    const double x = dval(q[0]), y = dval(q[1]), Rc = std::sqrt(x * x + y * y);
    double phi = std::acos(x / Rc);
    const double sinphi = y / Rc, cosphi = x / Rc;
    if (y < 0.) phi = 2. * M_PI - phi;
    const double vR = dval(q[3]) * cosphi + dval(q[4]) * sinphi;
    const double vT = -dval(q[3]) * sinphi + dval(q[4]) * cosphi;
    double FR, Fz, Fp;
    diss_forces(npot, P, Rc, dval(q[2]), phi, dval(t), vR, vT, dval(q[5]), FR, Fz, Fp);
    a[3] += R(cosphi * FR - 1. / Rc * sinphi * Fp);
    a[4] += R(sinphi * FR + 1. / Rc * cosphi * Fp);
    a[5] += R(Fz);
  }
}

// bovy_coords.c:101-115
template <class R> void cyl_to_rect(R* v) {
  R Rc = v[0], phi = v[5];
  R cp = cos(phi), sp = sin(phi);
  R vR = v[1], vT = v[2];
  v[0] = Rc * cp;
  v[1] = Rc * sp;
  v[2] = v[3];
  v[5] = v[4];
  v[3] = vR * cp - vT * sp;
  v[4] = vR * sp + vT * cp;
}
// bovy_coords.c:57-69 (planar: (R,vR,vT,phi) -> (x,y,vx,vy))
template <class R> void polar_to_rect(R* v) {
  R Rc = v[0], phi = v[3];
  R cp = cos(phi), sp = sin(phi);
  R vR = v[1], vT = v[2];
  v[0] = Rc * cp;
  v[1] = Rc * sp;
  v[2] = vR * cp - vT * sp;
  v[3] = vR * sp + vT * cp;
}
// bovy_coords.c:79-91 (planar: (x,y,vx,vy) -> (R,vR,vT,phi))
template <class R> void rect_to_polar(R* v) {
  R x = v[0], y = v[1], vx = v[2], vy = v[3];
  v[3] = atan2(y, x);
  R cp = cos(v[3]), sp = sin(v[3]);
  v[0] = sqrt(x * x + y * y);
  v[1] = vx * cp + vy * sp;
  v[2] = -vx * sp + vy * cp;
}
// bovy_coords.c:125-139
template <class R> void rect_to_cyl(R* v) {
  R x = v[0], y = v[1], vx = v[3], vy = v[4];
  v[3] = v[2];
  v[4] = v[5];
  v[5] = atan2(y, x);
  R cp = cos(v[5]), sp = sin(v[5]);
  v[0] = sqrt(x * x + y * y);
  v[1] = vx * cp + vy * sp;
  v[2] = -vx * sp + vy * cp;
}

}  // namespace orc
