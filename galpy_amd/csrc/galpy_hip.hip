// MI355X orbit-integration engine: kernels and C-ABI entry points.
//
// One orbit per lane, FP64 throughout, potential table in device memory read
// with wavefront-uniform (scalar) loads.  The C-ABI mirrors galpy's
// integrateFullOrbit (galpy/orbit/orbit_c_ext/integrateFullOrbit.c:1046-1150):
// cylindrical initial conditions in, cylindrical rows out, one call for all
// orbits.  Declarations and argument meaning: include/galpy_hip.h.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <signal.h>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <vector>
#include <algorithm>
#include <type_traits>
#include <initializer_list>

#include "../../include/galpy_hip.h"
#include "batch_dev.h"
#include "adjoint_dev.h"
#include "dxdv_dev.h"
#include "rk_dev.h"
#include "ias15_dev.h"

#define GH_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

constexpr int kBlock = 64;  // one wavefront per workgroup

thread_local char g_errbuf[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
  va_end(ap);
}

#define GH_CHECK(expr)                                                          \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                __LINE__);                                                      \
      return -1;                                                                \
    }                                                                           \
  } while (0)

// ---------------------------------------------------------------------------
// Interrupt flag: fine-grained pinned host memory that running kernels poll.
// ---------------------------------------------------------------------------
volatile int* g_irq_host = nullptr;
int* g_irq_dev = nullptr;

int* irq_device_ptr() {
  if (g_irq_dev) return g_irq_dev;
  void* p = nullptr;
  if (hipHostMalloc(&p, sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return nullptr;
  g_irq_host = static_cast<volatile int*>(p);
  *g_irq_host = 0;
  void* d = nullptr;
  if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) return nullptr;
  g_irq_dev = static_cast<int*>(d);
  return g_irq_dev;
}

// ---------------------------------------------------------------------------
// Host decode of galpy's packed potential into the device table.
// ---------------------------------------------------------------------------
// Arguments per component: integrateFullOrbit.c:80-345 for full orbits,
// integratePlanarOrbit.c:64-223 for planar ones (whose
// PowerSphericalPotentialwCutoff has no cache slots: 3 arguments, not 5).
int nargs_of(int code, bool planar) {
  switch (code) {
    case GH_LOGHALO: return 4;
    case GH_DEHNENBAR: return 6;
    case GH_MIYAMOTONAGAI: return 3;
    case GH_HERNQUIST: return 2;
    case GH_NFW: return 2;
    case GH_POWERSPHCUTOFF: return planar ? 3 : 5;
    // integrateFullOrbit.c:135-603 / integratePlanarOrbit.c:136-418
    case GH_POWERSPH: case GH_JAFFE: case GH_ISOCHRONE: case GH_PLUMMER: case GH_PSEUDOISO:
    case GH_KUZMINDISK: case GH_BURKERT: case GH_DEHNENCORESPH: return 2;
    case GH_DEHNENSPH: case GH_HOMOGSPHERE: return 3;
    case GH_FLATPOWER: return planar ? 3 : 4;  // planar: amp, alpha, core^2
    case GH_NULLPOT: return 0;
    case GH_KKSTAECKEL: return 3;  // amp, ac, Delta
    case GH_ELLIPDISK: return planar ? 6 : -1;  // planar only
    case GH_LOPSIDED: return planar ? 4 : -1;   // planar only
    case GH_STEADYSPIRAL: return planar ? 8 : -1;  // planar only
    case GH_SOFTNEEDLEBAR: return 13;  // amp, a, b, c^2, pa, omegab + 7 cache slots
    // wrappers' own arguments, read after the wrapped components
    // (integrateFullOrbit.c:821-895, integratePlanarOrbit.c:591-663)
    case GH_DEHNENSMOOTHWRAP: return 4;   // amp, tform, tsteady, grow
    case GH_SOLIDBODYROTWRAP: return 3;   // amp, omega, pa
    case GH_COROTATINGROTWRAP: return 5;  // amp, vpo, beta, pa, to
    case GH_GAUSSIANAMPWRAP: return 3;    // amp, to, sigma^2
    default: return -1;
  }
}

// smallest x >= a+1 with 2 e^-x x^(a-1) < 2^-60 Gamma(a): beyond it the upper
// incomplete gamma is below 1/128 ulp of Gamma(a)
double incgamma_xbig(double a, double gamma_a) {
  double x = std::max(a + 1.0, 2.0 * (a - 1.0) + 1.0);
  const double lim = std::ldexp(gamma_a, -60);
  while (2.0 * std::exp(-x + (a - 1.0) * std::log(x)) >= lim) x += 0.25;
  return x;
}

int fill_series(double a, double* ser) {
  // ser[n] = 1/(a (a+1) ... (a+n)); the last term at x = a+1 must be negligible
  double v = 1.0 / a;
  for (int n = 0; n < GH_IG_NSER; n++) {
    ser[n] = v;
    v /= (a + n + 1.0);
  }
  const double last = ser[GH_IG_NSER - 1] * std::pow(a + 1.0, GH_IG_NSER - 1);
  return (last < 1e-18 * ser[0]) ? 0 : -1;
}

// Upper incomplete gamma Gamma(a,x), x >= a+1, in long double: Legendre's
// continued fraction by the modified Lentz method (A&S 6.5.31).  Used only on
// the host to fit the device tables.
long double upper_gamma_ld(long double a, long double x) {
  const long double tiny = 1e-300L;
  long double b = x + 1.0L - a, c = 1.0L / tiny, d = 1.0L / b, h = d;
  for (int i = 1; i < 100000; i++) {
    const long double an = -i * (i - a);
    b += 2.0L;
    d = an * d + b;
    if (fabsl(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabsl(c) < tiny) c = tiny;
    d = 1.0L / d;
    const long double del = d * c;
    h *= del;
    if (fabsl(del - 1.0L) < 1e-19L) break;
  }
  return expl(a * logl(x) - x) * h;
}

// Fit the piecewise polynomial table of Gamma(a,x) on the octaves covering
// [a+1, xbig): GH_TAB_PER_OCT segments per octave, interpolation at the
// GH_TAB_NC Chebyshev nodes of each segment, stored as monomial coefficients
// in the local coordinate t in [-1, 1).  Returns the first octave's frexp
// exponent, or a large value if the range needs more than GH_TAB_MAXSEG.
int fit_upper_gamma_table(double a, double xbig, double tab[GH_TAB_MAXSEG][GH_TAB_NC]) {
  int e0, e1;
  std::frexp(a + 1.0, &e0);
  std::frexp(xbig, &e1);
  if ((e1 - e0 + 1) * GH_TAB_PER_OCT > GH_TAB_MAXSEG) return 1 << 20;
  const int N = GH_TAB_NC;
  for (int e = e0; e <= e1; e++) {
    for (int k = 0; k < GH_TAB_PER_OCT; k++) {
      long double f[GH_TAB_NC], cheb[GH_TAB_NC];
      for (int j = 0; j < N; j++) {
        const long double tj = cosl(M_PI * (j + 0.5L) / N);
        const long double m = 0.5L + (k + (tj + 1.0L) / 2.0L) / (2.0L * GH_TAB_PER_OCT);
        f[j] = upper_gamma_ld(a, ldexpl(m, e));
      }
      for (int n = 0; n < N; n++) {
        long double acc = 0.0L;
        for (int j = 0; j < N; j++) acc += f[j] * cosl(M_PI * n * (j + 0.5L) / N);
        cheb[n] = acc * 2.0L / N;
      }
      cheb[0] *= 0.5L;
      // Chebyshev -> monomial: T_{n+1} = 2 t T_n - T_{n-1}
      long double Tm1[GH_TAB_NC] = {0}, T0[GH_TAB_NC] = {0}, mono[GH_TAB_NC] = {0};
      T0[0] = 1.0L;
      for (int n = 0; n < N; n++) {
        for (int q = 0; q < N; q++) mono[q] += cheb[n] * T0[q];
        long double Tn1[GH_TAB_NC] = {0};
        for (int q = 0; q < N; q++) {
          if (q + 1 < N) Tn1[q + 1] += 2.0L * T0[q];
          Tn1[q] -= (n == 0) ? 0.0L : Tm1[q];
        }
        if (n == 0) {  // T_1 = t
          for (int q = 0; q < N; q++) Tn1[q] = 0.0L;
          Tn1[1] = 1.0L;
        }
        for (int q = 0; q < N; q++) {
          Tm1[q] = T0[q];
          T0[q] = Tn1[q];
        }
      }
      const int seg = (e - e0) * GH_TAB_PER_OCT + k;
      for (int q = 0; q < N; q++) tab[seg][q] = (double)mono[q];
    }
  }
  return e0;
}

// One (leaf) component's arguments and precomputed constants.
int decode_component(int code, const double*& pot_args, GhPotTable* T, int i, int& nslots,
                     bool planar) {
  const int na = nargs_of(code, planar);
  if (na < 0 || code < 0) {
    set_error("pot_type %d (component %d) has no HIP kernel", code, i);
    return -1;
  }
  {
    GhComponent& C = T->comp[i];
    C.type = code;
    for (int j = 0; j < na; j++) C.p[j] = *pot_args++;
    if (code == GH_FLATPOWER && planar) {  // full layout, q^2 = 1 (z = 0 in the plane)
      C.p[3] = C.p[2];
      C.p[2] = 1.0;
    }
    switch (code) {
      case GH_POWERSPHCUTOFF: {
        const double amp = C.p[0], alpha = C.p[1], rc = C.p[2];
        const double a = 1.5 - 0.5 * alpha, a2 = 1.0 - 0.5 * alpha;
        if (!(a > 0.0)) {
          set_error("PowerSphericalPotentialwCutoff alpha=%g needs alpha < 3", alpha);
          return -1;
        }
        C.c[0] = a;
        C.c[1] = std::tgamma(a);
        C.c[2] = amp * 2.0 * M_PI * std::pow(rc, 3.0 - alpha);
        C.c[3] = 1.0 / (rc * rc);
        C.c[4] = incgamma_xbig(a, C.c[1]);
        C.c[9] = std::ldexp(C.c[1], -58);  // continued-fraction stopping tolerance
        if (fill_series(a, T->ser[i][0]) != 0) {
          set_error("incomplete-gamma series too short for a=%g", a);
          return -1;
        }
        if (a2 > 0.0) {
          C.c[5] = a2;
          C.c[6] = std::tgamma(a2);
          C.c[7] = incgamma_xbig(a2, C.c[6]);
          C.c[10] = std::ldexp(C.c[6], -58);
          fill_series(a2, T->ser[i][1]);
        } else {  // galpy's Eval calls gsl_sf_gamma(1 - alpha/2): undefined here
          C.c[5] = a2;
          C.c[6] = NAN;
          C.c[7] = INFINITY;
          for (int n = 0; n < GH_IG_NSER; n++) T->ser[i][1][n] = NAN;
        }
        C.c[8] = 1.0 / rc;
        if (nslots >= GH_TAB_SLOTS) {
          set_error("at most %d PowerSphericalPotentialwCutoff components per potential", GH_TAB_SLOTS);
          return -1;
        }
        C.c[11] = nslots;
        C.c[12] = fit_upper_gamma_table(a, C.c[4], T->tab[nslots]);
        if (C.c[12] > 1000) {
          set_error("incomplete-gamma table range too wide for a=%g", a);
          return -1;
        }
        nslots++;
        break;
      }
      case GH_MIYAMOTONAGAI: C.c[0] = C.p[2] * C.p[2]; break;
      case GH_NFW: C.c[0] = 1.0 / C.p[1]; break;
      case GH_HERNQUIST: C.c[0] = 0.5 * C.p[0]; break;
      case GH_LOGHALO:
        C.c[0] = 1.0 / (C.p[1] * C.p[1]);
        C.nonaxi = C.p[3] < 1.0;
        break;
      case GH_DEHNENBAR:
        C.c[0] = 1.0 / C.p[3];
        C.nonaxi = 1;
        break;
      case GH_ISOCHRONE:
      case GH_PLUMMER: C.c[0] = C.p[1] * C.p[1]; break;  // b^2
      case GH_FLATPOWER: C.c[0] = 1.0 / C.p[2]; break;   // 1/q^2
      case GH_SOFTNEEDLEBAR: C.nonaxi = 1; break;
      case GH_STEADYSPIRAL: C.nonaxi = 1; break;
      case GH_LOPSIDED:
        C.c[0] = cos(C.p[3]);
        C.c[1] = sin(C.p[3]);
        C.nonaxi = 1;
        break;
      case GH_ELLIPDISK:
        C.c[0] = cos(2.0 * C.p[5]);
        C.c[1] = sin(2.0 * C.p[5]);
        C.nonaxi = 1;
        break;
      case GH_KKSTAECKEL:  // gamma, alpha, Delta^2
        C.c[2] = C.p[2] * C.p[2];
        C.c[0] = C.c[2] / (1.0 - C.p[1] * C.p[1]);
        C.c[1] = C.c[0] - C.c[2];
        break;
      case GH_PSEUDOISO:
      case GH_BURKERT: C.c[0] = 1.0 / C.p[1]; break;  // 1/a
      case GH_DEHNENSPH:
        if (!(C.p[2] >= 0.0 && C.p[2] < 3.0)) {
          set_error("DehnenSphericalPotential alpha=%g outside [0, 3)", C.p[2]);
          return -1;
        }
        C.c[0] = 1.0 / (3.0 - C.p[2]);
        break;
      default: break;
    }
    T->nonaxi |= C.nonaxi;
  }
  return 0;
}

// n components in galpy's order (parse_leapFuncArgs_Full,
// integrateFullOrbit.c:55-1045): a wrapper's code is followed by its wrapped
// components' codes; its arguments are nwrapped, the wrapped components'
// arguments, then its own (:1010-1031).  Leaves go to T->comp in the order
// they appear (galpy's summation order); every leaf lists the wrappers
// around it, outermost first.
int decode_components(int n, const int*& pot_type, const double*& pot_args, GhPotTable* T,
                      int& nleaf, int& nslots, int depth, bool planar) {
  for (int k = 0; k < n; k++) {
    const int code = *pot_type++;
    if (code >= 0) {
      if (nleaf >= GH_MAX_POT) {
        set_error("more than %d components (wrapped ones included)", GH_MAX_POT);
        return -1;
      }
      if (decode_component(code, pot_args, T, nleaf, nslots, planar) != 0) return -1;
      nleaf++;
      continue;
    }
    const int na = nargs_of(code, planar);
    if (na < 0) {
      set_error("wrapper pot_type %d has no HIP kernel", code);
      return -1;
    }
    if (depth >= GH_WRAP_DEPTH || T->nwraps >= GH_MAX_WRAP) {
      set_error("wrappers nested more than %d deep or more than %d in total", GH_WRAP_DEPTH,
                GH_MAX_WRAP);
      return -1;
    }
    const double nwd = *pot_args++;
    const int nw = (int)nwd;
    if (!(nwd >= 1.0) || nw != nwd) {
      set_error("wrapper pot_type %d wraps %g components", code, nwd);
      return -1;
    }
    const int first = nleaf;
    if (decode_components(nw, pot_type, pot_args, T, nleaf, nslots, depth + 1, planar) != 0)
      return -1;
    const int w = T->nwraps++;
    GhWrap& W = T->wrap[w];
    W.kind = code;
    for (int j = 0; j < na; j++) W.p[j] = *pot_args++;
    if (code == GH_GAUSSIANAMPWRAP && !(W.p[2] > 0.0)) {
      set_error("GaussianAmplitudeWrapperPotential sigma^2=%g must be positive", W.p[2]);
      return -1;
    }
    for (int l = first; l < nleaf; l++) {  // this wrapper encloses the ones already listed
      GhComponent& C = T->comp[l];
      for (int j = C.nwrap; j > 0; j--) C.wrap[j] = C.wrap[j - 1];
      C.wrap[0] = w;
      C.nwrap++;
    }
    T->has_wrap = 1;
  }
  return 0;
}

int decode_potential(int npot, const int* pot_type, const double* pot_args, GhPotTable* T,
                     bool planar = false) {
  std::memset(T, 0, sizeof(*T));
  if (npot < 1 || npot > GH_MAX_POT) {
    set_error("npot=%d outside [1, %d]", npot, GH_MAX_POT);
    return -1;
  }
  int nleaf = 0, nslots = 0;
  const int* pt = pot_type;
  const double* pa = pot_args;
  if (decode_components(npot, pt, pa, T, nleaf, nslots, 0, planar) != 0) return -1;
  T->npot = nleaf;
  return 0;
}

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

// wez_ias15 (wez_ias15.c:137-425), one orbit per lane.  The first step is
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

template <class Spec, class V>
__global__ void __launch_bounds__(kBlock)
gh_ias15_kernel(int nobj, double* __restrict__ yo, int nt, const double* __restrict__ t,
                int indiv_t, double dt_in, const GhPotTable* __restrict__ T,
                double* __restrict__ result, int* __restrict__ err, double* __restrict__ stats,
                const volatile int* cancel) {
  const long orbit = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (orbit >= nobj) return;
  constexpr int N = GhPhase<V>::N;
  constexpr int D = N / 2;
  const GhForce<Spec> F{T};
  const double* tt = t + (indiv_t ? orbit * nt : 0);
  double x[D], v[D];
  {
    double w[N], y[N];
#pragma unroll
    for (int k = 0; k < N; k++) w[k] = yo[orbit * N + k];
    GhPhase<V>::to_rect(w, y);
#pragma unroll
    for (int k = 0; k < N; k++) yo[orbit * N + k] = y[k];  // galpy converts yo in place
#pragma unroll
    for (int i = 0; i < D; i++) {
      x[i] = y[i];
      v[i] = y[D + i];
    }
  }
  double* out = result + orbit * (long)nt * N;
  auto put = [&](int row, const double* q, const double* p) {
    double y[N];
#pragma unroll
    for (int i = 0; i < D; i++) {
      y[i] = q[i];
      y[D + i] = p[i];
    }
    GhPhase<V>::store(out + (long)N * row, GhPhase<V>::q(y), GhPhase<V>::p(y));
  };
  put(0, x, v);
  double B[7][D], E[7][D], G[7][D], F0[D];
#pragma unroll
  for (int k = 0; k < 7; k++)
#pragma unroll
    for (int i = 0; i < D; i++) B[k][i] = E[k][i] = 0.0;
  const double init_dt = tt[1] - tt[0];
  double dt = dt_in == GH_DT_SENTINEL ? init_dt : dt_in;
  double to = tt[0];
  double time_remaining = fabs((double)nt * dt);
  int row = 1, e = 0;
  long nacc = 0, nrej = 0;
  while (time_remaining > 0.0) {
    if (((nacc + nrej) & 63) == 0) {
      if (cancel != nullptr && *cancel) {
        e = -10;
        break;
      }
      if (nacc + nrej >= kIas15MaxAttempts) {
        e = -3;
        break;
      }
    }
    const double dt_temp = dt;
    const double to_temp = to + dt_temp;
    const double dt_required =
        gh_ias15_attempt<GhForce<Spec>, V, D>(F, x, v, dt_temp, to_temp, F0, B, G);
    if (fabs(dt_temp) > fabs(dt_required)) {
      dt = dt_required;
      nrej++;
      continue;
    }
    nacc++;
    time_remaining -= fabs(dt);
    while (row < nt && (init_dt > 0.0 ? (to < tt[row] && tt[row] <= to_temp)
                                      : (to > tt[row] && tt[row] >= to_temp))) {
      const double hs = fabs(tt[row] - to) / fabs(to_temp - to);
      double xs[D], vs[D];
      gh_ias15_position<D>(xs, x, v, hs, dt_temp, F0, B);
      gh_ias15_velocity<D>(vs, v, hs, dt_temp, F0, B);
      put(row, xs, vs);
      row++;
    }
    gh_ias15_advance<D>(x, v, dt_temp, F0, B, E);
    to = to_temp;
    dt = dt_required;
  }
  if (e == 0) {
    double nan[D];
#pragma unroll
    for (int i = 0; i < D; i++) nan[i] = __builtin_nan("");
    for (; row < nt; row++) put(row, nan, nan);
  }
  err[orbit] = e;
  if (stats) {
    stats[orbit * 3 + 0] = (double)nacc;
    stats[orbit * 3 + 1] = (double)nrej;
    stats[orbit * 3 + 2] = dt;
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
    A14 = gh_accel(T, t_old + c14 * h,
                   gh_rkn_q14(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                              gh_job_vec<V>(qd, slot, 3 * D), A));
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  {
    V A[17];
    gh_job_load_acc(qd, slot, A);
    A[14] = A14;
    A15 = gh_accel(T, t_old + c15 * h,
                   gh_rkn_q15(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                              gh_job_vec<V>(qd, slot, 3 * D), A));
  }
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  {
    V A[17];
    gh_job_load_acc(qd, slot, A);
    A[14] = A14;
    A[15] = A15;
    A16 = gh_accel(T, t_old + c16 * h,
                   gh_rkn_q16(h, gh_job_vec<V>(qd, slot, 0), gh_job_vec<V>(qd, slot, D),
                              gh_job_vec<V>(qd, slot, 3 * D), A));
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

// Energy / Jacobi drift along stored cylindrical rows.
template <int N>
__global__ void __launch_bounds__(kBlock)
gh_energy_kernel(int nobj, int nt, const double* __restrict__ rows, const double* __restrict__ t,
                 int indiv_t, const GhPotTable* __restrict__ T, double omega,
                 double* __restrict__ e0, double* __restrict__ demax) {
  // N = 6: rows (R,vR,vT,z,vz,phi); N = 4: planar rows (R,vR,vT,phi), z = vz = 0
  const long orbit = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (orbit >= nobj) return;
  const double* tt = t + (indiv_t ? orbit * nt : 0);
  const double* r = rows + orbit * (long)nt * N;
  double E0 = 0.0, dmax = 0.0;
  for (int j = 0; j < nt; j++) {
    const double R = r[(long)N * j + 0], vR = r[(long)N * j + 1], vT = r[(long)N * j + 2];
    const double z = N == 6 ? r[(long)N * j + 3] : 0.0, vz = N == 6 ? r[(long)N * j + 4] : 0.0;
    const double phi = r[(long)N * j + N - 1];
    double sp, cp;
    sincos(phi, &sp, &cp);
    const double E = gh_potential(T, R, z, cp, sp, tt[j]) + 0.5 * (vR * vR + vT * vT + vz * vz) -
                     omega * R * vT;
    if (j == 0)
      E0 = E;
    else
      dmax = fmax(dmax, fabs((E - E0) / E0));
  }
  e0[orbit] = E0;
  demax[orbit] = dmax;
}

// Post-processing of stored trajectories (Orbits.py rperi :3733-3778,
// rap :3686-3731, e :3642-3684, zmax :3966-4010, E :2902-3115, Jacobi
// :3298-3360): r = sqrt(R^2 + z^2) over all outputs, rperi = min r,
// rap = max r, e = (rap - rperi) / (rap + rperi), zmax = max |z|; E at every
// output, and E - omega R vT.  One wavefront per orbit: lane k takes rows
// k, k + 64, ..., so a wavefront reads 64 consecutive rows (64 N doubles) at
// a time, coalesced, and E rows are written likewise; min/max/|z| reduce
// across the wavefront by shuffles.  No FMA contraction in r, so rperi, rap
// and zmax are exactly numpy's of the same rows.
constexpr int kSummaryOrbitsPerBlock = 4;

template <int N>
__global__ void __launch_bounds__(64 * kSummaryOrbitsPerBlock)
gh_summary_kernel(int nobj, int nt, const double* __restrict__ rows, const double* __restrict__ t,
                  int indiv_t, const GhPotTable* __restrict__ T, double omega,
                  double* __restrict__ rperi, double* __restrict__ rap, double* __restrict__ zmax,
                  double* __restrict__ ecc, double* __restrict__ E, double* __restrict__ EJ) {
#pragma clang fp contract(off)
  const long orbit = (long)blockIdx.x * kSummaryOrbitsPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (orbit >= nobj) return;  // uniform across the wavefront
  const double* __restrict__ r = rows + orbit * (long)nt * N;
  const double* __restrict__ tt = t + (indiv_t ? orbit * nt : 0);
  double rmin = INFINITY, rmax = -INFINITY, zm = 0.0;
  const bool energies = (E != nullptr) || (EJ != nullptr);
  for (int j = lane; j < nt; j += 64) {
    const double* __restrict__ row = r + (long)N * j;
    const double R = row[0];
    const double z = (N == 6) ? row[3] : 0.0;
    const double rs = sqrt(R * R + z * z);
    rmin = fmin(rmin, rs);
    rmax = fmax(rmax, rs);
    zm = fmax(zm, fabs(z));
    if (energies) {
      const double vR = row[1], vT = row[2], vz = (N == 6) ? row[4] : 0.0;
      double sp, cp;
      sincos(row[N - 1], &sp, &cp);
      const double En = gh_potential(T, R, z, cp, sp, tt[j]) + vR * vR / 2.0 + vT * vT / 2.0 +
                        vz * vz / 2.0;
      if (E) E[orbit * (long)nt + j] = En;
      if (EJ) EJ[orbit * (long)nt + j] = En - omega * R * vT;
    }
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    rmin = fmin(rmin, __shfl_xor(rmin, m));
    rmax = fmax(rmax, __shfl_xor(rmax, m));
    zm = fmax(zm, __shfl_xor(zm, m));
  }
  if (lane == 0) {
    if (rperi) rperi[orbit] = rmin;
    if (rap) rap[orbit] = rmax;
    if (zmax) zmax[orbit] = zm;
    if (ecc) ecc[orbit] = (rmax - rmin) / (rmax + rmin);
  }
}

__global__ void __launch_bounds__(kBlock)
gh_eval_kernel(int n, const double* __restrict__ R, const double* __restrict__ z,
               const double* __restrict__ phi, const double* __restrict__ t,
               const GhPotTable* __restrict__ T, double* __restrict__ pot,
               double* __restrict__ FR, double* __restrict__ Fz, double* __restrict__ Fphi) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double sp, cp;
  sincos(phi[i], &sp, &cp);
  if (pot) pot[i] = gh_potential(T, R[i], z[i], cp, sp, t[i]);
  // the wrapper-aware sum serves every table (evaluation is not a hot path)
  const GhCyl F = gh_cyl_forces<GhSpecWrapped>(T, R[i], R[i] * R[i], z[i], cp, sp, t[i]);
  if (FR) FR[i] = F.FR;
  if (Fz) Fz[i] = F.Fz;
  if (Fphi) Fphi[i] = F.Fphi;
}

// Work counters for the persistent kernels: one small device array per
// device, allocated on first use and kept for the life of the process; each
// launch takes the next slot, so concurrent launches on different streams
// never share a counter.
constexpr int kCounterSlots = 4096;
unsigned long long* g_counters[64] = {nullptr};
unsigned g_counter_next = 0;

unsigned long long* next_counter(int dev) {
  if (dev < 0 || dev >= 64) return nullptr;
  if (!g_counters[dev]) {
    void* p = nullptr;
    if (hipMalloc(&p, kCounterSlots * sizeof(unsigned long long)) != hipSuccess) return nullptr;
    g_counters[dev] = static_cast<unsigned long long*>(p);
  }
  const unsigned slot = __atomic_fetch_add(&g_counter_next, 1u, __ATOMIC_RELAXED) % kCounterSlots;
  return g_counters[dev] + slot;
}

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
int dop_waves_forced() {
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

// Host handle of an uploaded potential: the device table plus the component
// codes, so launches can pick a kernel specialised for the composition.
struct GhPotHandle {
  GhPotTable* dev;
  int npot;
  int types[GH_MAX_POT];
  int nonaxi[GH_MAX_POT];  // component exerts a phi torque (bar, triaxial halo)
  int has_wrap;            // wrapped components: run-time dispatch, no Hessian
};

void fill_handle(const GhPotTable& table, GhPotHandle* h) {
  h->npot = table.npot;
  h->has_wrap = table.has_wrap;
  for (int i = 0; i < GH_MAX_POT; i++) {
    h->types[i] = table.comp[i].type;
    h->nonaxi[i] = table.comp[i].nonaxi;
  }
}

// Components with a Hessian on the device (the adjoint's): time-independent
// and axisymmetric -- every type but the bar and a triaxial logarithmic halo.
bool hessian_capable(const GhPotHandle* h) {
  if (h->has_wrap) return false;  // time-dependent (and rotating) wrappers
  for (int i = 0; i < h->npot; i++) {
    if (h->types[i] == GH_DEHNENBAR || h->types[i] == GH_KKSTAECKEL || h->nonaxi[i]) return false;
  }
  return true;
}

// DOP853 launch at WPE waves per SIMD: optional pilot + longest-first
// order, then the persistent kernel.  The pilot's cost keys and the order
// live in one stream-ordered temporary allocation, freed after the launch.
template <class Spec, int WPE, class V>
int launch_dop853(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                  const GhPotTable* d_table, double rtol, double atol, double* d_result,
                  int* d_err, double* d_stats, const int* cancel, int dev, hipStream_t stream) {
  const int resident = dop_resident_blocks<Spec, WPE, V>(dev);
  if (resident <= 0) {
    set_error("could not query DOP853 occupancy on device %d", dev);
    return -1;
  }
  const unsigned nblocks =
      (unsigned)std::min<long>((long)((nobj + kBlock - 1) / kBlock), (long)resident);
  const unsigned long long lanes = (unsigned long long)nblocks * kBlock;
  unsigned long long* counter = next_counter(dev);
  if (counter == nullptr) {
    set_error("could not set up the DOP853 work counter on device %d", dev);
    return -1;
  }
  // One stream-ordered temporary allocation: the dense-output overflow area
  // (kOverflowBytes per workgroup) and, when some lane gets more than one
  // orbit, the pilot's cost keys and the longest-first order (otherwise all
  // orbits start at once and the order is moot).
  const bool ordered = !indiv_t && (unsigned long long)nobj > lanes && dop_order_enabled();
  size_t sort_bytes = 0;
  hipError_t e = hipSuccess;
  if (ordered)
    e = hipcub::DeviceRadixSort::SortPairsDescending(nullptr, sort_bytes, (const unsigned*)nullptr,
                                                     (unsigned*)nullptr, (const int*)nullptr,
                                                     (int*)nullptr, nobj, 0, 32, stream);
  const size_t ovf = ((size_t)nblocks * kOverflowBytes<V> + 255) / 256 * 256;
  const size_t n4 = ordered ? ((size_t)nobj * 4 + 255) / 256 * 256 : 0;
  void* work = nullptr;
  if (e == hipSuccess)
    e = hipMallocAsync(&work, ovf + (ordered ? 4 * n4 + 256 + sort_bytes : 0), stream);
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
  if (work) {
    const hipError_t ef = hipFreeAsync(work, stream);
    if (e == hipSuccess) e = ef;
  }
  if (e != hipSuccess) {
    set_error("DOP853 launch failed: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

// Planar DOP853 runs at 2 waves per SIMD only (its state is two thirds of a
// full orbit's, so it never needs the single-wave instance).
template <class Spec, class V>
int launch_spec(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                const GhPotTable* d_table, double dt, double rtol, double atol, double* d_result,
                int* d_err, int odeint_type, double* d_stats, hipStream_t stream) {
  const int* cancel = irq_device_ptr();
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
    case 1:  // RK4 and RK6: plain orbits
    case 2:
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        if (odeint_type == 1)
          hipLaunchKernelGGL((gh_rk_kernel<4, Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                             d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                             cancel);
        else
          hipLaunchKernelGGL((gh_rk_kernel<6, Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                             d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                             cancel);
        break;
      } else {
        set_error("rk4 and rk6 have no tangent-map (integrate_dxdv) stepper here");
        return -1;
      }
    case 5:  // DOPR54: plain orbits
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        hipLaunchKernelGGL((gh_dopr54_kernel<Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                           d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err, d_stats,
                           cancel);
        break;
      } else {
        set_error("dopr54 has no tangent-map (integrate_dxdv) stepper here");
        return -1;
      }
    case 7:  // IAS15: plain orbits
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        hipLaunchKernelGGL((gh_ias15_kernel<Spec, V>), grid, block, 0, stream, nobj, d_yo, nt,
                           d_t, indiv_t, dt, d_table, d_result, d_err, d_stats, cancel);
        break;
      } else {
        set_error("ias15 has no tangent-map (integrate_dxdv) stepper here");
        return -1;
      }
    case 4:  // symplec6: plain orbits (galpy has no symplec6 tangent map)
      if constexpr (std::is_same<V, GhVec3>::value || std::is_same<V, GhVec2>::value) {
        hipLaunchKernelGGL((gh_symplectic_kernel<4, Spec, V>), grid, block, 0, stream, nobj,
                           d_yo, nt, d_t, indiv_t, dt, rtol, atol, d_table, d_result, d_err,
                           d_stats, cancel);
        break;
      } else {
        set_error("symplec6 has no tangent-map (integrate_dxdv) stepper");
        return -1;
      }
    case 6: {
      int dev = 0;
      GH_CHECK(hipGetDevice(&dev));
      if constexpr (std::is_same<V, GhVec3>::value) {
        const bool two = dop_waves_forced() != 1;
        return two ? launch_dop853<Spec, 2, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                               d_result, d_err, d_stats, cancel, dev, stream)
                   : launch_dop853<Spec, 1, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                               d_result, d_err, d_stats, cancel, dev, stream);
      } else if constexpr (std::is_same<V, GhVec2>::value) {
        return launch_dop853<Spec, 2, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                         d_result, d_err, d_stats, cancel, dev, stream);
      } else {
        // variational states: twice the registers and LDS per job, one wave
        // per SIMD (the LDS queue allows four to six workgroups per CU)
        return launch_dop853<Spec, 1, V>(nobj, d_yo, nt, d_t, indiv_t, d_table, rtol, atol,
                                         d_result, d_err, d_stats, cancel, dev, stream);
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

bool types_are(const GhPotHandle* h, std::initializer_list<int> codes) {
  if (h->has_wrap || (int)codes.size() != h->npot) return false;
  int i = 0;
  for (int c : codes)
    if (h->types[i++] != c) return false;
  return true;
}

// Compositions with a specialised kernel: MWPotential2014, the rotating-bar
// model, and single Hernquist / logarithmic halos; anything else runs the
// generic kernel that dispatches on the component type at run time.
// V = GhVec3 integrates full orbits, GhVec2 planar ones.
template <class V>
int launch_integrate(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                     const GhPotHandle* h, double dt, double rtol, double atol, double* d_result,
                     int* d_err, int odeint_type, double* d_stats, hipStream_t stream) {
  if (nobj == 0) return 0;
  if (nobj < 0 || nt < 2) {
    set_error("invalid sizes nobj=%d nt=%d (need nt >= 2)", nobj, nt);
    return -1;
  }
#define GH_LAUNCH(SPEC)                                                                        \
  return launch_spec<SPEC, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result, \
                              d_err, odeint_type, d_stats, stream)
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  using LogBar = GhSpec<GH_LOGHALO, GH_DEHNENBAR>;
  using Hern = GhSpec<GH_HERNQUIST>;
  using Log = GhSpec<GH_LOGHALO>;
  using Any = GhSpec<>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW})) GH_LAUNCH(MW2014);
  if (types_are(h, {GH_LOGHALO, GH_DEHNENBAR})) GH_LAUNCH(LogBar);
  if (types_are(h, {GH_HERNQUIST})) GH_LAUNCH(Hern);
  if (types_are(h, {GH_LOGHALO})) GH_LAUNCH(Log);
  if (h->has_wrap) GH_LAUNCH(GhSpecWrapped);
  GH_LAUNCH(Any);
#undef GH_LAUNCH
}

// integrate_dxdv: V = GhVec6 (full orbits, rows of 12) or GhVec4 (planar,
// rows of 8); MWPotential2014 has a specialised kernel, any other
// composition with device Hessians runs the run-time dispatch.
template <class V>
int launch_integrate_dxdv(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                          const GhPotHandle* h, double dt, double rtol, double atol,
                          double* d_result, int* d_err, int odeint_type, double* d_stats,
                          hipStream_t stream) {
  if (nobj == 0) return 0;
  if (nobj < 0 || nt < 2) {
    set_error("invalid sizes nobj=%d nt=%d (need nt >= 2)", nobj, nt);
    return -1;
  }
  if (!hessian_capable(h)) {
    set_error("integrate_dxdv needs time-independent, axisymmetric components with a Hessian "
              "(every supported type but DehnenBar and a triaxial LogarithmicHalo, unwrapped)");
    return -1;
  }
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    return launch_spec<MW2014, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result,
                                  d_err, odeint_type, d_stats, stream);
  return launch_spec<GhSpec<>, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result,
                                  d_err, odeint_type, d_stats, stream);
}

// SIGINT during the blocking host call sets the polled flag, as galpy's
// handler sets `interrupted` (bovy_symplecticode.c:38-45).
void on_sigint(int) {
  if (g_irq_host) *g_irq_host = 1;
}

// What a host call integrates: galpy's full or planar orbits, or their
// variational equations (integrate_dxdv).
enum GhMode { GH_FULL = 0, GH_PLANAR = 1, GH_DXDV_FULL = 2, GH_DXDV_PLANAR = 3 };

// Blocking host-buffer integration, streamed through the GPU in chunks of
// orbits.  Rows: (R,vR,vT,z,vz,phi), (R,vR,vT,phi), or the rectangular
// (q, p | dq, dp) of 12 / 8 doubles.
int host_integrate(int nobj, double* yo, int nt, double* t, int indiv_t, int npot, int* pot_type,
                   double* pot_args, double dt, double rtol, double atol, double* result,
                   int* err, int odeint_type, void (*cb)(void), int mode) {
  const bool planar = mode == GH_PLANAR || mode == GH_DXDV_PLANAR;
  const size_t N = mode == GH_FULL ? 6 : mode == GH_PLANAR ? 4 : mode == GH_DXDV_FULL ? 12 : 8;
  if (nobj <= 0) return 0;
  if (nt < 2) {
    set_error("nt=%d: need at least two output times", nt);
    return -1;
  }
  if (odeint_type < 0 || odeint_type > 7) {
    set_error("odeint_type %d has no HIP integrator (supported: 0 leapfrog, 1 rk4, 2 rk6, "
              "3 symplec4, 4 symplec6, 5 dopr54, 6 dop853, 7 ias15)", odeint_type);
    return -1;
  }
  GhPotTable table;
  if (decode_potential(npot, pot_type, pot_args, &table, planar) != 0) return -1;
  hipStream_t stream;
  GH_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  void* d_table = nullptr;
  GhPotHandle handle;
  handle.dev = nullptr;
  fill_handle(table, &handle);
  double *d_yo = nullptr, *d_t = nullptr, *d_res = nullptr;
  int* d_err = nullptr;
  int rc = 0;
  struct sigaction old_action, action;
  std::memset(&action, 0, sizeof(action));
  action.sa_handler = on_sigint;
  irq_device_ptr();
  if (g_irq_host) *g_irq_host = 0;
  sigaction(SIGINT, &action, &old_action);
  do {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { rc = -1; set_error("hipMemGetInfo failed"); break; }
    const size_t per_orbit = (size_t)nt * N * 8 + N * 8 + 4 + (indiv_t ? (size_t)nt * 8 : 0);
    size_t budget = free_b / 2;
    long chunk = (long)std::min<size_t>((size_t)nobj, std::max<size_t>(1, budget / per_orbit));
    if (hipMalloc(&d_table, sizeof(GhPotTable)) != hipSuccess ||
        hipMalloc((void**)&d_yo, (size_t)chunk * N * 8) != hipSuccess ||
        hipMalloc((void**)&d_t, (size_t)nt * 8 * (indiv_t ? chunk : 1)) != hipSuccess ||
        hipMalloc((void**)&d_res, (size_t)chunk * nt * N * 8) != hipSuccess ||
        hipMalloc((void**)&d_err, (size_t)chunk * 4) != hipSuccess) {
      rc = -1;
      set_error("device allocation failed (chunk of %ld orbits, nt=%d)", chunk, nt);
      break;
    }
    if (hipMemcpyAsync(d_table, &table, sizeof(GhPotTable), hipMemcpyHostToDevice, stream) != hipSuccess) { rc = -1; break; }
    handle.dev = (GhPotTable*)d_table;
    if (!indiv_t &&
        hipMemcpyAsync(d_t, t, (size_t)nt * 8, hipMemcpyHostToDevice, stream) != hipSuccess) { rc = -1; break; }
    for (long o0 = 0; o0 < nobj && rc == 0; o0 += chunk) {
      const long n = std::min<long>(chunk, nobj - o0);
      if (hipMemcpyAsync(d_yo, yo + N * o0, (size_t)n * N * 8, hipMemcpyHostToDevice, stream) != hipSuccess) { rc = -1; break; }
      if (indiv_t && hipMemcpyAsync(d_t, t + (size_t)nt * o0, (size_t)n * nt * 8,
                                    hipMemcpyHostToDevice, stream) != hipSuccess) { rc = -1; break; }
      int lrc;
      switch (mode) {
        case GH_FULL:
          lrc = launch_integrate<GhVec3>((int)n, d_yo, nt, d_t, indiv_t, &handle, dt, rtol, atol,
                                         d_res, d_err, odeint_type, nullptr, stream);
          break;
        case GH_PLANAR:
          lrc = launch_integrate<GhVec2>((int)n, d_yo, nt, d_t, indiv_t, &handle, dt, rtol, atol,
                                         d_res, d_err, odeint_type, nullptr, stream);
          break;
        case GH_DXDV_FULL:
          lrc = launch_integrate_dxdv<GhVec6>((int)n, d_yo, nt, d_t, indiv_t, &handle, dt, rtol,
                                              atol, d_res, d_err, odeint_type, nullptr, stream);
          break;
        default:
          lrc = launch_integrate_dxdv<GhVec4>((int)n, d_yo, nt, d_t, indiv_t, &handle, dt, rtol,
                                              atol, d_res, d_err, odeint_type, nullptr, stream);
          break;
      }
      if (lrc != 0) { rc = -1; break; }
      if (hipMemcpyAsync(result + N * nt * o0, d_res, (size_t)n * nt * N * 8,
                         hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipMemcpyAsync(err + o0, d_err, (size_t)n * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipMemcpyAsync(yo + N * o0, d_yo, (size_t)n * N * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
          hipStreamSynchronize(stream) != hipSuccess) {
        rc = -1;
        set_error("device copy/sync failed: %s", hipGetErrorString(hipGetLastError()));
        break;
      }
      if (cb)
        for (long i = 0; i < n; i++) cb();
    }
  } while (0);
  sigaction(SIGINT, &old_action, nullptr);
  if (g_irq_host) *g_irq_host = 0;
  (void)hipStreamSynchronize(stream);
  (void)hipFree(d_table);
  (void)hipFree(d_yo);
  (void)hipFree(d_t);
  (void)hipFree(d_res);
  (void)hipFree(d_err);
  (void)hipStreamDestroy(stream);
  return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static void integrate_host_abi(const char* name, int nobj, double* yo, int nt, double* t,
                               int indiv_t, int npot, int* pot_type, double* pot_args,
                               void* pot_tfuncs, double dt, double rtol, double atol,
                               double* result, int* err, int odeint_type, void (*cb)(void),
                               int mode) {
  g_errbuf[0] = 0;
  int rc;
  if (pot_tfuncs != nullptr) {
    set_error("pot_tfuncs must be NULL: no supported potential uses time functions");
    rc = -1;
  } else {
    for (int i = 0; i < nobj; i++) err[i] = 0;
    rc = host_integrate(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args, dt, rtol, atol,
                        result, err, odeint_type, cb, mode);
  }
  if (rc != 0) {
    for (int i = 0; i < nobj; i++) err[i] = -1;
    fprintf(stderr, "galpy_hip: %s failed: %s\n", name, g_errbuf);
  }
}

GH_EXPORT void integrateFullOrbit(int nobj, double* yo, int nt, double* t, int indiv_t, int npot,
                                  int* pot_type, double* pot_args, void* pot_tfuncs, double dt,
                                  double rtol, double atol, double* result, int* err,
                                  int odeint_type, void (*cb)(void)) {
  integrate_host_abi("integrateFullOrbit", nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args,
                     pot_tfuncs, dt, rtol, atol, result, err, odeint_type, cb, GH_FULL);
}

GH_EXPORT void integratePlanarOrbit(int nobj, double* yo, int nt, double* t, int indiv_t,
                                    int npot, int* pot_type, double* pot_args, void* pot_tfuncs,
                                    double dt, double rtol, double atol, double* result, int* err,
                                    int odeint_type, void (*cb)(void)) {
  integrate_host_abi("integratePlanarOrbit", nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args,
                     pot_tfuncs, dt, rtol, atol, result, err, odeint_type, cb, GH_PLANAR);
}

GH_EXPORT void integrateFullOrbit_dxdv(double* yo, int nt, double* t, int npot, int* pot_type,
                                       double* pot_args, void* pot_tfuncs, double dt, double rtol,
                                       double atol, double* result, int* err, int odeint_type) {
  integrate_host_abi("integrateFullOrbit_dxdv", 1, yo, nt, t, 0, npot, pot_type, pot_args,
                     pot_tfuncs, dt, rtol, atol, result, err, odeint_type, nullptr, GH_DXDV_FULL);
}

GH_EXPORT void integratePlanarOrbit_dxdv(double* yo, int nt, double* t, int npot, int* pot_type,
                                         double* pot_args, void* pot_tfuncs, double dt,
                                         double rtol, double atol, double* result, int* err,
                                         int odeint_type, void (*cb)(void)) {
  (void)cb;  // galpy's integratePlanarOrbit_dxdv takes it and does not call it
  integrate_host_abi("integratePlanarOrbit_dxdv", 1, yo, nt, t, 0, npot, pot_type, pot_args,
                     pot_tfuncs, dt, rtol, atol, result, err, odeint_type, nullptr,
                     GH_DXDV_PLANAR);
}

GH_EXPORT void gh_integrate_dxdv(int nobj, int dim, double* yo, int nt, double* t, int indiv_t,
                                 int npot, int* pot_type, double* pot_args, double dt,
                                 double rtol, double atol, double* result, int* err,
                                 int odeint_type) {
  if (dim != 2 && dim != 3) {
    g_errbuf[0] = 0;
    set_error("gh_integrate_dxdv: dim must be 2 (planar) or 3, got %d", dim);
    for (int i = 0; i < nobj; i++) err[i] = -1;
    return;
  }
  integrate_host_abi("gh_integrate_dxdv", nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args,
                     nullptr, dt, rtol, atol, result, err, odeint_type, nullptr,
                     dim == 3 ? GH_DXDV_FULL : GH_DXDV_PLANAR);
}

GH_EXPORT int gh_integrate_dxdv_hip(int nobj, int dim, double* d_yo, int nt, const double* d_t,
                                    int indiv_t, const void* pot, double dt, double rtol,
                                    double atol, double* d_result, int* d_err, int odeint_type,
                                    double* d_stats, void* stream) {
  g_errbuf[0] = 0;
  if (pot == nullptr) {
    set_error("potential handle is NULL");
    return -1;
  }
  if (dim == 3)
    return launch_integrate_dxdv<GhVec6>(nobj, d_yo, nt, d_t, indiv_t, (const GhPotHandle*)pot,
                                         dt, rtol, atol, d_result, d_err, odeint_type, d_stats,
                                         (hipStream_t)stream);
  if (dim == 2)
    return launch_integrate_dxdv<GhVec4>(nobj, d_yo, nt, d_t, indiv_t, (const GhPotHandle*)pot,
                                         dt, rtol, atol, d_result, d_err, odeint_type, d_stats,
                                         (hipStream_t)stream);
  set_error("gh_integrate_dxdv_hip: dim must be 2 (planar) or 3, got %d", dim);
  return -1;
}

GH_EXPORT int integrateFullOrbit_hip(int nobj, double* d_yo, int nt, const double* d_t,
                                     int indiv_t, const void* pot, double dt, double rtol,
                                     double atol, double* d_result, int* d_err, int odeint_type,
                                     double* d_stats, void* stream) {
  g_errbuf[0] = 0;
  if (pot == nullptr) {
    set_error("potential handle is NULL");
    return -1;
  }
  return launch_integrate<GhVec3>(nobj, d_yo, nt, d_t, indiv_t, (const GhPotHandle*)pot, dt, rtol,
                                  atol, d_result, d_err, odeint_type, d_stats, (hipStream_t)stream);
}

GH_EXPORT int integratePlanarOrbit_hip(int nobj, double* d_yo, int nt, const double* d_t,
                                       int indiv_t, const void* pot, double dt, double rtol,
                                       double atol, double* d_result, int* d_err, int odeint_type,
                                       double* d_stats, void* stream) {
  g_errbuf[0] = 0;
  if (pot == nullptr) {
    set_error("potential handle is NULL");
    return -1;
  }
  return launch_integrate<GhVec2>(nobj, d_yo, nt, d_t, indiv_t, (const GhPotHandle*)pot, dt, rtol,
                                  atol, d_result, d_err, odeint_type, d_stats, (hipStream_t)stream);
}

static int upload_potential(int npot, const int* pot_type, const double* pot_args, void** pot,
                            bool planar) {
  g_errbuf[0] = 0;
  GhPotTable table;
  if (decode_potential(npot, pot_type, pot_args, &table, planar) != 0) return -1;
  void* d = nullptr;
  GH_CHECK(hipMalloc(&d, sizeof(GhPotTable)));
  GH_CHECK(hipMemcpy(d, &table, sizeof(GhPotTable), hipMemcpyHostToDevice));
  GhPotHandle* h = new GhPotHandle;
  h->dev = (GhPotTable*)d;
  fill_handle(table, h);
  *pot = h;
  return 0;
}

GH_EXPORT int gh_potential_upload(int npot, const int* pot_type, const double* pot_args,
                                  void** pot) {
  return upload_potential(npot, pot_type, pot_args, pot, false);
}

GH_EXPORT int gh_planar_potential_upload(int npot, const int* pot_type, const double* pot_args,
                                         void** pot) {
  return upload_potential(npot, pot_type, pot_args, pot, true);
}

GH_EXPORT int gh_potential_free(void* pot) {
  if (pot == nullptr) return 0;
  GhPotHandle* h = (GhPotHandle*)pot;
  const hipError_t e = hipFree(h->dev);
  delete h;
  if (e != hipSuccess) {
    set_error("hipFree failed: %s", hipGetErrorString(e));
    return -1;
  }
  return 0;
}

// device address of the table behind a handle (for callers that build their
// own launch descriptors, e.g. batched work lists)
GH_EXPORT const void* gh_potential_device_ptr(const void* pot) {
  return pot ? ((const GhPotHandle*)pot)->dev : nullptr;
}

GH_EXPORT int gh_energy_drift_hip(int nobj, int nt, const double* d_rows, const double* d_t,
                                  int indiv_t, const void* pot, double omega, double* d_e0,
                                  double* d_demax, void* stream) {
  g_errbuf[0] = 0;
  if (nobj <= 0) return 0;
  const dim3 grid((unsigned)((nobj + kBlock - 1) / kBlock)), block(kBlock);
  hipLaunchKernelGGL(gh_energy_kernel<6>, grid, block, 0, (hipStream_t)stream, nobj, nt, d_rows,
                     d_t, indiv_t, ((const GhPotHandle*)pot)->dev, omega, d_e0, d_demax);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_planar_energy_drift_hip(int nobj, int nt, const double* d_rows, const double* d_t,
                                         int indiv_t, const void* pot, double omega, double* d_e0,
                                         double* d_demax, void* stream) {
  g_errbuf[0] = 0;
  if (nobj <= 0) return 0;
  const dim3 grid((unsigned)((nobj + kBlock - 1) / kBlock)), block(kBlock);
  hipLaunchKernelGGL(gh_energy_kernel<4>, grid, block, 0, (hipStream_t)stream, nobj, nt, d_rows,
                     d_t, indiv_t, ((const GhPotHandle*)pot)->dev, omega, d_e0, d_demax);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_orbit_summary_hip(int nobj, int nt, int ncol, const double* d_rows,
                                   const double* d_t, int indiv_t, const void* pot, double omega,
                                   double* d_rperi, double* d_rap, double* d_zmax, double* d_e,
                                   double* d_E, double* d_EJ, void* stream) {
  g_errbuf[0] = 0;
  if (nobj <= 0) return 0;
  if (nt < 1 || (ncol != 6 && ncol != 4)) {
    set_error("gh_orbit_summary_hip: nt=%d ncol=%d (need nt >= 1, ncol 6 or 4)", nt, ncol);
    return -1;
  }
  const bool energies = d_E != nullptr || d_EJ != nullptr;
  if (energies && (pot == nullptr || d_t == nullptr)) {
    set_error("gh_orbit_summary_hip: E and the Jacobi integral need the potential and the times");
    return -1;
  }
  const GhPotTable* T = pot ? ((const GhPotHandle*)pot)->dev : nullptr;
  const dim3 grid((unsigned)((nobj + kSummaryOrbitsPerBlock - 1) / kSummaryOrbitsPerBlock));
  const dim3 block(64 * kSummaryOrbitsPerBlock);
  if (ncol == 6)
    hipLaunchKernelGGL(gh_summary_kernel<6>, grid, block, 0, (hipStream_t)stream, nobj, nt, d_rows,
                       d_t, indiv_t, T, omega, d_rperi, d_rap, d_zmax, d_e, d_E, d_EJ);
  else
    hipLaunchKernelGGL(gh_summary_kernel<4>, grid, block, 0, (hipStream_t)stream, nobj, nt, d_rows,
                       d_t, indiv_t, T, omega, d_rperi, d_rap, d_zmax, d_e, d_E, d_EJ);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_eval_hip(int n, const double* d_R, const double* d_z, const double* d_phi,
                          const double* d_t, const void* pot, double* d_pot, double* d_FR,
                          double* d_Fz, double* d_Fphi, void* stream) {
  g_errbuf[0] = 0;
  if (n <= 0) return 0;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock)), block(kBlock);
  hipLaunchKernelGGL(gh_eval_kernel, grid, block, 0, (hipStream_t)stream, n, d_R, d_z, d_phi, d_t,
                     ((const GhPotHandle*)pot)->dev, d_pot, d_FR, d_Fz, d_Fphi);
  GH_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Differentiation: forward flow and reverse-time adjoint
// ---------------------------------------------------------------------------
namespace {

// N full steps of leapfrog (METHOD 0) or symplec4 (METHOD 3) from the
// rectangular state w0 -- galpy's steppers on t = t0 + k dt with dt given,
// one step per output interval (bovy_symplecticode.c:132-377).
template <int METHOD, class Spec>
__global__ void __launch_bounds__(kBlock)
gh_flow_kernel(int n, const double* __restrict__ w0, double* __restrict__ wT, double t0, double dt,
               long long nsteps, const GhPotTable* __restrict__ T) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const GhForce<Spec> F{T};
  GhVec3 q{w0[6 * i], w0[6 * i + 1], w0[6 * i + 2]}, p{w0[6 * i + 3], w0[6 * i + 4], w0[6 * i + 5]};
  double to = t0;
  for (long long s = 0; s < nsteps; s++) {
    if (METHOD == 0) {
      const GhVec3 q12 = gh_drift(q, p, dt / 2.0);
      p = gh_drift(p, gh_accel(F, to + dt / 2.0, q12), dt);
      q = gh_drift(q12, p, dt / 2.0);
      to = to + dt;
    } else {
      q = gh_drift(q, p, GH_S4C1 * dt);
      to += GH_S4C1 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D1 * dt);
      q = gh_drift(q, p, GH_S4C2 * dt);
      to += GH_S4C2 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D2 * dt);
      q = gh_drift(q, p, GH_S4C3 * dt);
      to += GH_S4C3 * dt;
      p = gh_drift(p, gh_accel(F, to, q), GH_S4D3 * dt);
      q = gh_drift(q, p, GH_S4C4 * dt);
      to += GH_S4C4 * dt;
    }
  }
  wT[6 * i] = q.x; wT[6 * i + 1] = q.y; wT[6 * i + 2] = q.z;
  wT[6 * i + 3] = p.x; wT[6 * i + 4] = p.y; wT[6 * i + 5] = p.z;
}

// Reverse pass: undo the N steps sub-step by sub-step from wT, carrying
// lambda = dL/dw from gT back to t0 and accumulating dL/dtheta per orbit.
template <int METHOD, class Spec>
__global__ void __launch_bounds__(kBlock)
gh_adjoint_kernel(int n, const double* __restrict__ wT, const double* __restrict__ gT, double dt,
                  long long nsteps, const GhPotTable* __restrict__ T, double* __restrict__ lam0,
                  double* __restrict__ w0rec, double* __restrict__ dtheta) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  GhVec3 q{wT[6 * i], wT[6 * i + 1], wT[6 * i + 2]}, p{wT[6 * i + 3], wT[6 * i + 4], wT[6 * i + 5]};
  GhVec3 lq{gT[6 * i], gT[6 * i + 1], gT[6 * i + 2]}, lp{gT[6 * i + 3], gT[6 * i + 4], gT[6 * i + 5]};
  double dth[GH_MAX_POT][GH_ADJ_NPAR] = {};
  auto undo_drift = [&](double h) {
    q = gh_drift(q, p, -h);
    lp = gh_drift(lp, lq, h);
  };
  auto undo_kick = [&](double h) {
    const GhVec3 a = gh_kick_adjoint<Spec>(T, q, h, lp, lq, dth);
    p = gh_drift(p, a, -h);
  };
  for (long long s = 0; s < nsteps; s++) {
    if (METHOD == 0) {
      undo_drift(dt / 2.0);
      undo_kick(dt);
      undo_drift(dt / 2.0);
    } else {
      undo_drift(GH_S4C4 * dt);
      undo_kick(GH_S4D3 * dt);
      undo_drift(GH_S4C3 * dt);
      undo_kick(GH_S4D2 * dt);
      undo_drift(GH_S4C2 * dt);
      undo_kick(GH_S4D1 * dt);
      undo_drift(GH_S4C1 * dt);
    }
  }
  lam0[6 * i] = lq.x; lam0[6 * i + 1] = lq.y; lam0[6 * i + 2] = lq.z;
  lam0[6 * i + 3] = lp.x; lam0[6 * i + 4] = lp.y; lam0[6 * i + 5] = lp.z;
  if (w0rec) {
    w0rec[6 * i] = q.x; w0rec[6 * i + 1] = q.y; w0rec[6 * i + 2] = q.z;
    w0rec[6 * i + 3] = p.x; w0rec[6 * i + 4] = p.y; w0rec[6 * i + 5] = p.z;
  }
  if (dtheta) {
#pragma unroll
    for (int c = 0; c < GH_MAX_POT; c++)
#pragma unroll
      for (int j = 0; j < GH_ADJ_NPAR; j++)
        dtheta[(i * GH_MAX_POT + c) * GH_ADJ_NPAR + j] = dth[c][j];
  }
}


template <class Spec>
void launch_flow(int method, dim3 grid, hipStream_t st, int n, const double* w0, double* wT,
                 double t0, double dt, long long nsteps, const GhPotTable* T) {
  if (method == 0)
    hipLaunchKernelGGL((gh_flow_kernel<0, Spec>), grid, dim3(kBlock), 0, st, n, w0, wT, t0, dt,
                       nsteps, T);
  else
    hipLaunchKernelGGL((gh_flow_kernel<3, Spec>), grid, dim3(kBlock), 0, st, n, w0, wT, t0, dt,
                       nsteps, T);
}

template <class Spec>
void launch_adjoint(int method, dim3 grid, hipStream_t st, int n, const double* wT,
                    const double* gT, double dt, long long nsteps, const GhPotTable* T,
                    double* lam0, double* w0rec, double* dtheta) {
  if (method == 0)
    hipLaunchKernelGGL((gh_adjoint_kernel<0, Spec>), grid, dim3(kBlock), 0, st, n, wT, gT, dt,
                       nsteps, T, lam0, w0rec, dtheta);
  else
    hipLaunchKernelGGL((gh_adjoint_kernel<3, Spec>), grid, dim3(kBlock), 0, st, n, wT, gT, dt,
                       nsteps, T, lam0, w0rec, dtheta);
}

}  // namespace

GH_EXPORT int gh_flow_hip(int n, const double* d_w0, double* d_wT, double t0, double dt,
                          long long nsteps, const void* pot, int method, void* stream) {
  g_errbuf[0] = 0;
  if (n <= 0) return 0;
  if (method != 0 && method != 3) {
    set_error("gh_flow_hip supports leapfrog (0) and symplec4 (3), got %d", method);
    return -1;
  }
  const GhPotHandle* h = (const GhPotHandle*)pot;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  hipStream_t st = (hipStream_t)stream;
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    launch_flow<MW2014>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  else if (h->has_wrap)
    launch_flow<GhSpecWrapped>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  else
    launch_flow<GhSpec<>>(method, grid, st, n, d_w0, d_wT, t0, dt, nsteps, h->dev);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_adjoint_hip(int n, const double* d_wT, const double* d_gT, double dt,
                             long long nsteps, const void* pot, int method, double* d_lam0,
                             double* d_w0rec, double* d_dtheta, void* stream) {
  g_errbuf[0] = 0;
  if (n <= 0) return 0;
  const GhPotHandle* h = (const GhPotHandle*)pot;
  if (method != 0 && method != 3) {
    set_error("gh_adjoint_hip supports leapfrog (0) and symplec4 (3), got %d", method);
    return -1;
  }
  if (!hessian_capable(h)) {
    set_error("the adjoint needs time-independent, axisymmetric components with a Hessian "
              "(every supported type but DehnenBar and a triaxial LogarithmicHalo, unwrapped)");
    return -1;
  }
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  hipStream_t st = (hipStream_t)stream;
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    launch_adjoint<MW2014>(method, grid, st, n, d_wT, d_gT, dt, nsteps, h->dev, d_lam0, d_w0rec,
                           d_dtheta);
  else
    launch_adjoint<GhSpec<>>(method, grid, st, n, d_wT, d_gT, dt, nsteps, h->dev, d_lam0, d_w0rec,
                             d_dtheta);
  GH_CHECK(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Batched integration
// ---------------------------------------------------------------------------
namespace {

// one workgroup per tile: mark the tile's orbits unstarted, clear its flag
__global__ void gh_batch_reset_kernel(const gh_batch_group* __restrict__ groups,
                                      const gh_batch_tile* __restrict__ tiles, int ntiles,
                                      int* __restrict__ tile_done, int* __restrict__ remaining) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const gh_batch_tile tile = tiles[t];
  GhLaneState* state = (GhLaneState*)groups[tile.group].state;
  for (int k = threadIdx.x; k < tile.count; k += blockDim.x) state[tile.first + k].status = 0;
  if (threadIdx.x == 0) {
    tile_done[t] = 0;
    if (t == 0 && remaining) *remaining = ntiles;
  }
}

template <int OPL, bool DOP, bool GEN>
int batch_resident_blocks(int dev) {
  static int cached[64] = {0};
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  int per_cu = 0, ncu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gh_batch_kernel<OPL, DOP, GEN>, 64,
                                                   0) !=
          hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  const int n = std::max(1, per_cu) * std::max(1, ncu);
  if (dev >= 0 && dev < 64) cached[dev] = n;
  return n;
}

}  // namespace

GH_EXPORT size_t gh_batch_state_bytes(void) { return sizeof(GhLaneState); }

GH_EXPORT int gh_batch_reset(const gh_batch_group* d_groups, const gh_batch_tile* d_tiles,
                             int ntiles, int* d_tile_done, int* d_remaining, void* stream) {
  g_errbuf[0] = 0;
  irq_device_ptr();  // set up outside any graph capture of the launches
  if (ntiles <= 0) return 0;
  hipLaunchKernelGGL(gh_batch_reset_kernel, dim3((unsigned)ntiles), dim3(128), 0,
                     (hipStream_t)stream, d_groups, d_tiles, ntiles, d_tile_done, d_remaining);
  GH_CHECK(hipGetLastError());
  return 0;
}

GH_EXPORT int gh_batch_launch(const gh_batch_group* d_groups, const gh_batch_tile* d_tiles,
                              int ntiles, int* d_tile_done, unsigned long long* d_counter,
                              int* d_remaining, long long budget, int opl, int flags,
                              void* stream) {
  g_errbuf[0] = 0;
  if (ntiles <= 0) return 0;
  if (opl != 1 && opl != 2) {
    set_error("orbits per lane must be 1 or 2 (got %d)", opl);
    return -1;
  }
  int dev = 0;
  GH_CHECK(hipGetDevice(&dev));
  const bool dop = (flags & GH_BATCH_DOP) != 0;
  const bool gen = (flags & GH_BATCH_SPECIALISED) == 0;
  const int key = (opl - 1) * 4 + (dop ? 2 : 0) + (gen ? 1 : 0);
  int resident = 0;
  switch (key) {
    case 0: resident = batch_resident_blocks<1, false, false>(dev); break;
    case 1: resident = batch_resident_blocks<1, false, true>(dev); break;
    case 2: resident = batch_resident_blocks<1, true, false>(dev); break;
    case 3: resident = batch_resident_blocks<1, true, true>(dev); break;
    case 4: resident = batch_resident_blocks<2, false, false>(dev); break;
    case 5: resident = batch_resident_blocks<2, false, true>(dev); break;
    case 6: resident = batch_resident_blocks<2, true, false>(dev); break;
    default: resident = batch_resident_blocks<2, true, true>(dev); break;
  }
  if (resident <= 0) {
    set_error("occupancy query failed for the batch kernel");
    return -1;
  }
  const unsigned nblocks = (unsigned)std::min(ntiles, resident);
  const long long b = budget > 0 ? budget : (1ll << 62);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(gh_set_counter, dim3(1), dim3(1), 0, st, d_counter, 0ull);
  const int* cancel = irq_device_ptr();
#define GH_BATCH(O, D, G)                                                                    \
  hipLaunchKernelGGL((gh_batch_kernel<O, D, G>), dim3(nblocks), dim3(64), 0, st, d_groups,  \
                     d_tiles, ntiles, d_tile_done, d_counter, d_remaining, b, cancel)
  switch (key) {
    case 0: GH_BATCH(1, false, false); break;
    case 1: GH_BATCH(1, false, true); break;
    case 2: GH_BATCH(1, true, false); break;
    case 3: GH_BATCH(1, true, true); break;
    case 4: GH_BATCH(2, false, false); break;
    case 5: GH_BATCH(2, false, true); break;
    case 6: GH_BATCH(2, true, false); break;
    default: GH_BATCH(2, true, true); break;
  }
#undef GH_BATCH
  GH_CHECK(hipGetLastError());
  return 0;
}

// Host evaluation of the force-kernel incomplete gamma gamma(a, x) of
// component `comp` (a PowerSphericalPotentialwCutoff) exactly as the device
// computes it from the decoded table: table on [a+1, xbig), series below,
// Gamma(a) above.  For checks of the table fit without a GPU.
GH_EXPORT int gh_host_bulge_lower_gamma(int npot, const int* pot_type, const double* pot_args,
                                        int comp, int n, const double* x, double* out) {
  g_errbuf[0] = 0;
  static GhPotTable T;
  if (decode_potential(npot, pot_type, pot_args, &T) != 0) return -1;
  if (comp < 0 || comp >= T.npot || T.comp[comp].type != GH_POWERSPHCUTOFF) {
    set_error("component %d is not a PowerSphericalPotentialwCutoff", comp);
    return -1;
  }
  const GhComponent& C = T.comp[comp];
  const double a = C.c[0];
  const double* tab = &T.tab[(int)C.c[11]][0][0];
  for (int i = 0; i < n; i++) {
    const double xi = x[i];
    double g;
    if (xi >= C.c[4]) {
      g = C.c[1];
    } else if (xi <= 0.0) {
      g = 0.0;
    } else if (xi < a + 1.0) {
      const double pref = std::exp(a * std::log(xi) - xi);
      double sacc = T.ser[comp][0][GH_IG_NSER - 1];
      for (int q = GH_IG_NSER - 2; q >= 0; q--) sacc = std::fma(sacc, xi, T.ser[comp][0][q]);
      g = pref * sacc;
    } else {
      int e;
      const double m = std::frexp(xi, &e);
      const double u = (m - 0.5) * (2.0 * GH_TAB_PER_OCT);
      int k = (int)u;
      k = k > GH_TAB_PER_OCT - 1 ? GH_TAB_PER_OCT - 1 : k;
      const int seg = (e - (int)C.c[12]) * GH_TAB_PER_OCT + k;
      const double t = 2.0 * (u - k) - 1.0;
      const double* c = tab + seg * GH_TAB_NC;
      double v = c[GH_TAB_NC - 1];
      for (int q = GH_TAB_NC - 2; q >= 0; q--) v = std::fma(v, t, c[q]);
      g = C.c[1] - v;
    }
    out[i] = g;
  }
  return 0;
}

GH_EXPORT int gh_host_log_ge1(int n, const double* y, double* out) {
  for (int i = 0; i < n; i++) out[i] = gh_log_ge1(y[i]);
  return 0;
}

GH_EXPORT const char* galpy_hip_last_error(void) { return g_errbuf; }

GH_EXPORT void galpy_hip_set_interrupt(int value) {
  irq_device_ptr();
  if (g_irq_host) *g_irq_host = value;
}
