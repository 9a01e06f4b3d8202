// MI355X orbit-integration engine: host unit.  Decodes galpy's packed
// potential into the device table, chooses the kernel instance for a
// composition, drives host-buffer calls through the GPU, and exports the
// C-ABI.  The C-ABI mirrors galpy's integrateFullOrbit
// (galpy/orbit/orbit_c_ext/integrateFullOrbit.c:1046-1150): cylindrical
// initial conditions in, cylindrical rows out, one call for all orbits.
// Declarations and argument meaning: include/galpy_hip.h.
#include <hip/hip_runtime.h>
#include <signal.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "gh_internal.h"
#include "dxdv_dev.h"

namespace gh {

thread_local char g_errbuf[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// Interrupt flag: fine-grained pinned host memory that running kernels poll.
// ---------------------------------------------------------------------------
volatile int* g_irq_host = nullptr;
static int* g_irq_dev = nullptr;

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

// Work counters for the persistent kernels: one small device array per
// device, allocated on first use and kept for the life of the process; each
// launch takes the next slot, so concurrent launches on different streams
// never share a counter.
constexpr int kCounterSlots = 4096;
static unsigned long long* g_counters[64] = {nullptr};
static unsigned g_counter_next = 0;

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

namespace {

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
    case GH_EINASTO: case GH_EXPTRUNCNFW: return 3;  // amp, h, n / amp, a, rc
    case GH_TWOPOWERSPH: return 4;                    // amp, a, alpha, beta
    case GH_FLATPOWER: return planar ? 3 : 4;  // planar: amp, alpha, core^2
    case GH_NULLPOT: return 0;
    case GH_KKSTAECKEL: return 3;  // amp, ac, Delta
    case GH_ELLIPDISK: return planar ? 6 : -1;  // planar only
    case GH_LOPSIDED: return planar ? 4 : -1;   // planar only
    case GH_STEADYSPIRAL: return planar ? 8 : -1;  // planar only
    case GH_TRANSIENTSPIRAL: return planar ? 8 : -1;  // planar only
    case GH_COSMPHIDISK: return planar ? 9 : -1;      // planar only
    case GH_HENONHEILES: return planar ? 1 : -1;      // planar only
    case GH_SOFTNEEDLEBAR: return 13;  // amp, a, b, c^2, pa, omegab + 7 cache slots
    // wrappers' own arguments, read after the wrapped components
    // (integrateFullOrbit.c:821-895, integratePlanarOrbit.c:591-663)
    case GH_DEHNENSMOOTHWRAP: return 4;   // amp, tform, tsteady, grow
    case GH_SOLIDBODYROTWRAP: return 3;   // amp, omega, pa
    case GH_COROTATINGROTWRAP: return 5;  // amp, vpo, beta, pa, to
    case GH_GAUSSIANAMPWRAP: return 3;    // amp, to, sigma^2
    // amp, 7 cache slots, rot[9], rotSet, offsetSet, offset[3], 10 cache
    // slots (integrateFullOrbit.py:546-568; no planar version in galpy)
    case GH_ROTATETILTWRAP: return planar ? -1 : 32;
    case GH_KUZMINLIKEWRAP: return 3;  // amp, a, b^2
    case GH_CYLSEPWRAP: return 3;      // amp, Rp, Phi(Rp, 0)
    case GH_OBLATESTAECKELWRAP: return 5;  // amp, delta, u0, v0, refpot
    case GH_TIMEDEPAMPWRAP: return 1;  // amp (and one time function)
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

// GSL's natural cubic spline through (xs, ys), n >= 2 (gsl_interp_cspline,
// c_i = y''/2, c_0 = c_{n-1} = 0; Thomas' algorithm in long double) into
// T->ext in gh_spline_eval's layout; off receives its offset.
int build_spline(const double* xs, const double* ys, int n, GhPotTable* T, int& off) {
  if (n < 2 || T->next + 1 + n + 5 * (n - 1) > GH_EXT_LEN) {
    set_error("spline of %d knots does not fit the potential table", n);
    return -1;
  }
  for (int i = 0; i + 1 < n; i++)
    if (!(xs[i + 1] > xs[i])) {
      set_error("spline knots must increase");
      return -1;
    }
  std::vector<long double> c(n, 0.0L);
  if (n >= 3) {
    const int m = n - 2;
    std::vector<long double> diag(m), offd(m), rhs(m);
    for (int i = 1; i <= m; i++) {
      const long double h0 = (long double)xs[i] - xs[i - 1], h1 = (long double)xs[i + 1] - xs[i];
      diag[i - 1] = 2.0L * (h0 + h1);
      offd[i - 1] = h1;
      rhs[i - 1] = 3.0L * (((long double)ys[i + 1] - ys[i]) / h1 - ((long double)ys[i] - ys[i - 1]) / h0);
    }
    for (int i = 1; i < m; i++) {
      const long double w = offd[i - 1] / diag[i - 1];
      diag[i] -= w * offd[i - 1];
      rhs[i] -= w * rhs[i - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (rhs[i] - offd[i] * c[i + 2]) / diag[i];
  }
  double* e = T->ext + T->next;
  e[0] = n;
  for (int i = 0; i < n; i++) e[1 + i] = xs[i];
  long double integ = 0.0L;
  for (int i = 0; i + 1 < n; i++) {
    const long double h = (long double)xs[i + 1] - xs[i];
    const long double b = ((long double)ys[i + 1] - ys[i]) / h - h * (c[i + 1] + 2.0L * c[i]) / 3.0L;
    const long double d = (c[i + 1] - c[i]) / (3.0L * h);
    double* q = e + 1 + n + 5 * i;
    q[0] = ys[i];
    q[1] = (double)b;
    q[2] = (double)c[i];
    q[3] = (double)d;
    q[4] = (double)integ;
    integ += h * (ys[i] + h * (b / 2.0L + h * (c[i] / 3.0L + h * d / 4.0L)));
  }
  off = T->next;
  T->next += 1 + n + 5 * (n - 1);
  return 0;
}

bool is_ellipsoidal(int code) {
  return code == GH_TRIAXHERNQUIST || code == GH_TRIAXNFW || code == GH_TRIAXJAFFE ||
         code == GH_TWOPOWERTRIAX || code == GH_PERFECTELLIPSOID || code == GH_TRIAXGAUSSIAN ||
         code == GH_POWERTRIAX;
}

// An EllipsoidalPotential (integrateFullOrbit.py:221-300): amp, 15 cache
// slots, npsi, the psi arguments, b^2, c^2, aligned, rot[9], glorder, the
// nodes s_k and galpy's weights.  Keeps amp and the psi arguments in p, the
// rotation and per node (1/(1+tau), 1/(b^2+tau), 1/(c^2+tau), w) in ext,
// tau = 1/s^2 - 1 as EllipsoidalPotential.c:64 forms it.
int decode_ellipsoidal(int code, const double*& pot_args, GhPotTable* T, int i) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int npsi = (int)a[16];
  if (npsi < 1 || npsi > 7) {
    set_error("pot_type %d: %d psi arguments", code, npsi);
    return -1;
  }
  const double* e = a + 17 + npsi;
  const double b2 = e[0], c2 = e[1];
  const bool aligned = e[2] != 0.0;
  const int glorder = (int)e[12];
  if (glorder < 1 || T->next + 9 + 4 * glorder > GH_EXT_LEN) {
    set_error("pot_type %d: glorder %d does not fit the table (%d doubles left)", code, glorder,
              GH_EXT_LEN - T->next - 9);
    return -1;
  }
  C.type = code;
  C.p[0] = a[0];
  for (int j = 0; j < npsi; j++) C.p[1 + j] = a[17 + j];
  C.c[0] = b2;
  C.c[1] = c2;
  C.c[2] = aligned ? 1.0 : 0.0;
  C.c[3] = glorder;
  if (code == GH_TWOPOWERTRIAX) C.c[4] = 1.0 / C.p[1];
  C.ext = T->next;
  double* x = T->ext + T->next;
  for (int j = 0; j < 9; j++) x[j] = aligned ? (j % 4 == 0 ? 1.0 : 0.0) : e[3 + j];
  const double* glx = e + 13;
  const double* glw = glx + glorder;
  for (int k = 0; k < glorder; k++) {
    const double tau = 1.0 / glx[k] / glx[k] - 1.0;
    x[9 + 4 * k] = 1.0 / (1.0 + tau);
    x[10 + 4 * k] = 1.0 / (b2 + tau);
    x[11 + 4 * k] = 1.0 / (c2 + tau);
    x[12 + 4 * k] = glw[k];
  }
  T->next += 9 + 4 * glorder;
  T->has_wrap = 1;  // runs in the extended dispatch instances
  C.nonaxi = !aligned || std::fabs(std::sqrt(b2) - 1.0) > 1e-10;
  T->nonaxi |= C.nonaxi;
  pot_args += 30 + npsi + 2 * glorder;
  return 0;
}

// e^x E1(x) on the host in long double (A&S 5.1.11 / 5.1.22, as on the device)
double host_e1_scaled(double xd) {
  const long double x = xd;
  if (x <= 1.0L) {
    long double sum = 0.0L, term = 1.0L;
    for (int k = 1; k < 80; k++) {
      term *= -x / k;
      sum += term / k;
    }
    return (double)(expl(x) * (-0.5772156649015328606065120900824024L - logl(x) - sum));
  }
  long double b = x + 1.0L, c = 1e300L, d = 1.0L / b, h = d;
  for (int i = 1; i < 5000; i++) {
    const long double an = -(long double)i * i;
    b += 2.0L;
    d = 1.0L / (an * d + b);
    c = b + an / c;
    const long double del = c * d;
    h *= del;
    if (fabsl(del - 1.0L) < 1e-19L) break;
  }
  return (double)h;
}

// TwoPowerSphericalPotential: the enclosed-mass tables of gh_twopower_mass
// (series coefficients s_n = (B)_n/(A+1)_n for x <= 3; the tail expansion's
// d_k = c_k / (3 - beta - k), c_k = (-1)^k (B)_k / k!, the constant K and
// the logarithm's coefficient for x > 3), in long double.
int decode_twopower(GhPotTable* T, GhComponent& C) {
  const double a = C.p[1], alpha = C.p[2], beta = C.p[3];
  if (!(a > 0.0) || !(alpha < 3.0) || !(beta - alpha > 0.0)) {
    set_error("TwoPowerSphericalPotential a=%g alpha=%g beta=%g: need a > 0, alpha < 3, beta > alpha",
              a, alpha, beta);
    return -1;
  }
  const long double A = 3.0L - alpha, B = (long double)beta - alpha;
  const double amb = 3.0 - beta;
  double* e = T->ext + T->next;
  const int room = GH_EXT_LEN - T->next;
  // series at w <= 3/4: until the tail is below 1e-19 of the sum
  const long double w0 = 0.75L;
  long double sn = 1.0L, S = 0.0L, wn = 1.0L;
  int ns = 0;
  bool done = false;
  for (int n = 0; n < room && n < 3000 && !done; n++) {
    e[n] = (double)sn;
    S += sn * wn;
    ns = n + 1;
    // the terms shrink from here on (ratio < 1) and are negligible
    done = n > 8 && sn * wn < 1e-19L * S && (B + n) / (A + 1.0L + n) * w0 < 1.0L;
    sn *= (B + n) / (A + 1.0L + n);
    wn *= w0;
  }
  if (!done) {
    set_error("TwoPowerSphericalPotential alpha=%g beta=%g: series did not converge", alpha, beta);
    return -1;
  }
  const long double I3 = powl(3.0L, A) / A * powl(4.0L, -B) * S;
  // tail: k* = 3 - beta when a nonnegative integer
  const int kstar = (amb >= 0.0 && amb == std::floor(amb)) ? (int)amb : -1;
  double* d = e + ns;
  long double ck = 1.0L, K = I3, cstar = 0.0L;
  int nd = 0;
  done = false;
  for (int k = 0; k < room - ns && k < 2000 && !done; k++) {
    if (k > 0) ck *= -(B + k - 1.0L) / k;
    nd = k + 1;
    if (k == kstar) {
      cstar = ck;
      d[k] = 0.0;
      K -= ck * logl(3.0L);
      continue;
    }
    const long double dk = ck / ((long double)amb - k);
    d[k] = (double)dk;
    const long double at3 = dk * powl(3.0L, (long double)amb - k);
    K -= at3;
    done = k > kstar + 2 && k > 4 && fabsl(at3) < 1e-19L * fabsl(I3);
  }
  if (!done) {
    set_error("TwoPowerSphericalPotential alpha=%g beta=%g: series tables do not fit", alpha, beta);
    return -1;
  }
  C.ext = T->next;
  T->next += ns + nd;
  C.c[0] = (double)A;
  C.c[1] = (double)B;
  C.c[2] = amb;
  C.c[3] = ns;
  C.c[4] = nd;
  C.c[5] = (double)K;
  C.c[6] = (double)cstar;
  C.c[7] = 1.0 / a;
  C.c[8] = (double)(1.0L / A);
  return 0;
}

// SpiralArmsPotential (integrateFullOrbit.py:337-352, integratePlanarOrbit.py:
// 391-409, the same in both packings): nCs, amp, N, sin alpha, tan alpha,
// r_ref, phi_ref, Rs, H, omega, then C_1..C_nCs (kept in ext).
int decode_spiralarms(const double*& pot_args, GhPotTable* T, int i) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int nCs = (int)a[0];
  if (nCs < 1 || T->next + nCs > GH_EXT_LEN) {
    set_error("SpiralArmsPotential with %d Cs does not fit the table", nCs);
    return -1;
  }
  if (!(a[7] > 0.0) || !(a[8] > 0.0) || a[3] == 0.0 || a[4] == 0.0 || a[2] == 0.0) {
    set_error("SpiralArmsPotential needs Rs > 0, H > 0, N != 0 and a nonzero pitch angle");
    return -1;
  }
  C.type = GH_SPIRALARMS;
  for (int j = 0; j < 8; j++) C.p[j] = a[1 + j];  // amp, N, sin a, tan a, r_ref, phi_ref, Rs, H
  C.c[0] = a[9];
  C.c[1] = nCs;
  C.c[2] = 1.0 / a[7];
  C.c[3] = 1.0 / a[4];
  C.c[4] = std::log(a[5]);
  C.c[5] = a[8] * a[2] / a[3];
  C.ext = T->next;
  for (int k = 0; k < nCs; k++) T->ext[T->next + k] = a[10 + k];
  T->next += nCs;
  T->has_wrap = 1;  // extended dispatch instances
  C.nonaxi = 1;
  T->nonaxi = 1;
  pot_args += 10 + nCs;
  return 0;
}

// DoubleExponentialDiskPotential (integrateFullOrbit.py:93-107, the same in
// the planar packing): amp, -4 pi alpha amp, alpha, beta, de_n, then the J0
// nodes, J1 nodes, J0 weights and J1 weights, stored in ext as J0 (node,
// weight) pairs followed by J1 pairs.
int decode_dedisk(const double*& pot_args, GhPotTable* T, int i, bool planar) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int n = (int)a[4];
  if (n < 1 || T->next + 4 * n > GH_EXT_LEN) {
    set_error("DoubleExponentialDiskPotential with de_n=%d does not fit the table (%d doubles left)",
              n, GH_EXT_LEN - T->next);
    return -1;
  }
  C.type = GH_DEDISK;
  for (int j = 0; j < 4; j++) C.p[j] = a[j];
  C.c[0] = n;
  C.c[1] = a[2] * a[2];
  C.c[2] = a[3] * a[3];
  C.c[3] = planar ? 1.0 : 0.0;
  C.ext = T->next;
  double* e = T->ext + T->next;
  const double *j0x = a + 5, *j1x = a + 5 + n, *j0w = a + 5 + 2 * n, *j1w = a + 5 + 3 * n;
  for (int k = 0; k < n; k++) {
    e[2 * k] = j0x[k];
    e[2 * k + 1] = j0w[k];
    e[2 * n + 2 * k] = j1x[k];
    e[2 * n + 2 * k + 1] = j1w[k];
  }
  T->next += 4 * n;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args += 5 + 4 * n;
  return 0;
}

// interpSphericalPotential (integrateFullOrbit.py:366-373, the same in the
// planar packing): nr, rgrid, the force on rgrid, amp, rmin, rmax,
// M(<rmax), Phi0, Phimax; the force's natural spline in ext.
int decode_interpsph(const double*& pot_args, GhPotTable* T, int i) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int nr = (int)a[0];
  C.type = GH_INTERPSPH;
  int off = 0;
  if (build_spline(a + 1, a + 1 + nr, nr, T, off) != 0) return -1;
  C.ext = off;
  for (int j = 0; j < 6; j++) C.p[j] = a[1 + 2 * nr + j];
  T->has_wrap = 1;  // extended dispatch instances
  pot_args += 1 + 2 * nr + 6;
  return 0;
}

// SCFPotential (integratePlanarOrbit.py:726-762, the same in both
// packings): a, isNonAxi, N, L, M, Nt, the coefficient blocks (static
// Acos[, Asin] or tgrid and the time-cubic blocks), 11 cache slots; the
// blocks are copied to ext as they are.
int decode_scf(const double*& pot_args, GhPotTable* T, int i) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int nonaxi = (int)a[1], N = (int)a[2], L = (int)a[3], M = (int)a[4], Nt = (int)a[5];
  const long NLM = (long)N * L * M;
  const long nblk = Nt == 0 ? (1 + nonaxi) * NLM : Nt + (1 + nonaxi) * NLM * 4 * (Nt - 1);
  if (N < 1 || L < 1 || M < 1 || Nt == 1 || Nt < 0 || (nonaxi && M != L)) {
    set_error("SCFPotential N=%d L=%d M=%d Nt=%d isNonAxi=%d: unsupported shape", N, L, M, Nt, nonaxi);
    return -1;
  }
  if (T->next + nblk > GH_EXT_LEN) {
    set_error("SCFPotential coefficients (%ld doubles) do not fit the potential table", nblk);
    return -1;
  }
  C.type = GH_SCF;
  for (int j = 0; j < 6; j++) C.p[j] = a[j];
  C.ext = T->next;
  for (long k = 0; k < nblk; k++) T->ext[T->next + k] = a[6 + k];
  T->next += (int)nblk;
  C.nonaxi = nonaxi;
  T->nonaxi |= nonaxi;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args += 6 + nblk + 11;
  return 0;
}

// One [Sigma_i, h_i] pair of a DiskSCFPotential (integratePlanarOrbit.py:
// 778-812, the same in both packings): nsigma_args (3 or 4), Sigma type,
// amp, h[, Rhole], hz type, hz h.
int decode_diskscf(const double*& pot_args, GhPotTable* T, int i, bool planar) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int ns = (int)a[0];
  if ((ns != 3 && ns != 4) || ((int)a[1] == 1) != (ns == 4)) {
    set_error("DiskSCFPotential pair with %d Sigma arguments of type %g", ns, a[1]);
    return -1;
  }
  C.type = GH_DISKSCF;
  C.p[0] = a[1];
  C.p[1] = a[2];
  C.p[2] = a[3];
  C.p[3] = ns == 4 ? a[4] : 0.0;
  C.p[4] = a[1 + ns];
  C.p[5] = a[2 + ns];
  C.c[0] = 1.0 / C.p[2];
  C.c[1] = 1.0 / C.p[5];
  C.c[2] = planar ? 1.0 : 0.0;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args += ns + 3;
  return 0;
}

// interpRZPotential (integrateFullOrbit.py:111-173; full orbits only, as in
// galpy): nR, nz, three second-derivative flags, the grids, the B-spline
// coefficients of the potential, R force, z force (and of the flagged
// second derivatives, not needed here), amp, logR.
int decode_interprz(const double*& pot_args, GhPotTable* T, int i, bool planar) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int nR = (int)a[0], nz = (int)a[1];
  const int nextra = (int)a[2] + (int)a[3] + (int)a[4];
  if (planar) {
    set_error("interpRZPotential has no planar version");
    return -1;
  }
  const long n = (long)nR + nz + 3L * nR * nz;
  if (nR < 2 || nz < 2 || T->next + n > GH_EXT_LEN) {
    set_error("interpRZPotential grid %d x %d does not fit the potential table", nR, nz);
    return -1;
  }
  C.type = GH_INTERPRZ;
  const double* tail = a + 5 + nR + nz + (long)(3 + nextra) * nR * nz;
  C.p[0] = tail[0];
  C.p[1] = tail[1];
  C.c[0] = nR;
  C.c[1] = nz;
  C.ext = T->next;
  for (long k = 0; k < n; k++) T->ext[T->next + k] = a[5 + k];
  T->next += (int)n;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args = tail + 2;
  return 0;
}

// MultipoleExpansionPotential (MultipoleExpansionPotential.c:183-399; packing
// MultipoleExpansionPotential.py:2101-2216, the same in both packings): Nr,
// L, M, isNonAxi, rgrid[Nr], amp, Nt, then per (l, m < min(l + 1, M)) and
// trig (cos; sin when m > 0) either the quintic coefficients of I_inner and
// I_outer [Nr-1][6] and the density at the nodes [Nr] (static), or, after
// tgrid[Nt], their cubic-in-t blocks [Nt-1][6 (Nr-1)][4] and the density's
// [Nt-1][Nr][4].  Re-laid out interval-major for the device (one radial
// interval's coefficients of every block contiguous); the density is not
// needed for forces.
int decode_multipole(const double*& pot_args, GhPotTable* T, int i) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  const int Nr = (int)a[0], L = (int)a[1], M = (int)a[2], nonaxi = (int)a[3];
  if (Nr < 2 || L < 1 || M < 1 || (!nonaxi && M != 1)) {
    set_error("MultipoleExpansionPotential Nr=%d L=%d M=%d isNonAxi=%d: unsupported shape", Nr, L, M,
              nonaxi);
    return -1;
  }
  const double* rg = a + 4;
  const double amp = rg[Nr];
  const int Nt = (int)rg[Nr + 1];
  if (Nt == 1 || Nt < 0) {
    set_error("MultipoleExpansionPotential with Nt=%d time nodes", Nt);
    return -1;
  }
  int nblk = 0;
  for (int l = 0; l < L; l++) nblk += 2 * std::min(l + 1, M) - 1;
  const long nr = 6L * (Nr - 1);
  const long stride = Nt > 0 ? 48 : 12;
  const long nint = Nt > 0 ? (long)(Nt - 1) * (Nr - 1) : (long)(Nr - 1);
  const long need = Nr + (Nt > 0 ? Nt : 0) + nint * nblk * stride;
  if (T->next + need > GH_EXT_LEN) {
    set_error("MultipoleExpansionPotential coefficients (%ld doubles) do not fit the potential table", need);
    return -1;
  }
  C.type = GH_MULTIPOLE;
  C.p[0] = amp;
  C.c[0] = Nr;
  C.c[1] = L;
  C.c[2] = M;
  C.c[3] = Nt;
  C.c[4] = rg[0];
  C.c[5] = rg[Nr - 1];
  C.c[6] = nblk;
  C.ext = T->next;
  double* e = T->ext + T->next;
  for (int k = 0; k < Nr; k++) e[k] = rg[k];
  const double* src = rg + Nr + 2;
  double* dst = e + Nr;
  if (Nt > 0) {
    for (int k = 0; k < Nt; k++) dst[k] = src[k];
    src += Nt;
    dst += Nt;
    const long per = (long)(Nt - 1) * (2 * nr + Nr) * 4;  // one block's doubles in the packing
    for (int k = 0; k < nblk; k++, src += per)
      for (int it = 0; it < Nt - 1; it++)
        for (int ir = 0; ir < Nr - 1; ir++) {
          double* d = dst + (((long)it * (Nr - 1) + ir) * nblk + k) * 48;
          for (int half = 0; half < 2; half++) {  // I_inner, I_outer
            const double* s0 = src + half * (long)(Nt - 1) * nr * 4 + ((long)it * nr + 6L * ir) * 4;
            for (int j = 0; j < 24; j++) d[half * 24 + j] = s0[j];
          }
        }
  } else {
    const long per = 2 * nr + Nr;
    for (int k = 0; k < nblk; k++, src += per)
      for (int ir = 0; ir < Nr - 1; ir++) {
        double* d = dst + ((long)ir * nblk + k) * 12;
        for (int j = 0; j < 6; j++) {
          d[j] = src[6L * ir + j];
          d[6 + j] = src[nr + 6L * ir + j];
        }
      }
  }
  T->next += (int)need;
  C.nonaxi = nonaxi;
  T->nonaxi |= nonaxi;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args = src;
  return 0;
}

// One (leaf) component's arguments and precomputed constants.
int decode_component(int code, const double*& pot_args, GhPotTable* T, int i, int& nslots,
                     bool planar) {
  if (is_ellipsoidal(code)) return decode_ellipsoidal(code, pot_args, T, i);
  if (code == GH_SPIRALARMS) return decode_spiralarms(pot_args, T, i);
  if (code == GH_DEDISK) return decode_dedisk(pot_args, T, i, planar);
  if (code == GH_INTERPSPH) return decode_interpsph(pot_args, T, i);
  if (code == GH_SCF) return decode_scf(pot_args, T, i);
  if (code == GH_DISKSCF) return decode_diskscf(pot_args, T, i, planar);
  if (code == GH_INTERPRZ) return decode_interprz(pot_args, T, i, planar);
  if (code == GH_MULTIPOLE) return decode_multipole(pot_args, T, i);
  const int na = nargs_of(code, planar);
  if (na < 0 || code < 0) {
    set_error("pot_type %d (component %d) has no HIP kernel", code, i);
    return -1;
  }
  {
    GhComponent& C = T->comp[i];
    C.type = code;
    // the first 8 arguments in p; the few components with more keep the
    // rest in c (decoded below)
    double extra[8];
    for (int j = 0; j < na; j++) {
      if (j < 8) C.p[j] = *pot_args++;
      else extra[j - 8] = *pot_args++;
    }
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
      case GH_TRANSIENTSPIRAL: C.nonaxi = 1; break;
      case GH_HENONHEILES: C.nonaxi = 1; break;
      case GH_COSMPHIDISK:
        C.p[4] = (double)(int)C.p[4];  // m, read as an int (CosmphiDiskPotential.c:13)
        C.c[0] = extra[0];             // r1p
        C.c[2] = C.p[1] / C.p[4];      // mphio / m
        C.nonaxi = 1;
        break;
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
      case GH_EINASTO: {
        const double h = C.p[1], n = C.p[2];
        if (!(h > 0.0) || !(n > 0.0)) {
          set_error("EinastoPotential h=%g n=%g must be positive", h, n);
          return -1;
        }
        C.c[0] = 3.0 * n;
        C.c[1] = std::tgamma(3.0 * n);
        C.c[2] = std::ldexp(C.c[1], -58);
        C.c[3] = 2.0 * n;
        C.c[4] = std::tgamma(2.0 * n);
        C.c[5] = std::ldexp(C.c[4], -58);
        C.c[6] = 4.0 * M_PI * h * h * h * n;
        C.c[7] = 1.0 / h;
        C.c[8] = 1.0 / n;
        T->has_wrap = 1;  // extended dispatch
        break;
      }
      case GH_EXPTRUNCNFW: {  // ExpTruncNFWPotential.c:11-28
        const double a = C.p[1], rc = C.p[2];
        if (!(a > 0.0) || !(rc > 0.0)) {
          set_error("ExpTruncNFWPotential a=%g rc=%g must be positive", a, rc);
          return -1;
        }
        C.c[0] = a / rc;
        C.c[1] = host_e1_scaled(a / rc);
        C.c[2] = 1.0 / rc;
        C.c[3] = 1e-3 * (a < rc ? a : rc);
        C.c[4] = 0.5;
        C.c[5] = -(1.0 / rc + 2.0 / a) / 3.0;
        C.c[6] = (0.5 / (rc * rc) + 2.0 / (rc * a) + 3.0 / (a * a)) / 4.0;
        C.c[7] = -(1.0 / (6.0 * rc * rc * rc) + 1.0 / (rc * rc * a) + 3.0 / (rc * a * a) +
                   4.0 / (a * a * a)) / 5.0;
        C.c[8] = 1.0 / (a * a);
        T->has_wrap = 1;
        break;
      }
      case GH_TWOPOWERSPH:
        if (decode_twopower(T, C) != 0) return -1;
        T->has_wrap = 1;
        break;
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
// The caller's time functions (galpy's pot_tfuncs: C function pointers,
// ctypes callbacks into Python for galpy) and the time range the tables
// must cover; consumed in galpy's order (wrapped components first).
struct GhTfuncs {
  double (**fns)(double);
  int used;
  double ta, tb;
};

// Piecewise Chebyshev table of f over [ta, tb] in T->ext (layout in
// gh_tfunc_eval): 18 coefficients per segment, computed in long double from
// f at the Chebyshev nodes; segments doubled from 16 until f at three
// off-node points per segment is reproduced to 4e-16 of max |f| (or 2048
// segments, for functions that are not smooth).
int build_tfunc_table(double (*f)(double), double ta, double tb, GhPotTable* T, int& off) {
  const int nc = 18;
  if (!(tb > ta)) {
    ta -= 1.0;
    tb += 1.0;
  }
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int nseg = 16;; nseg *= 2) {
    if (T->next + 5 + nseg * nc > GH_EXT_LEN) {
      set_error("time-function table of %d segments does not fit the potential table", nseg);
      return -1;
    }
    double* e = T->ext + T->next;
    const long double h = ((long double)tb - ta) / nseg;
    double err = 0.0, scale = 0.0;
    for (int k = 0; k < nseg; k++) {
      long double fv[nc];
      for (int j = 0; j < nc; j++) {
        const long double x = cosl(pi * (j + 0.5L) / nc);
        fv[j] = f((double)(ta + h * (k + 0.5L * (x + 1.0L))));
      }
      double* c = e + 5 + k * nc;
      for (int m = 0; m < nc; m++) {
        long double sum = 0.0L;
        for (int j = 0; j < nc; j++) sum += fv[j] * cosl(pi * m * (j + 0.5L) / nc);
        c[m] = (double)(sum * 2.0L / nc * (m == 0 ? 0.5L : 1.0L));
      }
      for (double u : {-0.8, 0.1, 0.77}) {
        long double b1 = 0.0L, b2 = 0.0L;
        for (int m = nc - 1; m >= 1; m--) {
          const long double b0 = 2.0L * u * b1 - b2 + c[m];
          b2 = b1;
          b1 = b0;
        }
        const double approx = (double)(u * b1 - b2 + c[0]);
        const double exact = f((double)(ta + h * (k + 0.5L * (u + 1.0L))));
        err = std::max(err, std::fabs(approx - exact));
        scale = std::max(scale, std::fabs(exact));
      }
    }
    if (err <= 4e-16 * scale || nseg >= 2048) {
      e[0] = ta;
      e[1] = tb;
      e[2] = nseg;
      e[3] = nc;
      e[4] = (double)(nseg / ((long double)tb - ta));
      off = T->next;
      T->next += 5 + nseg * nc;
      return 0;
    }
  }
}


// n components in galpy's order (parse_leapFuncArgs_Full,
// integrateFullOrbit.c:55-1045): a wrapper's code is followed by its wrapped
// components' codes; its arguments are nwrapped, the wrapped components'
// arguments, then its own (:1010-1031).  Leaves go to T->comp in the order
// they appear (galpy's summation order); every leaf lists the wrappers
// around it, outermost first.
// NonInertialFrameForce (integrateFullOrbit.c:689-721; packing
// integratePlanarOrbit.py:835-965, the same in both packings): pot_type 39
// takes 23 arguments (amp, 10 cache slots, rot_acc, lin_acc, omegaz_only,
// const_freq, Omega_as_func, Omega[3], Omega^2, Omegadot[3]) and its time
// functions from pot_tfuncs (a0[3], x0[3] and v0[3] when it also rotates,
// then Omega and Omegadot); 45 prefixes n_spline, nPts, tgrid[nPts] and the
// n_spline sampled functions (natural splines, initNonInertialFrameForceSplines
// :1435-1462) and appends tmin, tmax.  Each time input becomes a table or a
// spline in ext; their offsets are listed at C.ext.
int decode_noninertial(int code, const double*& pot_args, GhPotTable* T, int i, GhTfuncs* tf) {
  GhComponent& C = T->comp[i];
  const double* a = pot_args;
  int nin = 0;
  std::vector<int> offs;
  if (code == GH_NONINERTIALINTERP) {
    const int ns = (int)a[0], npts = (int)a[1];
    if (ns < 0 || npts < 2) {
      set_error("NonInertialFrameForce with %d splines of %d points", ns, npts);
      return -1;
    }
    const double* tg = a + 2;
    for (int k = 0; k < ns; k++) {
      int off = 0;
      if (build_spline(tg, tg + (long)(k + 1) * npts, npts, T, off) != 0) return -1;
      offs.push_back(off);
    }
    nin = ns;
    a += 2 + (long)(1 + ns) * npts;
  } else {
    const int ntf = (int)(3 * a[12] * (1 + 2 * a[11]) + (6 - 4 * a[13]) * a[15]);
    if (ntf > 0 && (!tf || !tf->fns)) {
      set_error("NonInertialFrameForce needs its %d time functions (pot_tfuncs)", ntf);
      return -1;
    }
    for (int k = 0; k < ntf; k++) {
      int off = 0;
      if (build_tfunc_table(tf->fns[tf->used++], tf->ta, tf->tb, T, off) != 0) return -1;
      offs.push_back(off);
    }
    nin = ntf;
  }
  if (T->next + nin > GH_EXT_LEN) {
    set_error("NonInertialFrameForce tables do not fit the potential table");
    return -1;
  }
  C.type = code;
  C.p[0] = a[0];
  for (int j = 0; j < 5; j++) C.p[1 + j] = a[11 + j];
  C.p[6] = code == GH_NONINERTIALINTERP ? 1.0 : 0.0;
  for (int j = 0; j < 7; j++) C.c[j] = a[16 + j];
  C.p[7] = code == GH_NONINERTIALINTERP ? a[23] : 0.0;
  C.c[7] = code == GH_NONINERTIALINTERP ? a[24] : 0.0;
  C.ext = T->next;
  for (int k = 0; k < nin; k++) T->ext[T->next + k] = offs[k];
  T->next += nin;
  C.nonaxi = 1;
  T->nonaxi = 1;
  T->ndiss++;
  T->has_wrap = 1;  // extended dispatch instances
  pot_args = a + (code == GH_NONINERTIALINTERP ? 25 : 23);
  return 0;
}

int decode_components(int n, const int*& pot_type, const double*& pot_args, GhPotTable* T,
                      int& nleaf, int& nslots, int depth, bool planar, GhTfuncs* tf) {
  for (int k = 0; k < n; k++) {
    const int code = *pot_type++;
    if (code >= 0) {
      if (nleaf >= GH_MAX_POT) {
        set_error("more than %d components (wrapped ones included)", GH_MAX_POT);
        return -1;
      }
      if (code == GH_NONINERTIAL || code == GH_NONINERTIALINTERP) {
        if (decode_noninertial(code, pot_args, T, nleaf, tf) != 0) return -1;
      } else if (decode_component(code, pot_args, T, nleaf, nslots, planar) != 0) {
        return -1;
      }
      nleaf++;
      continue;
    }
    if (code == GH_CHANDRASEKHAR && planar) {  // galpy has no planar dynamical friction
      set_error("ChandrasekharDynamicalFrictionForce has no planar version");
      return -1;
    }
    if (code == GH_FDM && planar) {
      set_error("FDMDynamicalFrictionForce has no planar version");
      return -1;
    }
    const int na = (code == GH_MOVINGOBJWRAP || code == GH_CHANDRASEKHAR || code == GH_FDM)
                       ? 0
                       : nargs_of(code, planar);
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
    if (decode_components(nw, pot_type, pot_args, T, nleaf, nslots, depth + 1, planar, tf) != 0)
      return -1;
    const int w = T->nwraps++;
    GhWrap& W = T->wrap[w];
    W.kind = code;
    W.first = first;
    if (code == GH_FDM) {
      // FDMDynamicalFrictionForce (integrateFullOrbit.c:973-985, packing
      // integrateFullOrbit.py:468-515): it wraps the Chandrasekhar force of
      // the same satellite and scales its amplitude; its own arguments repeat
      // the sigma_r samples and the 16, then m/hbar and the constant factor
      // (< 0: none).  The wrapped Chandrasekhar entry takes the two numbers.
      const double* a = pot_args;
      const int n = (int)a[0];
      const double* own = a + 1 + 2L * n;
      GhWrap* C7 = nullptr;
      for (int k = T->nwraps - 2; k >= 0 && !C7; k--)
        if (T->wrap[k].kind == GH_CHANDRASEKHAR && T->wrap[k].first >= first) C7 = &T->wrap[k];
      if (!C7 || nw != 1) {
        set_error("FDMDynamicalFrictionForce must wrap one ChandrasekharDynamicalFrictionForce");
        return -1;
      }
      C7->p[10] = own[16];
      C7->p[11] = own[17];
      C7->p[12] = 1.0;
      pot_args = own + 18;
      continue;
    }
    if (code == GH_CHANDRASEKHAR) {
      // ChandrasekharDynamicalFrictionForce (integrateFullOrbit.c:917-927,
      // initChandrasekharDynamicalFrictionSplines :1464-1491): nPts, r, sigma_r
      // (a natural spline over (r - ro) / (rf - ro)), then amp, 8 cache slots,
      // ms, rhm, gamma^2, lnLambda, minr^2, ro, rf.  Its wrapped components
      // give only the density, not forces.
      const double* a = pot_args;
      const int n = (int)a[0];
      const double* ra = a + 1;
      const double* own = ra + 2L * n;
      std::vector<double> rn(n);
      for (int j = 0; j < n; j++) rn[j] = (ra[j] - own[14]) / (own[15] - own[14]);
      int off = 0;
      if (build_spline(rn.data(), ra + n, n, T, off) != 0) return -1;
      W.p[0] = own[0];
      for (int j = 0; j < 7; j++) W.p[1 + j] = own[9 + j];  // ms, rhm, gamma^2, lnLambda, minr^2, ro, rf
      W.p[8] = off;
      W.p[9] = nleaf - first;
      for (int l = first; l < nleaf; l++) {
        if (!gh_has_density(T->comp[l].type) || T->comp[l].nwrap) {
          set_error("ChandrasekharDynamicalFrictionForce: no device density for pot_type %d",
                    T->comp[l].type);
          return -1;
        }
        T->comp[l].densonly = 1;
      }
      T->ndiss++;
      T->has_wrap = 1;
      pot_args = own + 16;
      continue;
    }
    if (code == GH_MOVINGOBJWRAP) {
      // initMovingObjectSplines (integrateFullOrbit.c:1392-1431; planar
      // integratePlanarOrbit.c:1177-1215): nPts, t, x, y[, z], amp, t0, tf;
      // natural splines of the positions over (t - t0) / (tf - t0)
      const double* a = pot_args;
      const int n = (int)a[0], nc = planar ? 2 : 3;
      const double* tarr = a + 1;
      const double to = tarr[(nc + 1) * n + 1], tf = tarr[(nc + 1) * n + 2];
      std::vector<double> tt(n);
      for (int j = 0; j < n; j++) tt[j] = (tarr[j] - to) / (tf - to);
      W.p[0] = tarr[(nc + 1) * n];
      W.p[1] = to;
      W.p[2] = tf;
      W.p[5] = -1.0;
      for (int k = 0; k < nc; k++) {
        int off = 0;
        if (build_spline(tt.data(), tarr + (k + 1) * n, n, T, off) != 0) return -1;
        W.p[3 + k] = off;
      }
      pot_args += 1 + (nc + 1) * n + 3;
    } else if (code == GH_ROTATETILTWRAP) {  // amp, rot, rotSet, offsetSet, offset (caches skipped)
      const double* a = pot_args;
      W.p[0] = a[0];
      for (int j = 0; j < 14; j++) W.p[1 + j] = a[8 + j];
      pot_args += na;
    } else {
      for (int j = 0; j < na; j++) W.p[j] = *pot_args++;
    }
    if (code == GH_TIMEDEPAMPWRAP) {  // its time function A(t), tabulated
      if (!tf || !tf->fns) {
        set_error("TimeDependentAmplitudeWrapperPotential needs its time function (pot_tfuncs)");
        return -1;
      }
      int off = 0;
      if (build_tfunc_table(tf->fns[tf->used++], tf->ta, tf->tb, T, off) != 0) return -1;
      W.p[1] = off;
    }
    if (code == GH_OBLATESTAECKELWRAP) {
      W.p[5] = planar ? 1.0 : 0.0;  // the planar force (no V term)
      if (!(W.p[1] > 0.0)) {
        set_error("OblateStaeckelWrapperPotential delta=%g must be positive", W.p[1]);
        return -1;
      }
    }
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
                     bool planar = false, GhTfuncs* tf = nullptr) {
  std::memset(T, 0, offsetof(GhPotTable, ext));  // ext is read only up to next
  if (npot < 1 || npot > GH_MAX_POT) {
    set_error("npot=%d outside [1, %d]", npot, GH_MAX_POT);
    return -1;
  }
  int nleaf = 0, nslots = 0;
  const int* pt = pot_type;
  const double* pa = pot_args;
  if (decode_components(npot, pt, pa, T, nleaf, nslots, 0, planar, tf) != 0) return -1;
  T->npot = nleaf;
  // leaves whose wrappers move the evaluation point take the general walk
  // (two passes under a CylindricallySeparable wrapper)
  for (int l = 0; l < nleaf; l++) {
    GhComponent& C = T->comp[l];
    int moved = 0, cylsep = 0, oblate = 0;
    for (int k = 0; k < C.nwrap; k++) {
      const int kind = T->wrap[C.wrap[k]].kind;
      moved |= kind == GH_ROTATETILTWRAP || kind == GH_KUZMINLIKEWRAP || kind == GH_CYLSEPWRAP ||
               kind == GH_MOVINGOBJWRAP || kind == GH_OBLATESTAECKELWRAP;
      cylsep += kind == GH_CYLSEPWRAP;
      oblate += kind == GH_OBLATESTAECKELWRAP;
      if (kind == GH_OBLATESTAECKELWRAP && k != C.nwrap - 1) {
        set_error("an OblateStaeckelWrapperPotential around another wrapper is not supported");
        return -1;
      }
    }
    if (cylsep + oblate > 1) {
      set_error("a component under more than one CylindricallySeparable or OblateStaeckel wrapper");
      return -1;
    }
    C.mapwrap = moved ? 1 + cylsep + oblate : 0;
    C.mappot = oblate;
    if (moved) T->nonaxi = 1;
  }
  return 0;
}
// ---------------------------------------------------------------------------
// Evaluation kernels along stored rows and at points (not integrators)
// ---------------------------------------------------------------------------
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
// Compositions with a specialised kernel: MWPotential2014, the rotating-bar
// model, and single Hernquist / logarithmic halos; anything else runs the
// generic kernel that dispatches on the component type at run time.
// V = GhVec3 integrates full orbits, GhVec2 planar ones.
template <class V>
int launch_integrate(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                     const GhPotHandle* h, double dt, double rtol, double atol, double* d_result,
                     int* d_err, int odeint_type, double* d_stats, hipStream_t stream,
                     const GhWorkspace* ws = nullptr) {
  if (nobj == 0) return 0;
  if (nobj < 0 || nt < 2) {
    set_error("invalid sizes nobj=%d nt=%d (need nt >= 2)", nobj, nt);
    return -1;
  }
  if (h->ndiss && (odeint_type == 0 || odeint_type == 3 || odeint_type == 4 || odeint_type == 7)) {
    // the symplectic and IAS15 integrators take position-only forces
    // (evalRectForce, integrateFullOrbit.c:1091-1132); galpy reroutes them
    set_error("velocity-dependent forces need odeint_type 1, 2, 5 or 6 (RK4, RK6, DOPR54, DOP853), "
              "not %d", odeint_type);
    return -1;
  }
#define GH_LAUNCH(SPEC)                                                                        \
  return launch_spec<SPEC, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result, \
                              d_err, odeint_type, d_stats, stream, ws)
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
    set_error("integrate_dxdv needs components with a Hessian (every supported type but "
              "KuzminKutuzovStaeckel, SoftenedNeedleBar and the planar-only disks, unwrapped)");
    return -1;
  }
  using MW2014 = GhSpec<GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW>;
  if (types_are(h, {GH_POWERSPHCUTOFF, GH_MIYAMOTONAGAI, GH_NFW}))
    return launch_spec<MW2014, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result,
                                  d_err, odeint_type, d_stats, stream, nullptr);
  return launch_spec<GhSpec<>, V>(nobj, d_yo, nt, d_t, indiv_t, h->dev, dt, rtol, atol, d_result,
                                  d_err, odeint_type, d_stats, stream, nullptr);
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
                   double* pot_args, double (**tfuncs)(double), double dt, double rtol,
                   double atol, double* result, int* err, int odeint_type, void (*cb)(void),
                   int mode) {
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
  std::unique_ptr<GhPotTable> table_ptr(new GhPotTable);
  GhPotTable& table = *table_ptr;
  // time functions are tabulated over the output times, 5% wider each side
  GhTfuncs tf = {tfuncs, 0, 0.0, 0.0};
  if (tfuncs) {
    const size_t nall = (size_t)nt * (indiv_t ? nobj : 1);
    double lo = t[0], hi = t[0];
    for (size_t k = 1; k < nall; k++) {
      lo = std::min(lo, t[k]);
      hi = std::max(hi, t[k]);
    }
    const double pad = 0.05 * (hi - lo);
    tf.ta = lo - pad;
    tf.tb = hi + pad;
  }
  if (decode_potential(npot, pot_type, pot_args, &table, planar, &tf) != 0) return -1;
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
  action.sa_flags = SA_RESTART;  // the runtime's own waits resume after the handler
  irq_device_ptr();
  if (g_irq_host) *g_irq_host = 0;
  sigaction(SIGINT, &action, &old_action);
  do {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { rc = -1; set_error("hipMemGetInfo failed"); break; }
    const size_t per_orbit = (size_t)nt * N * 8 + N * 8 + 4 + (indiv_t ? (size_t)nt * 8 : 0);
    size_t budget = free_b / 2;
    long chunk = (long)std::min<size_t>((size_t)nobj, std::max<size_t>(1, budget / per_orbit));
    if (hipMalloc(&d_table, gh_table_alloc_bytes(table)) != hipSuccess ||
        hipMalloc((void**)&d_yo, (size_t)chunk * N * 8) != hipSuccess ||
        hipMalloc((void**)&d_t, (size_t)nt * 8 * (indiv_t ? chunk : 1)) != hipSuccess ||
        hipMalloc((void**)&d_res, (size_t)chunk * nt * N * 8) != hipSuccess ||
        hipMalloc((void**)&d_err, (size_t)chunk * 4) != hipSuccess) {
      rc = -1;
      set_error("device allocation failed (chunk of %ld orbits, nt=%d)", chunk, nt);
      break;
    }
    if (hipMemcpyAsync(d_table, &table, gh_table_bytes(table), hipMemcpyHostToDevice, stream) != hipSuccess) { rc = -1; break; }
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

int launch_full_orbits(int nobj, double* d_yo, int nt, const double* d_t, int indiv_t,
                       const GhPotHandle* h, double dt, double rtol, double atol, double* d_result,
                       int* d_err, int odeint_type, double* d_stats, hipStream_t stream,
                       const GhWorkspace* ws) {
  return launch_integrate<GhVec3>(nobj, d_yo, nt, d_t, indiv_t, h, dt, rtol, atol, d_result, d_err,
                                  odeint_type, d_stats, stream, ws);
}

const int* irq_device_ptr_if_ready() { return g_irq_dev; }

}  // namespace gh

using namespace gh;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static void integrate_host_abi(const char* name, int nobj, double* yo, int nt, double* t,
                               int indiv_t, int npot, int* pot_type, double* pot_args,
                               void* pot_tfuncs, double dt, double rtol, double atol,
                               double* result, int* err, int odeint_type, void (*cb)(void),
                               int mode) {
  g_errbuf[0] = 0;
  for (int i = 0; i < nobj; i++) err[i] = 0;
  const int rc = host_integrate(nobj, yo, nt, t, indiv_t, npot, pot_type, pot_args,
                                (double (**)(double))pot_tfuncs, dt, rtol, atol, result, err,
                                odeint_type, cb, mode);
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
                            bool planar, GhTfuncs* tf = nullptr) {
  g_errbuf[0] = 0;
  std::unique_ptr<GhPotTable> table_ptr(new GhPotTable);
  GhPotTable& table = *table_ptr;
  if (decode_potential(npot, pot_type, pot_args, &table, planar, tf) != 0) return -1;
  void* d = nullptr;
  GH_CHECK(hipMalloc(&d, gh_table_alloc_bytes(table)));
  GH_CHECK(hipMemcpy(d, &table, gh_table_bytes(table), hipMemcpyHostToDevice));
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

// Host evaluation of the device's time-function table of fns[0] over
// [t0, t1] at ts[n] (for CPU tests of the table).
GH_EXPORT int gh_host_tfunc_eval(double (**fns)(double), double t0, double t1, int n,
                                 const double* ts, double* out) {
  g_errbuf[0] = 0;
  static GhPotTable T;
  T.next = 0;
  int off = 0;
  if (build_tfunc_table(fns[0], std::min(t0, t1), std::max(t0, t1), &T, off) != 0) return -1;
  for (int k = 0; k < n; k++) out[k] = gh_tfunc_eval(&T, off, ts[k]);
  return 0;
}

GH_EXPORT int gh_potential_upload_tfuncs(int npot, const int* pot_type, const double* pot_args,
                                         double (**pot_tfuncs)(double), double t0, double t1,
                                         int planar, void** pot) {
  GhTfuncs tf = {pot_tfuncs, 0, std::min(t0, t1), std::max(t0, t1)};
  return upload_potential(npot, pot_type, pot_args, pot, planar != 0, &tf);
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

GH_EXPORT int gh_host_sincos(int n, const double* x, double* s, double* c) {
  for (int i = 0; i < n; i++) {
    const GhSinCos v = gh_sincos(x[i]);
    s[i] = v.s;
    c[i] = v.c;
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