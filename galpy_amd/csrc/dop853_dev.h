// Device DOP853 (Dormand-Prince 8(5,3) with dense output), one orbit per lane.
//
// Follows galpy/util/leung_dop853.c:231-517 step for step: the initial step
// heuristic, the 12 stages, the mixed 5th/3rd-order error norm, the step-size
// controller (safe=0.9, beta=0, fac1=0.333, fac2=6) and the 7th-order dense
// output evaluated at every requested time passed by an accepted step.  Each
// lane owns its own step size, so the accepted/rejected sequence of every
// orbit is galpy's; lanes of a wavefront execute the stages together and only
// the accept branch (FSAL evaluation + dense output) is masked per lane.
// The tableau is the published one (Hairer, Norsett & Wanner, dop853.f).
#pragma once
#include "integrators_dev.h"

namespace ghdp {
constexpr double c2 = 0.526001519587677318785587544488e-1, c3 = 0.789002279381515978178381316732e-1,
  c4 = 0.118350341907227396726757197510, c5 = 0.281649658092772603273242802490,
  c6 = 0.333333333333333333333333333333, c7 = 0.25, c8 = 0.307692307692307692307692307692,
  c9 = 0.651282051282051282051282051282, c10 = 0.6, c11 = 0.857142857142857142857142857142,
  c14 = 0.1, c15 = 0.2, c16 = 0.777777777777777777777777777778;
constexpr double a21 = 5.26001519587677318785587544488e-2,
  a31 = 1.97250569845378994544595329183e-2, a32 = 5.91751709536136983633785987549e-2,
  a41 = 2.95875854768068491816892993775e-2, a43 = 8.87627564304205475450678981324e-2,
  a51 = 2.41365134159266685502369798665e-1, a53 = -8.84549479328286085344864962717e-1,
  a54 = 9.24834003261792003115737966543e-1,
  a61 = 3.7037037037037037037037037037e-2, a64 = 1.70828608729473871279604482173e-1,
  a65 = 1.25467687566822425016691814123e-1,
  a71 = 3.7109375e-2, a74 = 1.70252211019544039314978060272e-1,
  a75 = 6.02165389804559606850219397283e-2, a76 = -1.7578125e-2,
  a81 = 3.70920001185047927108779319836e-2, a84 = 1.70383925712239993810214054705e-1,
  a85 = 1.07262030446373284651809199168e-1, a86 = -1.53194377486244017527936158236e-2,
  a87 = 8.27378916381402288758473766002e-3,
  a91 = 6.24110958716075717114429577812e-1, a94 = -3.36089262944694129406857109825e0,
  a95 = -8.68219346841726006818189891453e-1, a96 = 2.75920996994467083049415600797e1,
  a97 = 2.01540675504778934086186788979e1, a98 = -4.34898841810699588477366255144e1,
  a101 = 4.77662536438264365890433908527e-1, a104 = -2.48811461997166764192642586468e0,
  a105 = -5.90290826836842996371446475743e-1, a106 = 2.12300514481811942347288949897e1,
  a107 = 1.52792336328824235832596922938e1, a108 = -3.32882109689848629194453265587e1,
  a109 = -2.03312017085086261358222928593e-2,
  a111 = -9.3714243008598732571704021658e-1, a114 = 5.18637242884406370830023853209e0,
  a115 = 1.09143734899672957818500254654e0, a116 = -8.14978701074692612513997267357e0,
  a117 = -1.85200656599969598641566180701e1, a118 = 2.27394870993505042818970056734e1,
  a119 = 2.49360555267965238987089396762e0, a1110 = -3.0467644718982195003823669022e0,
  a121 = 2.27331014751653820792359768449e0, a124 = -1.05344954667372501984066689879e1,
  a125 = -2.00087205822486249909675718444e0, a126 = -1.79589318631187989172765950534e1,
  a127 = 2.79488845294199600508499808837e1, a128 = -2.85899827713502369474065508674e0,
  a129 = -8.87285693353062954433549289258e0, a1210 = 1.23605671757943030647266201528e1,
  a1211 = 6.43392746015763530355970484046e-1,
  a141 = 5.61675022830479523392909219681e-2, a147 = 2.53500210216624811088794765333e-1,
  a148 = -2.46239037470802489917441475441e-1, a149 = -1.24191423263816360469010140626e-1,
  a1410 = 1.5329179827876569731206322685e-1, a1411 = 8.20105229563468988491666602057e-3,
  a1412 = 7.56789766054569976138603589584e-3, a1413 = -8.298e-3,
  a151 = 3.18346481635021405060768473261e-2, a156 = 2.83009096723667755288322961402e-2,
  a157 = 5.35419883074385676223797384372e-2, a158 = -5.49237485713909884646569340306e-2,
  a1511 = -1.08347328697249322858509316994e-4, a1512 = 3.82571090835658412954920192323e-4,
  a1513 = -3.40465008687404560802977114492e-4, a1514 = 1.41312443674632500278074618366e-1,
  a161 = -4.28896301583791923408573538692e-1, a166 = -4.69762141536116384314449447206e0,
  a167 = 7.68342119606259904184240953878e0, a168 = 4.06898981839711007970213554331e0,
  a169 = 3.56727187455281109270669543021e-1, a1613 = -1.39902416515901462129418009734e-3,
  a1614 = 2.9475147891527723389556272149e0, a1615 = -9.15095847217987001081870187138e0;
constexpr double b1 = 5.42937341165687622380535766363e-2, b6 = 4.45031289275240888144113950566,
  b7 = 1.89151789931450038304281599044, b8 = -5.8012039600105847814672114227,
  b9 = 3.1116436695781989440891606237e-1, b10 = -1.52160949662516078556178806805e-1,
  b11 = 2.01365400804030348374776537501e-1, b12 = 4.47106157277725905176885569043e-2;
constexpr double bhh1 = 0.244094488188976377952755905512, bhh2 = 0.733846688281611857341361741547,
  bhh3 = 0.220588235294117647058823529412e-1;
constexpr double er1 = 0.1312004499419488073250102996e-1, er6 = -0.1225156446376204440720569753e+1,
  er7 = -0.4957589496572501915214079952, er8 = 0.1664377182454986536961530415e+1,
  er9 = -0.3503288487499736816886487290, er10 = 0.3341791187130174790297318841,
  er11 = 0.8192320648511571246570742613e-1, er12 = -0.2235530786388629525884427845e-1;
constexpr double d41 = -0.84289382761090128651353491142e+1, d46 = 0.56671495351937776962531783590,
  d47 = -0.30689499459498916912797304727e+1, d48 = 0.23846676565120698287728149680e+1,
  d49 = 0.21170345824450282767155149946e+1, d410 = -0.87139158377797299206789907490,
  d411 = 0.22404374302607882758541771650e+1, d412 = 0.63157877876946881815570249290,
  d413 = -0.88990336451333310820698117400e-1, d414 = 0.18148505520854727256656404962e+2,
  d415 = -0.91946323924783554000451984436e+1, d416 = -0.44360363875948939664310572000e+1;
constexpr double d51 = 0.10427508642579134603413151009e+2, d56 = 0.24228349177525818288430175319e+3,
  d57 = 0.16520045171727028198505394887e+3, d58 = -0.37454675472269020279518312152e+3,
  d59 = -0.22113666853125306036270938578e+2, d510 = 0.77334326684722638389603898808e+1,
  d511 = -0.30674084731089398182061213626e+2, d512 = -0.93321305264302278729567221706e+1,
  d513 = 0.15697238121770843886131091075e+2, d514 = -0.31139403219565177677282850411e+2,
  d515 = -0.93529243588444783865713862664e+1, d516 = 0.35816841486394083752465898540e+2;
constexpr double d61 = 0.19985053242002433820987653617e+2, d66 = -0.38703730874935176555105901742e+3,
  d67 = -0.18917813819516756882830838328e+3, d68 = 0.52780815920542364900561016686e+3,
  d69 = -0.11573902539959630126141871134e+2, d610 = 0.68812326946963000169666922661e+1,
  d611 = -0.10006050966910838403183860980e+1, d612 = 0.77771377980534432092869265740,
  d613 = -0.27782057523535084065932004339e+1, d614 = -0.60196695231264120758267380846e+2,
  d615 = 0.84320405506677161018159903784e+2, d616 = 0.11992291136182789328035130030e+2;
constexpr double d71 = -0.25693933462703749003312586129e+2, d76 = -0.15418974869023643374053993627e+3,
  d77 = -0.23152937917604549567536039109e+3, d78 = 0.35763911791061412378285349910e+3,
  d79 = 0.93405324183624310003907691704e+2, d710 = -0.37458323136451633156875139351e+2,
  d711 = 0.10409964950896230045147246184e+3, d712 = 0.29840293426660503123344363579e+2,
  d713 = -0.43533456590011143754432175058e+2, d714 = 0.96324553959188282948394950600e+2,
  d715 = -0.39177261675615439165231486172e+2, d716 = -0.14972683625798562581422125276e+3;
constexpr double uround = 2.3e-16;
}  // namespace ghdp

#include "dop853_rkn.h"  // Nystrom-form sums over this tableau (generated)

// RHS of the first-order system: (v, a(q)) -- evalRectDeriv,
// integrateFullOrbit.c:1326-1351 (evalPlanarRectDeriv,
// integratePlanarOrbit.c:1032-1055, for planar orbits)
template <class F, class V>
__device__ __forceinline__ void gh_deriv(const F& T, double t, const double* y, double* k) {
  GhPhase<V>::set(k, GhPhase<V>::p(y), gh_accel_v(T, t, GhPhase<V>::q(y), GhPhase<V>::p(y)));
}

__device__ __forceinline__ double gh_csign(double x, double y) {
  return (y > 0.0) ? fabs(x) : -fabs(x);
}

// Per-lane DOP853 state between steps (leung_dop853.c:231-517 locals), for
// a full (V = GhVec3, dim 6) or planar (V = GhVec2, dim 4) orbit.
template <class V>
struct GhDopStateT {
  double y[GhPhase<V>::N];   // current state (galpy's y0)
  double k1[GhPhase<V>::N];  // derivative at y (FSAL)
  double h, t_cur, t_old, t_old_older, hmax, pos_neg;
  int done;      // output rows written after row 0
  int reject;
  long long nacc, nrej;
};
using GhDopState = GhDopStateT<GhVec3>;

// Start an orbit: row 0, first derivative and the initial step heuristic
// (leung_dop853.c:276-325).  s.y holds the rectangular initial condition.
// STORE = false (a pilot run that only counts steps) writes no output rows.
template <class F, bool STORE = true, class V>
__device__ __forceinline__ void gh_dop_init(const F& T,
                                            const double* __restrict__ t, int nt, double rtol,
                                            double atol, GhDopStateT<V>& s,
                                            double* __restrict__ out) {
  constexpr int N = GhPhase<V>::N;
  s.hmax = t[nt - 1] - t[0];
  s.pos_neg = gh_csign(1.0, s.hmax);
  s.done = 0;
  s.reject = 0;
  s.nacc = 0;
  s.nrej = 0;
  if (STORE) GhPhase<V>::store(out, GhPhase<V>::q(s.y), GhPhase<V>::p(s.y));
  gh_deriv<F, V>(T, t[0], s.y, s.k1);
  double dnf = 0.0, dny = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double sk = atol + rtol * fabs(s.y[i]);
    double sqr = s.k1[i] / sk;
    dnf += sqr * sqr;
    sqr = s.y[i] / sk;
    dny += sqr * sqr;
  }
  double h = gh_csign(fmin(sqrt(dny / dnf) * 0.01, fabs(s.hmax)), s.pos_neg);
  double k2[N], k3[N];
#pragma unroll
  for (int i = 0; i < N; i++) k3[i] = s.y[i] + h * s.k1[i];
  gh_deriv<F, V>(T, t[0] + h, k3, k2);
  double der2 = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double sk = atol + rtol * fabs(s.y[i]);
    const double sqr = (k2[i] - s.k1[i]) / sk;
    der2 += sqr * sqr;
  }
  der2 = sqrt(der2) / h;
  const double der12 = fmax(fabs(der2), sqrt(dnf));
  const double h1 = pow(0.01 / der12, 1.0 / 8.0);
  s.h = gh_csign(fmin(100.0 * fabs(h), fmin(fabs(h1), fabs(s.hmax))), s.pos_neg);
  s.t_cur = t[0];
  s.t_old = t[0];
  s.t_old_older = t[0];
}

// The step and the dense output are written in Nystrom form (dop853_rkn.h,
// generated from the tableau above): the twelve stages keep only their
// accelerations A[j] and form only their positions, since for y' = (p, a(q))
// the stage velocities are linear in p0 and the A's.  This is the same method
// as leung_dop853.c, with the position sums regrouped (same value in exact
// arithmetic); it halves the registers the stages occupy.

// Stage velocities of the first-order form, p0 + h sum_j a_ij A_j with the
// sums in leung_dop853.c's order (:351-452): formed only when the forces
// depend on the velocity (the Nystrom sums above need no stage velocity).
template <class V>
__device__ __forceinline__ V gh_lin(V s) {
  return s;
}
template <class V, class... Rest>
__device__ __forceinline__ V gh_lin(V s, double a, V x, Rest... rest) {
  return gh_lin(gh_drift(s, x, a), rest...);
}
template <class V, class... Rest>
__device__ __forceinline__ V gh_pstage(double h, V p0, double a, V x, Rest... rest) {
  return gh_drift(p0, gh_lin(gh_drift(V{}, x, a), rest...), h);
}
// a stage's acceleration, with its velocity p0 + h sum a_j A_j when the
// forces need it (the pairs by value: nothing of the stages is addressed,
// and without velocity-dependent forces the sum is dead code)
template <class F, class V, class... Rest>
__device__ __forceinline__ V gh_stage_acc(const F& T, double t, V q, double h, V p0, Rest... rest) {
  if (T.diss()) return gh_accel_v(T, t, q, gh_pstage(h, p0, rest...));
  return gh_accel(T, t, q);
}
template <class F, class V>
__device__ __forceinline__ V gh_stage_acc_pn(const F& T, double t, V q, V pn) {
  if (T.diss()) return gh_accel_v(T, t, q, pn);
  return gh_accel(T, t, q);
}

// What the dense output of one accepted step needs (leung_dop853.c:430-474):
// the step's start and end states, the accelerations of the stages the
// continuous extension uses (A1, A4..A13; A[j] of stage j, A13 = FSAL), the
// step's start time and size, and the number of output rows written before
// it.  A lane can hand this record to another lane (deferred dense output).
template <class V>
struct GhDenseJobT {
  V q0, p0, qn, pn;
  V A[11];  // stages 1, 4, 5, ..., 13
  double t_old, h;
  int done0;
};

// Dense output of one accepted step: the three extra stages 14-16 and the
// continuous extension at every requested time inside the step, written as
// rows done0+1, done0+2, ...  Returns the number of rows written.
template <class F, bool STORE = true, class V>
__device__ __forceinline__ int gh_dop_dense(const F& T, const double* __restrict__ t, int nt,
                                         const GhDenseJobT<V>& J, double* __restrict__ out) {
  using namespace ghdp;
  constexpr int N = GhPhase<V>::N;
  const double h = J.h;
  V A[17];
  A[1] = J.A[0];
#pragma unroll
  for (int j = 4; j <= 13; j++) A[j] = J.A[j - 3];
  A[2] = A[3] = V{};  // unused by stages 14-16 and the extension
  const V p0 = J.p0;
  A[14] = gh_stage_acc(T, J.t_old + c14 * h, gh_rkn_q14(h, J.q0, J.p0, J.pn, A), h, p0, a141, A[1], a147, A[7], a148, A[8], a149, A[9], a1410, A[10], a1411, A[11], a1412, A[12], a1413, A[13]);
  A[15] = gh_stage_acc(T, J.t_old + c15 * h, gh_rkn_q15(h, J.q0, J.p0, J.pn, A), h, p0, a151, A[1], a156, A[6], a157, A[7], a158, A[8], a1511, A[11], a1512, A[12], a1513, A[13], a1514, A[14]);
  A[16] = gh_stage_acc(T, J.t_old + c16 * h, gh_rkn_q16(h, J.q0, J.p0, J.pn, A), h, p0, a161, A[1], a166, A[6], a167, A[7], a168, A[8], a169, A[9], a1613, A[13], a1614, A[14], a1615, A[15]);
  // the continuous extension is re-formed per requested time (a step holds
  // one output time in the common case), so its 48 coefficients are never
  // live at once
  const double pos_neg = (h > 0.0) ? 1.0 : -1.0;
  const double t_cur = J.t_old + h;
  int done = J.done0;
  while (done < nt - 1 && pos_neg * t[done + 1] < pos_neg * t_cur) {
    const double sd = (t[done + 1] - J.t_old) / h;
    const double s1 = 1.0 - sd;
    double yo[N];  // positions, then velocities, in component order
    gh_rkn_dense_eval(h, sd, s1, J.q0, J.p0, J.qn, J.pn, A, yo);
    done++;
    if (STORE) GhPhase<V>::store(out + (long)N * done, gh_vec_at<V>(yo), gh_vec_at<V>(yo + N / 2));
  }
  return done - J.done0;
}

// Dense-output handling of gh_dop_step.
enum GhDenseMode {
  GH_DENSE_SKIP = 0,    // count the passed output times only (pilot runs)
  GH_DENSE_INLINE = 1,  // compute them in this lane now
  GH_DENSE_DEFER = 2    // hand the step's GhDenseJob to sink(J) and count them
};

struct GhNoSink {
  template <class J>
  __device__ void operator()(const J&) const {}
};

// One attempted step (leung_dop853.c:335-492); returns true once every
// requested time has been handled.  With GH_DENSE_DEFER, sink(J) is called
// (from inside the accept branch, by exactly the lanes whose step passed
// output times) as soon as the job is formed, and s.done already counts its
// rows.
template <class F, int DENSE = GH_DENSE_INLINE, class Sink = GhNoSink, class V>
__device__ __forceinline__ bool gh_dop_step(const F& T,
                                            const double* __restrict__ t, int nt, double rtol,
                                            double atol, GhDopStateT<V>& s,
                                            double* __restrict__ out,
                                            const Sink& sink = Sink()) {
  using namespace ghdp;
  constexpr int N = GhPhase<V>::N;
  const double safe = 0.9;
  const double facc1 = 1.0 / 0.333, facc2 = 1.0 / 6.0;
  const double h = s.pos_neg * fmax(fabs(s.h), 1e3 * uround);
  const double tc = s.t_cur;
  const V q0 = GhPhase<V>::q(s.y), p0 = GhPhase<V>::p(s.y);
  const V none{};
  // galpy's dop853 ignores the result of the twelve stages when rejected, so
  // galpy's accept/reject flow is kept exactly; only the accept branch is
  // masked per lane.  (galpy also tracks facold for Lund stabilisation, but
  // with beta = 0 pow(facold, beta) is exactly 1 and never enters a result.)
  V A[17];
  A[1] = GhPhase<V>::p(s.k1);
  A[2] = gh_stage_acc(T, tc + c2 * h, gh_rkn_q2(h, q0, p0, none, A), h, p0, a21, A[1]);
  A[3] = gh_stage_acc(T, tc + c3 * h, gh_rkn_q3(h, q0, p0, none, A), h, p0, a31, A[1], a32, A[2]);
  A[4] = gh_stage_acc(T, tc + c4 * h, gh_rkn_q4(h, q0, p0, none, A), h, p0, a41, A[1], a43, A[3]);
  A[5] = gh_stage_acc(T, tc + c5 * h, gh_rkn_q5(h, q0, p0, none, A), h, p0, a51, A[1], a53, A[3], a54, A[4]);
  A[6] = gh_stage_acc(T, tc + c6 * h, gh_rkn_q6(h, q0, p0, none, A), h, p0, a61, A[1], a64, A[4], a65, A[5]);
  A[7] = gh_stage_acc(T, tc + c7 * h, gh_rkn_q7(h, q0, p0, none, A), h, p0, a71, A[1], a74, A[4], a75, A[5], a76, A[6]);
  A[8] = gh_stage_acc(T, tc + c8 * h, gh_rkn_q8(h, q0, p0, none, A), h, p0, a81, A[1], a84, A[4], a85, A[5], a86, A[6], a87, A[7]);
  A[9] = gh_stage_acc(T, tc + c9 * h, gh_rkn_q9(h, q0, p0, none, A), h, p0, a91, A[1], a94, A[4], a95, A[5], a96, A[6], a97, A[7], a98, A[8]);
  A[10] = gh_stage_acc(T, tc + c10 * h, gh_rkn_q10(h, q0, p0, none, A), h, p0, a101, A[1], a104, A[4], a105, A[5], a106, A[6], a107, A[7], a108, A[8], a109, A[9]);
  A[11] = gh_stage_acc(T, tc + c11 * h, gh_rkn_q11(h, q0, p0, none, A), h, p0, a111, A[1], a114, A[4], a115, A[5], a116, A[6], a117, A[7], a118, A[8], a119, A[9], a1110, A[10]);
  const V q12 = gh_rkn_q12(h, q0, p0, none, A);
  s.t_old_older = s.t_old;
  s.t_old = tc;
  s.t_cur = tc + h;
  A[12] = gh_stage_acc(T, s.t_cur, q12, h, p0, a121, A[1], a124, A[4], a125, A[5], a126, A[6], a127, A[7], a128, A[8], a129, A[9], a1210, A[10], a1211, A[11]);
  V qn, pn;
  double err5, err3;
  gh_rkn_solution(h, q0, p0, A, atol, rtol, qn, pn, err5, err3);
  double deno = err5 + 0.01 * err3;
  if (deno <= 0.0) deno = 1.0;
  const double errv = fabs(h) * err5 * sqrt(1.0 / (deno * (double)N));
  // errv^(1/8) (leung_dop853.c uses pow) as three square roots
  const double fac11 = sqrt(sqrt(sqrt(errv)));
  const double fac = fmax(facc2, fmin(facc1, fac11 / safe));
  double hnew = h / fac;
  if (errv <= 1.0) {
    s.nacc++;
    A[13] = gh_stage_acc_pn(T, s.t_cur, qn, pn);  // FSAL: k13 = f(y_new)
    // dense output only when a requested time falls inside this step; the
    // three extra stages feed nothing else, so they are skipped otherwise --
    // outputs are unchanged
    if (s.pos_neg * t[s.done + 1] < s.pos_neg * s.t_cur) {
      if (DENSE == GH_DENSE_SKIP) {
        while (s.done < nt - 1 && s.pos_neg * t[s.done + 1] < s.pos_neg * s.t_cur) s.done++;
      } else {
        GhDenseJobT<V> J;
        J.q0 = q0;
        J.p0 = p0;
        J.qn = qn;
        J.pn = pn;
        J.A[0] = A[1];
#pragma unroll
        for (int j = 4; j <= 13; j++) J.A[j - 3] = A[j];
        J.t_old = tc;
        J.h = h;
        J.done0 = s.done;
        if (DENSE == GH_DENSE_INLINE) {
          s.done += gh_dop_dense<F, true>(T, t, nt, J, out);
        } else {
          sink(J);
          while (s.done < nt - 1 && s.pos_neg * t[s.done + 1] < s.pos_neg * s.t_cur) s.done++;
        }
      }
    }
    GhPhase<V>::set(s.y, qn, pn);
    GhPhase<V>::set(s.k1, pn, A[13]);
    hnew = (fabs(hnew) > fabs(s.hmax)) ? s.pos_neg * s.hmax : hnew;
    if (s.reject) hnew = s.pos_neg * fmin(fabs(hnew), fabs(h));
    s.reject = 0;
  } else {
    s.nrej++;
    hnew = h / fmin(facc1, fac11 / safe);
    s.reject = 1;
    s.t_cur = s.t_old;
    s.t_old = s.t_old_older;
  }
  s.h = hnew;
  return s.done >= nt - 1;
}
