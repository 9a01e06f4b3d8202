/*
 * galpy_hip.h -- C ABI of the MI355X (gfx950) orbit-integration engine.
 *
 * Drop-in boundary for galpy's C orbit integration: the host packs the
 * potential exactly as galpy's `_parse_pot` does (integer pot_type codes,
 * float64 pot_args), and calls one function for all orbits, as galpy's
 * Python layer calls `integrateFullOrbit` (galpy/orbit/integrateFullOrbit.py:
 * 687-734) or, for planar orbits, `integratePlanarOrbit`
 * (galpy/orbit/integratePlanarOrbit.py:1059-1110) through ctypes.  Plain
 * pointers and sizes only; no HIP or torch types appear in these signatures.
 *
 * Supported potentials (pot_type codes, integrateFullOrbit.c:80-345,
 * integratePlanarOrbit.c:64-223):
 *   0 LogarithmicHaloPotential, 1 DehnenBarPotential, 5 MiyamotoNagaiPotential,
 *   7 PowerSphericalPotential (and KeplerPotential), 8 HernquistPotential,
 *   9 NFWPotential, 10 JaffePotential, 14 IsochronePotential,
 *   15 PowerSphericalPotentialwCutoff, 17 PlummerPotential,
 *   18 PseudoIsothermalPotential, 19 KuzminDiskPotential, 20 BurkertPotential,
 *   33 DehnenCoreSphericalPotential, 34 DehnenSphericalPotential,
 *   35 HomogeneousSpherePotential, 12 FlattenedPowerPotential, 40 NullPotential,
 *   16 KuzminKutuzovStaeckelPotential, 25 SoftenedNeedleBarPotential (13
 *   arguments, the last 7 galpy's cache slots, unread); neither takes
 *   integrate_dxdv (no device Hessian); for planar orbits only, 4
 *   EllipticalDiskPotential, 6 LopsidedDiskPotential and 3
 *   SteadyLogSpiralPotential (integratePlanarOrbit.c's codes);
 * and the wrappers (integrateFullOrbit.c:821-895, packed as galpy packs them:
 * the wrapper's code, then the wrapped components' codes; as arguments
 * nwrapped, the wrapped components' arguments, then the wrapper's own)
 *   -1 DehnenSmoothWrapperPotential, -2 SolidBodyRotationWrapperPotential,
 *   -4 CorotatingRotationWrapperPotential, -5 GaussianAmplitudeWrapperPotential,
 * nested up to 3 deep; at most 8 (leaf) components and 8 wrappers in all.
 * npot counts the top-level components, as in galpy.
 * Supported odeint_type codes (integrateFullOrbit.c:1091-1132,
 * integratePlanarOrbit.c:788-825):
 *   0 leapfrog, 1 rk4, 2 rk6, 3 symplec4, 4 symplec6, 5 dopr54, 6 dop853, 7 ias15
 * (integrate_dxdv: 0, 3 and 6, as galpy's tangent-map steppers and DOP853).
 * Any other code, or nt < 2, sets err[i] = -1 for every orbit and records a
 * message readable with galpy_hip_last_error().
 */
#ifndef GALPY_HIP_H
#define GALPY_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces galpy/orbit/orbit_c_ext/integrateFullOrbit.c:1046-1061
 * (`integrateFullOrbit`), same signature and argument meaning:
 *   yo[nobj*6]      initial conditions (R,vR,vT,z,vz,phi); like galpy, it is
 *                   overwritten with the rectangular state (the initial one
 *                   for leapfrog/symplec4, the last integrator state for
 *                   dop853) -- callers pass a copy
 *   t[nt] or t[nobj*nt] (indiv_t=1) output times, equally spaced
 *   dt              fixed step, or -9999.99 to estimate one per orbit
 *   rtol, atol      natural logarithms of the tolerances
 *   result[nobj*nt*6] output rows (R,vR,vT,z,vz,phi); row 0 is the IC
 *   err[nobj]       0 ok, -10 interrupted by SIGINT, -1 unsupported input;
 *                   dopr54 as bovy_dopr54 (bovy_rk.c:501-616): odd when a
 *                   step was forced through at 1/10000 of the interval's first
 *                   step, -3 when one output interval took more than 2^22
 *                   step attempts (a bound the reference does not have);
 *                   ias15: -3 after 2^22 step attempts in all; rows past
 *                   the end of the run (a dt shorter than the grid spacing,
 *                   wez_ias15.c:280) are NaN, uninitialised in the reference
 *   pot_tfuncs      the components' time functions double f(double t), in
 *                   galpy's order (integrateFullOrbit.c:1034-1038), or NULL
 *                   when none takes one; each is tabulated on the host as a
 *                   piecewise Chebyshev series over the output times (5%
 *                   wider each side) and the device evaluates the table
 *   cb              called once per finished orbit (may be NULL)
 * All pointers are host memory.  Work is streamed through the GPU in chunks
 * of orbits sized to free device memory. */
void integrateFullOrbit(int nobj, double *yo, int nt, double *t, int indiv_t, int npot,
                        int *pot_type, double *pot_args, void *pot_tfuncs, double dt,
                        double rtol, double atol, double *result, int *err, int odeint_type,
                        void (*cb)(void));

/* Replaces galpy/orbit/orbit_c_ext/integratePlanarOrbit.c:743-758
 * (`integratePlanarOrbit`), same signature and argument meaning as
 * integrateFullOrbit above for planar orbits:
 *   yo[nobj*4]      initial conditions (R,vR,vT,phi), overwritten with the
 *                   rectangular state (x,y,vx,vy) as galpy does
 *   result[nobj*nt*4] output rows (R,vR,vT,phi); row 0 is the IC
 *   pot_type/pot_args in galpy's planar packing
 *                   (integratePlanarOrbit.py:33-663; integratePlanarOrbit.c:55-739):
 *                   the same codes, PowerSphericalPotentialwCutoff with 3
 *                   arguments (amp, alpha, rc) instead of 5
 * The planar forces are galpy's planar kernels (planarRforce,
 * planarphitorque), which equal the full forces in the plane z = 0. */
void integratePlanarOrbit(int nobj, double *yo, int nt, double *t, int indiv_t, int npot,
                          int *pot_type, double *pot_args, void *pot_tfuncs, double dt,
                          double rtol, double atol, double *result, int *err, int odeint_type,
                          void (*cb)(void));

/* Replaces integrateFullOrbit.c:1234-1306 (`integrateFullOrbit_dxdv`), same
 * signature and meaning: ONE orbit and one phase-space deviation, all
 * rectangular -- yo[12] = (x,y,z,vx,vy,vz, dx,dy,dz,dvx,dvy,dvz), result[nt*12]
 * rows in that layout, *err one int.  odeint_type 0 / 3: galpy's variational
 * steppers leapfrog_dxdv / symplec4_dxdv (bovy_symplecticode.c:1053-1268,
 * step estimated from the orbit alone, orbit bit-identical to a plain run);
 * 6: DOP853 on the 12-dim variational system (evalRectDeriv_dxdv,
 * integrateFullOrbit.c:1496-1600).  The kick of the deviation is K dq with K
 * the Cartesian Hessian of -Phi: every time-independent axisymmetric
 * component has one on the device; DehnenBar and a triaxial
 * LogarithmicHalo set *err = -1. */
void integrateFullOrbit_dxdv(double *yo, int nt, double *t, int npot, int *pot_type,
                             double *pot_args, void *pot_tfuncs, double dt, double rtol,
                             double atol, double *result, int *err, int odeint_type);

/* Replaces integratePlanarOrbit.c:940-1012 (`integratePlanarOrbit_dxdv`):
 * the planar twin, yo[8] = (x,y,vx,vy, dx,dy,dvx,dvy), result[nt*8], planar
 * packing; DOP853 on the 8-dim system (evalPlanarRectDeriv_dxdv,
 * integratePlanarOrbit.c:1124-1175).  cb is accepted and, as in galpy, not
 * called. */
void integratePlanarOrbit_dxdv(double *yo, int nt, double *t, int npot, int *pot_type,
                               double *pot_args, void *pot_tfuncs, double dt, double rtol,
                               double atol, double *result, int *err, int odeint_type,
                               void (*cb)(void));

/* integrate_dxdv for many orbits in one call (no reference counterpart:
 * galpy maps the single-orbit call over orbits from Python,
 * integrateFullOrbit.py:980-999, integratePlanarOrbit.py:1497-1516).
 * dim 3: yo[nobj*12], result[nobj*nt*12], full packing; dim 2: yo[nobj*8],
 * result[nobj*nt*8], planar packing.  t[nt] or t[nobj*nt] (indiv_t). */
void gh_integrate_dxdv(int nobj, int dim, double *yo, int nt, double *t, int indiv_t, int npot,
                       int *pot_type, double *pot_args, double dt, double rtol, double atol,
                       double *result, int *err, int odeint_type);
/* The same with device buffers, a potential handle (gh_potential_upload for
 * dim 3, gh_planar_potential_upload for dim 2) and a stream; d_stats as for
 * integrateFullOrbit_hip. */
int gh_integrate_dxdv_hip(int nobj, int dim, double *d_yo, int nt, const double *d_t, int indiv_t,
                          const void *pot, double dt, double rtol, double atol, double *d_result,
                          int *d_err, int odeint_type, double *d_stats, void *stream);

/* Device-resident twin of integrateFullOrbit for batching/serving callers
 * (no reference counterpart; it is integrateFullOrbit.c:1046 with the
 * buffers already in HBM).  d_yo, d_t, d_result, d_err and d_stats are device
 * pointers; pot is a handle from gh_potential_upload.  d_stats (nullable) gets
 * 3 doubles per orbit: orbit-steps (accepted steps for dop853), rejected
 * steps, step size used (symplectic).  stream is a hipStream_t (NULL = the
 * default stream).  The call is asynchronous with respect to the host.
 * Returns 0 on success, nonzero on a launch/argument error. */
int integrateFullOrbit_hip(int nobj, double *d_yo, int nt, const double *d_t, int indiv_t,
                           const void *pot, double dt, double rtol, double atol,
                           double *d_result, int *d_err, int odeint_type, double *d_stats,
                           void *stream);

/* Bytes of the per-GPU scratch array behind integrateFullOrbit_hip_noalloc. */
#define GH_NOALLOC_SCRATCH_BYTES (80u << 20)

/* integrateFullOrbit_hip without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and errors as integrateFullOrbit_hip; the launch
 * temporaries (DOP853 dense-output overflow area, pilot order and sort
 * temporary, work counter) live in one static device array of
 * GH_NOALLOC_SCRATCH_BYTES per GPU, kept until the process ends.  A DOP853
 * call whose temporaries do not fit (roughly 48 MiB of overflow area at the
 * MI355X's 2,048 resident workgroups plus 16 bytes per orbit when there are
 * more orbits than resident lanes: about 2 million orbits) returns -1 with
 * a message and launches nothing.  DOPR54 and IAS15 use only the work
 * counter there and claim orbits in index order (no pilot, no sort), which
 * changes no result.  Calls that overlap on one GPU race on the array.  The interrupt flag is polled only if it was set up before (by a
 * host-buffer call or galpy_hip_set_interrupt). */
int integrateFullOrbit_hip_noalloc(int nobj, double *d_yo, int nt, const double *d_t,
                                   int indiv_t, const void *pot, double dt, double rtol,
                                   double atol, double *d_result, int *d_err, int odeint_type,
                                   double *d_stats, void *stream);

/* Device-resident twin of integratePlanarOrbit (integratePlanarOrbit.c:743
 * with the buffers in HBM): as integrateFullOrbit_hip with d_yo[nobj*4]
 * and d_result[nobj*nt*4]; pot from gh_planar_potential_upload. */
int integratePlanarOrbit_hip(int nobj, double *d_yo, int nt, const double *d_t, int indiv_t,
                             const void *pot, double dt, double rtol, double atol,
                             double *d_result, int *d_err, int odeint_type, double *d_stats,
                             void *stream);

/* Decodes galpy's packed potential (integrateFullOrbit.c:69-1045 semantics)
 * into a device-resident table; *pot receives an opaque host handle (device
 * table + component codes, used to select a kernel specialised for the
 * composition).  gh_potential_free releases both. */
int gh_potential_upload(int npot, const int *pot_type, const double *pot_args, void **pot);
/* The same from galpy's planar packing (integratePlanarOrbit.c:55-739). */
int gh_planar_potential_upload(int npot, const int *pot_type, const double *pot_args,
                               void **pot);
/* gh_potential_upload / gh_planar_potential_upload (planar != 0) for a
 * potential whose components take time functions (pot_tfuncs as for
 * integrateFullOrbit): each is tabulated on the host over [t0, t1], the
 * times the handle will be integrated over. */
int gh_potential_upload_tfuncs(int npot, const int *pot_type, const double *pot_args,
                               double (**pot_tfuncs)(double), double t0, double t1, int planar,
                               void **pot);
/* Host evaluation of the table gh_potential_upload_tfuncs builds for the
 * time function fns[0] over [t0, t1], at ts[n] (for CPU tests; no
 * reference counterpart). */
int gh_host_tfunc_eval(double (**fns)(double), double t0, double t1, int n, const double *ts,
                       double *out);
int gh_potential_free(void *pot);
/* Device address of the table behind a handle. */
const void *gh_potential_device_ptr(const void *pot);

/* Energy (omega = 0) or Jacobi integral E - omega*Lz along stored
 * trajectories (Orbits.py:2902 E, :3298 Jacobi): d_rows[nobj*nt*6]
 * cylindrical rows, d_t[nt] (or [nobj*nt] with indiv_t).  Writes per orbit
 * d_e0[i] = integral at row 0 and d_demax[i] = max_j |E_j - E_0| / |E_0|.
 * Asynchronous on stream. */
int gh_energy_drift_hip(int nobj, int nt, const double *d_rows, const double *d_t, int indiv_t,
                        const void *pot, double omega, double *d_e0, double *d_demax,
                        void *stream);

/* Post-processing of stored trajectories on the device (galpy computes these
 * on the host from the integrated rows: Orbits.py rperi :3733-3778, rap
 * :3686-3731, e :3642-3684, zmax :3966-4010, E :2902-3115, Jacobi
 * :3298-3360), so only per-orbit numbers need to cross PCIe.  d_rows is
 * [nobj][nt][ncol] with ncol 6 (R,vR,vT,z,vz,phi) or 4 (planar R,vR,vT,phi);
 * per orbit: d_rperi = min r, d_rap = max r (r = sqrt(R^2 + z^2) over all
 * outputs), d_zmax = max |z| (0 for planar rows), d_e = (rap - rperi) /
 * (rap + rperi); per output ([nobj][nt]): d_E = Phi + v^2/2 and d_EJ =
 * E - omega R vT.  Every output pointer may be NULL; the energies need pot
 * (an uploaded potential) and d_t ([nt], or [nobj][nt] with indiv_t).
 * Device pointers; asynchronous on stream.  No reference counterpart as a C
 * function. */
int gh_orbit_summary_hip(int nobj, int nt, int ncol, const double *d_rows, const double *d_t,
                         int indiv_t, const void *pot, double omega, double *d_rperi,
                         double *d_rap, double *d_zmax, double *d_e, double *d_E, double *d_EJ,
                         void *stream);
/* The same along planar rows d_rows[nobj*nt*4] (R,vR,vT,phi), z = vz = 0. */
int gh_planar_energy_drift_hip(int nobj, int nt, const double *d_rows, const double *d_t,
                               int indiv_t, const void *pot, double omega, double *d_e0,
                               double *d_demax, void *stream);

/* Potential value and cylindrical forces at n points (device pointers). */
int gh_eval_hip(int n, const double *d_R, const double *d_z, const double *d_phi,
                const double *d_t, const void *pot, double *d_pot, double *d_FR,
                double *d_Fz, double *d_Fphi, void *stream);

/* Host evaluation of the force-kernel incomplete gamma of component `comp`
 * (a PowerSphericalPotentialwCutoff) from the decoded table, exactly as the
 * device computes it; for checks of the table without a GPU. */
int gh_host_bulge_lower_gamma(int npot, const int *pot_type, const double *pot_args, int comp,
                              int n, const double *x, double *out);

/* Host evaluation of the force kernels' logarithm of y >= 1 (the
 * log(1 + r/a) of NFWPotential.c:26), the same code the device runs; for
 * checks without a GPU. */
int gh_host_log_ge1(int n, const double *y, double *out);

/* Host evaluation of the force kernels' sine and cosine of one argument (the
 * pattern angles of DehnenBarPotential.c:18-124, SoftenedNeedleBarPotential.c
 * and the rotating wrappers), the same code the device runs; for checks
 * without a GPU. */
int gh_host_sincos(int n, const double *x, double *s, double *c);

/* ------------------------------------------------------------------------
 * Differentiation (no reference counterpart; galpy offers the forward
 * tangent map integrate_dxdv, bovy_symplecticode.c:1053-1268).
 * ------------------------------------------------------------------------ */
/* N full leapfrog (method 0) or symplec4 (3) steps of size dt from the
 * rectangular states d_w0[n][6] (x,y,z,vx,vy,vz) to d_wT[n][6] -- galpy's
 * steppers on t = t0 + k dt.  Device pointers; asynchronous on stream. */
int gh_flow_hip(int n, const double *d_w0, double *d_wT, double t0, double dt, long long nsteps,
                const void *pot, int method, void *stream);

/* Reverse-time adjoint of gh_flow_hip for time-independent potentials:
 * from the final states d_wT and d_gT = dL/dw(T) ([n][6]) it recomputes the
 * trajectory backwards by reversibility (nothing stored) and writes
 * d_lam0 = dL/dw(0) ([n][6]), d_w0rec = the recomputed initial states
 * (nullable) and d_dtheta[n][8][3] = dL/d(packed argument j of component c)
 * per orbit (nullable; slots: MiyamotoNagai amp,a,b; NFW amp,a; Hernquist
 * amp,a; every other component amp only).  DehnenBar and a triaxial
 * LogarithmicHalo are rejected. */
int gh_adjoint_hip(int n, const double *d_wT, const double *d_gT, double dt, long long nsteps,
                   const void *pot, int method, double *d_lam0, double *d_w0rec, double *d_dtheta,
                   void *stream);

/* ------------------------------------------------------------------------
 * Batched integration (no reference counterpart: galpy integrates one group
 * of orbits per call).  Many groups -- each with its own potential, output
 * times, step and method -- run in one persistent launch.  All pointers in
 * the descriptors are device pointers.
 * ------------------------------------------------------------------------ */
#include <stddef.h>

typedef struct gh_batch_group {
  const void *table;   /* gh_potential_device_ptr(handle) */
  const double *t;     /* [nt] output times */
  double *yo;          /* [n][6] initial conditions (R,vR,vT,z,vz,phi); overwritten as galpy does */
  double *result;      /* [n][nt][6] output rows */
  int *err;            /* [n] */
  double *stats;       /* [n][3] orbit-steps, rejected steps, step used; may be NULL */
  void *state;         /* [n] resumable state records (gh_batch_state_bytes() each) */
  double dt;           /* fixed step, or -9999.99 to estimate per orbit (symplectic) */
  double rtol, atol;   /* natural logarithms */
  int n, nt;
  int method;          /* 0 leapfrog, 3 symplec4, 6 dop853 */
  int pad;
} gh_batch_group;

typedef struct gh_batch_tile {
  int group;           /* index into the group array */
  int first;           /* first orbit of the group in this tile */
  int count;           /* orbits in the tile, <= 64 * orbits-per-lane */
  int pad;
} gh_batch_tile;

/* Bytes of resumable state per orbit. */
size_t gh_batch_state_bytes(void);

/* Prepares tiles for their first launch: marks the state records of their
 * orbits unstarted, clears their tile flags and sets *d_remaining = ntiles
 * (d_remaining may be NULL).  Kernels only, so the sequence may be captured
 * in a graph. */
int gh_batch_reset(const gh_batch_group *d_groups, const gh_batch_tile *d_tiles, int ntiles,
                   int *d_tile_done, int *d_remaining, void *stream);

/* Flags of gh_batch_launch. */
#define GH_BATCH_DOP 1         /* some group uses dop853 */
#define GH_BATCH_SPECIALISED 2 /* every group's potential is MWPotential2014 ([15,5,9]),
                                  LogarithmicHalo+DehnenBar ([0,1]), Hernquist ([8]) or
                                  LogarithmicHalo ([0]): run only the forces specialised
                                  for those (higher occupancy); a group with another
                                  composition then finishes at once with err = -1 */

/* One persistent launch: wavefronts claim unfinished tiles from *d_counter
 * (reset to 0 here by a kernel) and advance each orbit by at most `budget`
 * steps (<= 0: to completion).  opl = orbits per lane (1 or 2; tiles must
 * hold at most 64*opl orbits).  flags: GH_BATCH_DOP must be set if any group
 * uses dop853 (selects the kernel instance that carries it); see
 * GH_BATCH_SPECIALISED.  *d_remaining counts unfinished tiles.  Asynchronous
 * on stream. */
int gh_batch_launch(const gh_batch_group *d_groups, const gh_batch_tile *d_tiles, int ntiles,
                    int *d_tile_done, unsigned long long *d_counter, int *d_remaining,
                    long long budget, int opl, int flags, void *stream);

/* Message of the last error on the calling thread ("" if none). */
const char *galpy_hip_last_error(void);

/* Sets the interrupt flag polled by running kernels (as galpy's SIGINT
 * handler sets `interrupted`, bovy_symplecticode.c:38-45); orbits that see
 * it stop with err = -10.  Pass 0 to clear. */
void galpy_hip_set_interrupt(int value);

#ifdef __cplusplus
}
#endif
#endif /* GALPY_HIP_H */
