"""Shared test helpers: the CPU oracle (oracle/liboracle.so) as the checker,
seeded initial conditions and the benchmark potentials.

Only tests, the smoke check and the CPU-baseline leg of bench.py use the
oracle; it is never part of the product path."""

import ctypes
import os

import numpy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
LN1EM12 = -12.0 * numpy.log(10.0)
DT_SENTINEL = -9999.99

ORACLE_FMA_PATH = os.path.join(ROOT, "oracle", "liboracle_fma.so")

_oracle = None
_oracle_fma = None


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _oracle = ctypes.CDLL(ORACLE_PATH)
    return _oracle


def oracle_fma():
    """The oracle compiled with fused multiply-adds: the same algorithm with
    a different, equally valid rounding.  Its distance from oracle() measures
    how strongly a problem amplifies rounding differences."""
    global _oracle_fma
    if _oracle_fma is None:
        if not os.path.exists(ORACLE_FMA_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _oracle_fma = ctypes.CDLL(ORACLE_FMA_PATH)
    return _oracle_fma


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# packed arguments per component (integrateFullOrbit.py:65-365; the planar
# packing gives PowerSphericalPotentialwCutoff 3) and wrappers' own arguments
_NARGS = {0: 4, 1: 6, 2: 8, 3: 8, 28: 9, 29: 1, 4: 6, 5: 3, 6: 4, 16: 3, 25: 13, 7: 2, 8: 2, 9: 2, 10: 2, 14: 2, 15: 5, 17: 2, 18: 2, 19: 2, 20: 2,
          33: 2, 34: 3, 35: 3, 12: 4, 40: 0, 41: 3, 42: 4, 46: 3, -1: 4, -2: 3, -4: 5, -5: 3, -8: 32, -9: 1, -10: 3, -12: 3, -3: 5, 39: 23}


_ELLIPSOIDAL = (21, 22, 23, 30, 37, 38, 43)


def _nargs(code, pot_args, ai, planar):
    """Arguments of one component or wrapper starting at pot_args[ai]; an
    EllipsoidalPotential has 30 + npsi + 2 glorder (integrateFullOrbit.c:437)."""
    if code in _ELLIPSOIDAL:
        npsi = int(pot_args[ai + 16])
        return 30 + npsi + 2 * int(pot_args[ai + 29 + npsi])
    if code == 27:  # SpiralArmsPotential: 10 + nCs
        return 10 + int(pot_args[ai])
    if code == 11:  # DoubleExponentialDiskPotential: 5 + 4 de_n
        return 5 + 4 * int(pot_args[ai + 4])
    if code == 24:  # SCFPotential
        nonaxi, NLM, Nt = int(pot_args[ai + 1]), int(pot_args[ai + 2] * pot_args[ai + 3] * pot_args[ai + 4]), int(pot_args[ai + 5])
        return 6 + (1 + nonaxi) * NLM + 11 if Nt == 0 else 6 + Nt + (1 + nonaxi) * NLM * 4 * (Nt - 1) + 11
    if code == 13:  # interpRZPotential
        nR, nz = int(pot_args[ai]), int(pot_args[ai + 1])
        return 5 + nR + nz + (3 + int(pot_args[ai + 2] + pot_args[ai + 3] + pot_args[ai + 4])) * nR * nz + 2
    if code == 44:  # MultipoleExpansionPotential
        Nr, L, M = int(pot_args[ai]), int(pot_args[ai + 1]), int(pot_args[ai + 2])
        Nt = int(pot_args[ai + 4 + Nr + 1])
        nblk = sum(2 * min(l + 1, M) - 1 for l in range(L))
        per = 12 * (Nr - 1) + Nr
        return 4 + Nr + 2 + Nt + nblk * (per if Nt == 0 else 4 * (Nt - 1) * per)
    if code == -11:  # FDMDynamicalFrictionForce's own: nPts, r, sigma_r, 16, m/hbar, factor
        return 1 + 2 * int(pot_args[ai]) + 18
    if code == -7:  # ChandrasekharDynamicalFrictionForce's own: nPts, r, sigma_r, 16
        return 1 + 2 * int(pot_args[ai]) + 16
    if code == 45:  # NonInertialFrameForce with splines: n_spline, nPts, tgrid, arrays, 23 + 2
        return 2 + (1 + int(pot_args[ai])) * int(pot_args[ai + 1]) + 25
    if code == 26:  # DiskSCFPotential pair: nsigma + 3
        return int(pot_args[ai]) + 3
    if code == 36:  # interpSphericalPotential: 1 + 2 nr + 6
        return 1 + 2 * int(pot_args[ai]) + 6
    if code == -6:  # MovingObjectPotential's own: 1 + (3|4) n + 3
        return 1 + (3 if planar else 4) * int(pot_args[ai]) + 3
    return 3 if (planar and code in (15, 12)) else _NARGS[code]


def npot_of(pot_type, pot_args, planar=False):
    """galpy's npot for packed arrays: the number of top-level components
    (a wrapper and the components it wraps count once)."""
    ti = ai = 0

    def walk():
        nonlocal ti, ai
        code = int(pot_type[ti])
        ti += 1
        if code < 0:
            nw = int(pot_args[ai])
            ai += 1
            for _ in range(nw):
                walk()
        ai += _nargs(code, pot_args, ai, planar)

    n = 0
    while ti < len(pot_type):
        walk()
        n += 1
    return n


class oracle_tfuncs:
    """Context: the oracle's parses take these time functions (galpy's
    pot_tfuncs, in packing order) until the context ends."""

    def __init__(self, tfuncs):
        from galpy_amd.integrate import tfunc_array

        self.arr = tfunc_array(tfuncs)

    def __enter__(self):
        lib = oracle()
        lib.oracle_set_tfuncs(ctypes.cast(self.arr, ctypes.c_void_p) if self.arr is not None else None,
                              0 if self.arr is None else len(self.arr))
        return self

    def __exit__(self, *exc):
        oracle().oracle_set_tfuncs(None, 0)
        return False


def oracle_integrate(pot_type, pot_args, yo, t, method, dt=None, rtol=LN1EM12, atol=LN1EM12,
                     nthreads=0, planar=False, fma=False):
    """Run the oracle's integrateFullOrbit (returns result, err, stats[N,4],
    yo_after) where stats = (steps, rejected, rhs, dt).  planar=True runs
    its integratePlanarOrbit instead: rows of (R,vR,vT,phi) and the planar
    packing.  fma=True runs the FMA-rounded build (oracle_fma)."""
    lib = oracle_fma() if fma else oracle()
    yo = numpy.array(numpy.atleast_2d(yo), dtype=numpy.float64, order="C")
    t = numpy.ascontiguousarray(t, dtype=numpy.float64)
    n = len(yo)
    nt = t.shape[-1]
    indiv = int(t.ndim > 1)
    ncol = 4 if planar else 6
    assert yo.shape[1] == ncol
    res = numpy.empty((n, nt, ncol))
    err = numpy.zeros(n, dtype=numpy.int32)
    steps = numpy.zeros(n, dtype=numpy.int64)
    rej = numpy.zeros(n, dtype=numpy.int64)
    rhs = numpy.zeros(n, dtype=numpy.int64)
    dtu = numpy.zeros(n)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    fn = lib.oracle_integrate_planar_stats if planar else lib.oracle_integrate_stats
    fn(
        ctypes.c_int(n), _p(yo), ctypes.c_int(nt), _p(t), ctypes.c_int(indiv), ctypes.c_int(npot_of(pt, pa, planar)),
        _p(pt), _p(pa), ctypes.c_double(DT_SENTINEL if dt is None else dt), ctypes.c_double(rtol),
        ctypes.c_double(atol), _p(res), _p(err), ctypes.c_int(method), _p(steps), _p(rej), _p(rhs),
        _p(dtu), ctypes.c_int(nthreads),
    )
    stats = numpy.stack([steps, rej, rhs, dtu], axis=1).astype(float)
    return res, err, stats, yo


def oracle_forces(pot_type, pot_args, R, z, phi=None, t=None):
    lib = oracle()
    R = numpy.ascontiguousarray(R, dtype=float)
    n = R.size
    z = numpy.ascontiguousarray(numpy.broadcast_to(z, R.shape), dtype=float)
    phi = numpy.ascontiguousarray(numpy.zeros(n) if phi is None else numpy.broadcast_to(phi, R.shape), dtype=float)
    t = numpy.ascontiguousarray(numpy.zeros(n) if t is None else numpy.broadcast_to(t, R.shape), dtype=float)
    out = [numpy.empty(n) for _ in range(3)]
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_forces(npot_of(pt, pa), _p(pt), _p(pa), n, _p(R), _p(z), _p(phi), _p(t), *[_p(o) for o in out])
    assert rc == 0
    return tuple(out)


def oracle_density(pot_type, pot_args, R, z, phi=None):
    lib = oracle()
    R = numpy.ascontiguousarray(R, dtype=float)
    z = numpy.ascontiguousarray(numpy.broadcast_to(z, R.shape), dtype=float)
    phi = numpy.ascontiguousarray(numpy.zeros(R.size) if phi is None else numpy.broadcast_to(phi, R.shape), dtype=float)
    out = numpy.empty(R.size)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    assert lib.oracle_eval_density(npot_of(pt, pa), _p(pt), _p(pa), R.size, _p(R), _p(z), _p(phi), _p(out)) == 0
    return out


def oracle_diss_forces(pot_type, pot_args, R, z, phi, t, vR, vT, vz):
    lib = oracle()
    arrs = [numpy.ascontiguousarray(numpy.broadcast_to(v, numpy.shape(R)), dtype=float) for v in (R, z, phi, t, vR, vT, vz)]
    out = [numpy.empty(arrs[0].size) for _ in range(3)]
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    assert lib.oracle_eval_diss_forces(npot_of(pt, pa), _p(pt), _p(pa), arrs[0].size, *[_p(a) for a in arrs],
                                       *[_p(o) for o in out]) == 0
    return tuple(out)


def oracle_potential(pot_type, pot_args, R, z, phi=None, t=None):
    lib = oracle()
    R = numpy.ascontiguousarray(R, dtype=float)
    n = R.size
    z = numpy.ascontiguousarray(numpy.broadcast_to(z, R.shape), dtype=float)
    phi = numpy.ascontiguousarray(numpy.zeros(n) if phi is None else numpy.broadcast_to(phi, R.shape), dtype=float)
    t = numpy.ascontiguousarray(numpy.zeros(n) if t is None else numpy.broadcast_to(t, R.shape), dtype=float)
    out = numpy.empty(n)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_potential(npot_of(pt, pa), _p(pt), _p(pa), n, _p(R), _p(z), _p(phi), _p(t), _p(out))
    assert rc == 0
    return out


def oracle_count_ops(pot_type, pot_args, yo, t, method, dt=None, planar=False):
    lib = oracle()
    yo = numpy.array(numpy.atleast_2d(yo), dtype=numpy.float64, order="C")
    t = numpy.ascontiguousarray(t, dtype=numpy.float64)
    n, nt = len(yo), t.shape[-1]
    res = numpy.empty((n, nt, 4 if planar else 6))
    err = numpy.zeros(n, dtype=numpy.int32)
    out = numpy.zeros(8)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    fn = lib.oracle_count_ops_planar if planar else lib.oracle_count_ops
    rc = fn(
        ctypes.c_int(n), _p(yo), ctypes.c_int(nt), _p(t), ctypes.c_int(int(t.ndim > 1)), ctypes.c_int(npot_of(pt, pa, planar)),
        _p(pt), _p(pa), ctypes.c_double(DT_SENTINEL if dt is None else dt), ctypes.c_double(LN1EM12),
        ctypes.c_double(LN1EM12), _p(res), _p(err), ctypes.c_int(method), _p(out),
    )
    assert rc == 0
    keys = ["add", "mul", "div", "sqrt", "special", "steps", "rejected", "rhs"]
    return dict(zip(keys, out))


def ics(n, seed=0):
    """Seeded initial conditions of the benchmark configurations:
    R~U(0.5,2), z~N(0,0.1), phi~U(0,2pi), vR,vz~N(0,0.1), vT~N(1,0.1);
    rows are (R,vR,vT,z,vz,phi)."""
    rng = numpy.random.default_rng(seed)
    R = rng.uniform(0.5, 2.0, n)
    z = rng.normal(0.0, 0.1, n)
    phi = rng.uniform(0.0, 2.0 * numpy.pi, n)
    vR = rng.normal(0.0, 0.1, n)
    vz = rng.normal(0.0, 0.1, n)
    vT = rng.normal(1.0, 0.1, n)
    return numpy.stack([R, vR, vT, z, vz, phi], axis=1)


def oracle_integrate_dxdv(pot_type, pot_args, w, t, method, dim=3, dt=None, rtol=LN1EM12,
                          atol=LN1EM12, nthreads=0, steps=None):
    """The oracle's integrate_dxdv for many orbits: w [n, 4 dim] rectangular
    (q, p | dq, dp) -> (result [n, nt, 4 dim], err); ``steps`` (an int64
    array of n, optional) receives the orbit-steps."""
    lib = oracle()
    w = numpy.array(numpy.atleast_2d(w), dtype=numpy.float64, order="C")
    t = numpy.ascontiguousarray(t, dtype=numpy.float64)
    n, nt = len(w), t.shape[-1]
    assert w.shape[1] == 4 * dim
    res = numpy.empty((n, nt, 4 * dim))
    err = numpy.zeros(n, dtype=numpy.int32)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    lib.oracle_integrate_dxdv(
        ctypes.c_int(n), ctypes.c_int(dim), _p(w), ctypes.c_int(nt), _p(t), ctypes.c_int(int(t.ndim > 1)),
        ctypes.c_int(npot_of(pt, pa, dim == 2)), _p(pt), _p(pa), ctypes.c_double(DT_SENTINEL if dt is None else dt),
        ctypes.c_double(rtol), ctypes.c_double(atol), _p(res), _p(err), ctypes.c_int(method),
        ctypes.c_int(nthreads), None if steps is None else _p(steps),
    )
    return res, err


def oracle_count_ops_dxdv(pot_type, pot_args, w, t, method, dim=3, dt=None):
    """Counted FP64 ops of the oracle's integrate_dxdv (see oracle_count_ops)."""
    lib = oracle()
    w = numpy.array(numpy.atleast_2d(w), dtype=numpy.float64, order="C")
    t = numpy.ascontiguousarray(t, dtype=numpy.float64)
    n, nt = len(w), len(t)
    res = numpy.empty((n, nt, 4 * dim))
    err = numpy.zeros(n, dtype=numpy.int32)
    out = numpy.zeros(6)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_count_ops_dxdv(
        ctypes.c_int(n), ctypes.c_int(dim), _p(w), ctypes.c_int(nt), _p(t), ctypes.c_int(npot_of(pt, pa, dim == 2)), _p(pt),
        _p(pa), ctypes.c_double(DT_SENTINEL if dt is None else dt), ctypes.c_double(LN1EM12),
        ctypes.c_double(LN1EM12), _p(res), _p(err), ctypes.c_int(method), _p(out),
    )
    assert rc == 0
    return dict(zip(["add", "mul", "div", "sqrt", "special", "steps"], out))


def planar_ics(n, seed=0):
    """Planar rows (R,vR,vT,phi) of the same law as ics (the z, vz draws are
    dropped)."""
    return numpy.ascontiguousarray(ics(n, seed)[:, [0, 1, 2, 5]])


def in_plane(planar_rows):
    """Full rows (R,vR,vT,0,0,phi) of planar rows (R,vR,vT,phi)."""
    p = numpy.asarray(planar_rows)
    z = numpy.zeros(p.shape[:-1] + (1,))
    return numpy.concatenate([p[..., :3], z, z, p[..., 3:]], axis=-1)


def oracle_planar_forces(pot_type, pot_args, R, phi=None, t=None):
    """The oracle's calcPlanarRforce / calcPlanarphitorque (planar packing)."""
    lib = oracle()
    R = numpy.ascontiguousarray(R, dtype=float)
    n = R.size
    phi = numpy.ascontiguousarray(numpy.zeros(n) if phi is None else numpy.broadcast_to(phi, R.shape), dtype=float)
    t = numpy.ascontiguousarray(numpy.zeros(n) if t is None else numpy.broadcast_to(t, R.shape), dtype=float)
    FR, Fphi = numpy.empty(n), numpy.empty(n)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_planar_forces(npot_of(pt, pa, True), _p(pt), _p(pa), n, _p(R), _p(phi), _p(t), _p(FR), _p(Fphi))
    assert rc == 0
    return FR, Fphi


def bar_potential():
    """LogarithmicHaloPotential(q=0.9) + DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01)."""
    from galpy_amd.potential import DehnenBarPotential, LogarithmicHaloPotential

    return [LogarithmicHaloPotential(q=0.9), DehnenBarPotential(omegab=1.85, rb=0.8, Af=0.01)]


def spherical_a():
    """Plummer + Kuzmin disk + Burkert + Jaffe, normalized so that the
    circular velocity at R = 1 is 1."""
    from galpy_amd.potential import BurkertPotential, JaffePotential, KuzminDiskPotential, PlummerPotential

    return [PlummerPotential(b=0.8, normalize=0.3), KuzminDiskPotential(a=0.5, normalize=0.3),
            BurkertPotential(a=2.0, normalize=0.2), JaffePotential(a=2.0, normalize=0.2)]


def spherical_b():
    """Isochrone + pseudo-isothermal + cored Dehnen + Dehnen (alpha 1.5) +
    homogeneous sphere + power law + Kepler (seven components)."""
    from galpy_amd.potential import (
        DehnenCoreSphericalPotential,
        DehnenSphericalPotential,
        HomogeneousSpherePotential,
        IsochronePotential,
        KeplerPotential,
        PowerSphericalPotential,
        PseudoIsothermalPotential,
    )

    return [IsochronePotential(b=0.7, normalize=0.2), PseudoIsothermalPotential(a=1.3, normalize=0.15),
            DehnenCoreSphericalPotential(a=1.2, normalize=0.15),
            DehnenSphericalPotential(a=0.9, alpha=1.5, normalize=0.15),
            HomogeneousSpherePotential(R=1.5, normalize=0.1), PowerSphericalPotential(alpha=1.7, normalize=0.15),
            KeplerPotential(normalize=0.1)]


def orbit_diffs(a, b, phi_last=True):
    """Per-orbit max over times and columns of |a - b| (relative, unit
    floor); with phi_last the last column is an angle, compared mod 2 pi.
    a, b: [N, nt, ncol]."""
    d = numpy.abs(a - b) / numpy.maximum(numpy.abs(b), 1.0)
    if phi_last:
        d[..., -1] = numpy.abs((a[..., -1] - b[..., -1] + numpy.pi) % (2 * numpy.pi) - numpy.pi)
    return d.reshape(d.shape[0], -1).max(axis=1)


def ulp_sensitivity(run, y0, phi_last=True, nrep=3):
    """Conditioning of each orbit problem: the per-orbit change of the
    oracle's own result when every initial coordinate moves by one ulp
    (max over nrep random sign patterns).  run(y0) returns the oracle's
    [N, nt, ncol] rows for initial states y0."""
    ref = run(y0)
    worst = numpy.zeros(len(y0))
    for k in range(nrep):
        sign = numpy.random.default_rng(1000 + k).choice([-1.0, 1.0], size=numpy.shape(y0))
        worst = numpy.maximum(worst, orbit_diffs(run(y0 * (1.0 + sign * 2.0**-52)), ref, phi_last))
    return ref, worst


def assert_parity(res, ref, bar, sens=None, factor=0.0, phi_last=True, label=""):
    """Every orbit within `bar`, or, for an orbit whose problem is so
    ill-conditioned that the oracle's own result moves by sens under a
    one-ulp change of its inputs, within factor * sens (rounding
    differences are then amplified beyond the bar by the problem itself)."""
    d = orbit_diffs(res, ref, phi_last)
    allowed = numpy.full(len(d), float(bar)) if sens is None else numpy.maximum(bar, factor * sens)
    bad = numpy.flatnonzero(d > allowed)
    assert bad.size == 0, (label, bad[:5], d[bad[:5]], allowed[bad[:5]])
    return d


def phase_diff(a, b):
    """Max over rows of |a-b| per component with unit floor, phi modulo 2 pi.
    a, b: [..., 6] rows (R,vR,vT,z,vz,phi) or [..., 4] planar rows
    (R,vR,vT,phi); phi is the last column of both."""
    d = numpy.abs(a - b)
    dphi = numpy.abs((a[..., -1] - b[..., -1] + numpy.pi) % (2 * numpy.pi) - numpy.pi)
    d[..., -1] = dphi
    scale = numpy.maximum(numpy.abs(b), 1.0)
    scale[..., -1] = numpy.pi
    return (d / scale).max()


def oracle_dxdv(pot_type, pot_args, w, nde, t, method, dt=None):
    """Oracle tangent-map integration: w [n, 6*(1+nde)] rectangular base
    state + deviation columns -> final augmented states."""
    lib = oracle()
    w = numpy.ascontiguousarray(numpy.atleast_2d(w), dtype=numpy.float64)
    t = numpy.ascontiguousarray(t, dtype=numpy.float64)
    out = numpy.empty_like(w)
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_symplectic_dxdv(
        ctypes.c_int(len(w)), _p(w), ctypes.c_int(nde), ctypes.c_int(len(t)), _p(t), ctypes.c_int(npot_of(pt, pa)),
        _p(pt), _p(pa), ctypes.c_double(DT_SENTINEL if dt is None else dt), ctypes.c_double(LN1EM12),
        ctypes.c_double(LN1EM12), ctypes.c_int(method), _p(out),
    )
    assert rc == 0
    return out


def oracle_stm(pot_type, pot_args, w0, t, method, dt=None):
    """Full 6x6 state-transition matrices M[n, 6, 6] (M[:, i, j] = dw_i(T)/dw_j(0))
    and the final base states, from the oracle's tangent map with the six
    unit deviation columns."""
    w0 = numpy.atleast_2d(w0)
    n = len(w0)
    aug = numpy.zeros((n, 6 * 7))
    aug[:, :6] = w0
    for j in range(6):
        aug[:, 6 * (j + 1) + j] = 1.0
    out = oracle_dxdv(pot_type, pot_args, aug, 6, t, method, dt=dt)
    M = out[:, 6:].reshape(n, 6, 6).transpose(0, 2, 1)
    return M, out[:, :6]


def oracle_hessian(pot_type, pot_args, R, z):
    lib = oracle()
    R = numpy.ascontiguousarray(R, dtype=float)
    z = numpy.ascontiguousarray(numpy.broadcast_to(z, R.shape), dtype=float)
    out = [numpy.empty(R.size) for _ in range(3)]
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_hessian(npot_of(pt, pa), _p(pt), _p(pa), R.size, _p(R), _p(z), *[_p(o) for o in out])
    assert rc == 0
    return tuple(out)


def cyl_to_rect(v):
    v = numpy.atleast_2d(v)
    R, vR, vT, z, vz, phi = v.T
    c, s = numpy.cos(phi), numpy.sin(phi)
    return numpy.stack([R * c, R * s, z, vR * c - vT * s, vR * s + vT * c, vz], axis=1)


def oracle_hessian_full(pot_type, pot_args, R, z, phi, t):
    """The oracle's six cylindrical second derivatives (R2, z2, Rz, phi2,
    Rphi, zphi) at (R, z, phi, t); full packing."""
    lib = oracle()
    arrs = [numpy.ascontiguousarray(numpy.broadcast_to(v, numpy.shape(R)), dtype=float) for v in (R, z, phi, t)]
    n = arrs[0].size
    out = [numpy.empty(n) for _ in range(6)]
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_hessian_full(npot_of(pt, pa), _p(pt), _p(pa), n, *[_p(a) for a in arrs],
                                      *[_p(o) for o in out])
    assert rc == 0
    return tuple(out)


def oracle_planar_hessian(pot_type, pot_args, R, phi, t):
    """The oracle's planar second derivatives (R2, phi2, Rphi); planar packing."""
    lib = oracle()
    arrs = [numpy.ascontiguousarray(numpy.broadcast_to(v, numpy.shape(R)), dtype=float) for v in (R, phi, t)]
    n = arrs[0].size
    out = [numpy.empty(n) for _ in range(3)]
    pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
    pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
    rc = lib.oracle_eval_planar_hessian(npot_of(pt, pa, True), _p(pt), _p(pa), n, *[_p(a) for a in arrs],
                                        *[_p(o) for o in out])
    assert rc == 0
    return tuple(out)
