"""Host front end of the HIP integrators (galpy's integrateFullOrbit_c twin).

``integrateFullOrbit_hip`` keeps the argument meaning and return values of
galpy's ``integrateFullOrbit_c`` (galpy/orbit/integrateFullOrbit.py:616-749):
packing through ``parse_pot``, tolerances as natural logs
(integratePlanarOrbit.py:712-724), ``dt=None`` -> -9999.99, a preallocated
``(nobj, nt, 6)`` result, a per-orbit ``err``, and ``KeyboardInterrupt`` when
an orbit reports -10.  The call goes through ctypes to the C-ABI symbol
``integrateFullOrbit`` of libgalpy_hip.so.  ``integratePlanarOrbit_hip`` is
the same for planar orbits (galpy's ``integratePlanarOrbit_c``,
galpy/orbit/integratePlanarOrbit.py:989-1122): rows of (R,vR,vT,phi), the
planar packing ``parse_pot_planar`` and the symbol ``integratePlanarOrbit``.

``integrate_device`` is the HBM-resident path: inputs and outputs are torch
tensors on the GPU and the kernels run asynchronously on torch's current
stream (used by the benchmark, batching and serving layers).
"""

import ctypes

import numpy

from . import _lib
from .potential import parse_pot, parse_pot_planar

DT_SENTINEL = -9999.99

METHOD_CODES = {
    "leapfrog_hip": 0,
    "symplec4_hip": 3,
    "symplec6_hip": 4,
    "rk4_hip": 1,
    "rk6_hip": 2,
    "dopr54_hip": 5,
    "dop853_hip": 6,
    "ias15_hip": 7,
}


def _parse_integrator(int_method):
    """integratePlanarOrbit.py:690-709 for the HIP method names (unknown -> 0)."""
    return METHOD_CODES.get(int_method.lower(), 0)


def _parse_tol(rtol, atol):
    """integratePlanarOrbit.py:712-724: natural logs, default 1e-12."""
    rtol = -12.0 * numpy.log(10.0) if rtol is None else numpy.log(rtol)
    atol = -12.0 * numpy.log(10.0) if atol is None else numpy.log(atol)
    return (rtol, atol)


def integrateFullOrbit_hip(pot, yo, t, int_method, rtol=None, atol=None, progressbar=False, dt=None):
    """Integrate N full 3D orbits on the GPU.

    Parameters mirror galpy's ``integrateFullOrbit_c``: ``yo`` is ``[N,6]`` or
    ``[6]`` in (R,vR,vT,z,vz,phi); ``t`` is 1D (shared) or ``[N,nt]``.
    Returns ``(result, err)`` with ``result`` of shape ``(N,nt,6)`` (or
    ``(nt,6)`` for a single orbit).
    """
    return _integrate_host("integrateFullOrbit", parse_pot, 6, pot, yo, t, int_method, rtol, atol,
                           progressbar, dt)


def integratePlanarOrbit_hip(pot, yo, t, int_method, rtol=None, atol=None, progressbar=False,
                             dt=None):
    """Integrate N planar orbits on the GPU (galpy's ``integratePlanarOrbit_c``,
    integratePlanarOrbit.py:989-1122): ``yo`` is ``[N,4]`` or ``[4]`` in
    (R,vR,vT,phi); returns ``(result, err)`` with ``result`` of shape
    ``(N,nt,4)`` (or ``(nt,4)``)."""
    return _integrate_host("integratePlanarOrbit", parse_pot_planar, 4, pot, yo, t, int_method,
                           rtol, atol, progressbar, dt)


TFUNC_TYPE = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double)


def tfunc_array(tfuncs):
    """galpy's ``_prep_tfuncs`` (integrateFullOrbit.py:970-999 without numba):
    the time functions as a C array of ``double f(double)`` pointers, or None
    when there are none.  The returned object keeps the callbacks alive."""
    if not tfuncs:
        return None
    cbs = [TFUNC_TYPE(lambda t, f=f: float(f(t))) for f in tfuncs]
    arr = (TFUNC_TYPE * len(cbs))(*cbs)
    arr._keep = cbs
    return arr


def _integrate_host(symbol, parse, ncol, pot, yo, t, int_method, rtol, atol, progressbar, dt):
    lib = _lib.load()
    single_obj = len(numpy.shape(yo)) == 1
    yo = numpy.require(numpy.atleast_2d(yo), dtype=numpy.float64, requirements=["C", "W"])
    if yo.shape[-1] != ncol:
        raise ValueError(f"{symbol} takes initial conditions of {ncol} phase-space coordinates")
    nobj = len(yo)
    rtol, atol = _parse_tol(rtol, atol)
    npot, pot_type, pot_args, tfuncs = parse(pot, t)
    tf = tfunc_array(tfuncs)
    int_method_c = _parse_integrator(int_method)
    if dt is None:
        dt = DT_SENTINEL
    t = numpy.require(t, dtype=numpy.float64, requirements=["C", "W"])
    indiv_t = t.ndim > 1
    nt = t.shape[-1]
    result = numpy.empty((nobj, nt, ncol))
    err = numpy.zeros(nobj, dtype=numpy.int32)
    cb = None
    if progressbar and nobj > 1:
        try:
            import tqdm

            pbar = tqdm.tqdm(total=nobj, leave=False)
            cb = ctypes.CFUNCTYPE(None)(pbar.update)
        except ImportError:  # pragma: no cover
            cb = None
    getattr(lib, symbol)(
        ctypes.c_int(nobj), yo, ctypes.c_int(nt), t, ctypes.c_int(int(indiv_t)),
        ctypes.c_int(npot), pot_type, pot_args, ctypes.cast(tf, ctypes.c_void_p) if tf is not None else None,
        ctypes.c_double(dt),
        ctypes.c_double(rtol), ctypes.c_double(atol), result, err,
        ctypes.c_int(int_method_c), ctypes.cast(cb, ctypes.c_void_p) if cb is not None else None,
    )
    if cb is not None:
        pbar.close()
    if numpy.any(err == -10):
        raise KeyboardInterrupt("Orbit integration interrupted by CTRL-C (SIGINT)")
    if numpy.any(err == -1):
        raise RuntimeError(f"HIP orbit integration failed: {_lib.last_error()}")
    if single_obj:
        return (result[0], err[0])
    return (result, err)


class PotentialTable:
    """A potential decoded into a device-resident table (gh_potential_upload,
    or gh_planar_potential_upload from the planar packing when ``planar``;
    a potential with time functions needs ``trange``, the time range they
    are tabulated over)."""

    def __init__(self, pot, planar=False, trange=None):
        self.pot = pot
        self.planar = bool(planar)
        npot, pot_type, pot_args, tfuncs = (parse_pot_planar if planar else parse_pot)(
            pot, None if trange is None else numpy.asarray(trange, dtype=float))
        self.npot, self.pot_type, self.pot_args = npot, pot_type, pot_args
        lib = _lib.load()
        handle = ctypes.c_void_p()
        if tfuncs:
            # time functions are tabulated over the times the table serves
            if trange is None:
                raise ValueError("this potential has time functions: give trange=(t0, t1), the times it "
                                 "will be integrated over")
            self._tf = tfunc_array(tfuncs)
            upload = "gh_potential_upload_tfuncs"
            rc = lib.gh_potential_upload_tfuncs(npot, pot_type.ctypes.data, pot_args.ctypes.data,
                                                ctypes.cast(self._tf, ctypes.c_void_p), float(trange[0]),
                                                float(trange[1]), int(self.planar), ctypes.byref(handle))
        else:
            upload = "gh_planar_potential_upload" if planar else "gh_potential_upload"
            rc = getattr(lib, upload)(npot, pot_type.ctypes.data, pot_args.ctypes.data, ctypes.byref(handle))
        _lib.check(rc, upload)
        self.handle = handle

    @property
    def ptr(self):
        return self.handle.value

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                _lib.load().gh_potential_free(self.handle)
                self.handle = None
        except Exception:  # interpreter shutdown
            pass


def _ptr(x):
    return 0 if x is None else x.data_ptr()


def integrate_device(table, yo, t, int_method, rtol=None, atol=None, dt=None, result=None,
                     err=None, stats=None, stream=None):
    """Launch an integration on device tensors (asynchronous on ``stream``).

    ``yo``: float64 ``[N,6]`` cuda tensor (overwritten with the rectangular
    state, as galpy does), ``t``: ``[nt]`` or ``[N,nt]``; ``result`` ``[N,nt,6]``,
    ``err`` int32 ``[N]`` and ``stats`` float64 ``[N,3]`` are allocated when
    not given.  Returns ``(result, err, stats)``.  A ``yo`` of ``[N,4]``
    rows (R,vR,vT,phi) integrates planar orbits (integratePlanarOrbit_hip,
    ``result`` ``[N,nt,4]``); ``table`` must then come from the planar packing.
    """
    import torch

    assert yo.is_cuda and yo.dtype == torch.float64 and yo.is_contiguous()
    ncol = yo.shape[-1]
    assert yo.dim() == 2 and ncol in (4, 6)
    planar = ncol == 4
    if planar != getattr(table, "planar", False):
        raise ValueError("planar orbits need a PotentialTable(pot, planar=True), full orbits a full one")
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    nobj = yo.shape[0]
    indiv_t = t.dim() > 1
    nt = t.shape[-1]
    if indiv_t:
        assert t.shape == (nobj, nt)
    if result is None:
        result = torch.empty((nobj, nt, ncol), dtype=torch.float64, device=yo.device)
    if err is None:
        err = torch.zeros(nobj, dtype=torch.int32, device=yo.device)
    assert result.shape == (nobj, nt, ncol) and result.is_contiguous()
    assert err.shape == (nobj,) and err.dtype == torch.int32
    if stats is not None:
        assert stats.shape == (nobj, 3) and stats.dtype == torch.float64
    rtol, atol = _parse_tol(rtol, atol)
    if stream is None:
        stream = torch.cuda.current_stream(yo.device)
    code = METHOD_CODES[int_method] if isinstance(int_method, str) else int(int_method)
    lib = _lib.load()
    symbol = "integratePlanarOrbit_hip" if planar else "integrateFullOrbit_hip"
    rc = getattr(lib, symbol)(
        nobj, yo.data_ptr(), nt, t.data_ptr(), int(indiv_t), table.ptr,
        DT_SENTINEL if dt is None else float(dt), rtol, atol, result.data_ptr(), err.data_ptr(),
        code, _ptr(stats), stream.cuda_stream,
    )
    _lib.check(rc, symbol)
    return result, err, stats


def energy_drift_device(table, rows, t, omega=0.0, stream=None):
    """Per-orbit E(0) and max |E(t)-E(0)|/|E(0)| (Jacobi with ``omega``) on device,
    along full ``[N,nt,6]`` or planar ``[N,nt,4]`` rows."""
    import torch

    nobj, nt, ncol = rows.shape
    assert ncol in (4, 6) and rows.is_contiguous()
    e0 = torch.empty(nobj, dtype=torch.float64, device=rows.device)
    demax = torch.empty(nobj, dtype=torch.float64, device=rows.device)
    if stream is None:
        stream = torch.cuda.current_stream(rows.device)
    fn = "gh_planar_energy_drift_hip" if ncol == 4 else "gh_energy_drift_hip"
    rc = getattr(_lib.load(), fn)(
        nobj, nt, rows.data_ptr(), t.data_ptr(), int(t.dim() > 1), table.ptr, float(omega),
        e0.data_ptr(), demax.data_ptr(), stream.cuda_stream,
    )
    _lib.check(rc, fn)
    return e0, demax


def orbit_summary_device(rows, t=None, table=None, omega=0.0, energies=False, jacobi=False,
                         stream=None):
    """Per-orbit rperi, rap, zmax, e of stored rows on the device (galpy's
    ``Orbit.rperi/rap/zmax/e`` with ``analytic=False``, Orbits.py:3642-4010),
    and optionally E and the Jacobi integral E - omega Lz at every output
    (``Orbit.E``, ``Orbit.Jacobi``).  ``rows``: float64 cuda tensor
    ``[N, nt, 6]`` (R,vR,vT,z,vz,phi) or planar ``[N, nt, 4]``; the energies
    need ``t`` ([nt] or [N, nt]) and ``table`` (a PotentialTable).  Returns a
    dict of device tensors; asynchronous on ``stream``."""
    import torch

    nobj, nt, ncol = rows.shape
    if ncol not in (4, 6) or not rows.is_contiguous() or rows.dtype != torch.float64:
        raise ValueError("rows must be a contiguous float64 [N, nt, 6] or [N, nt, 4] tensor")
    if (energies or jacobi) and (table is None or t is None):
        raise ValueError("E and the Jacobi integral need the potential table and the times")
    dev = rows.device
    out = {k: torch.empty(nobj, dtype=torch.float64, device=dev) for k in ("rperi", "rap", "e")}
    if ncol == 6:
        out["zmax"] = torch.empty(nobj, dtype=torch.float64, device=dev)
    if energies:
        out["E"] = torch.empty((nobj, nt), dtype=torch.float64, device=dev)
    if jacobi:
        out["Jacobi"] = torch.empty((nobj, nt), dtype=torch.float64, device=dev)
    if stream is None:
        stream = torch.cuda.current_stream(dev)
    p = lambda k: out[k].data_ptr() if k in out else None
    rc = _lib.load().gh_orbit_summary_hip(
        nobj, nt, ncol, rows.data_ptr(), None if t is None else t.data_ptr(),
        int(t is not None and t.dim() > 1), None if table is None else table.ptr, float(omega),
        p("rperi"), p("rap"), p("zmax"), p("e"), p("E"), p("Jacobi"), stream.cuda_stream,
    )
    _lib.check(rc, "gh_orbit_summary_hip")
    return out


def eval_device(table, R, z, phi, t, stream=None):
    """Potential value and cylindrical forces at points (device tensors)."""
    import torch

    n = R.numel()
    outs = [torch.empty(n, dtype=torch.float64, device=R.device) for _ in range(4)]
    if stream is None:
        stream = torch.cuda.current_stream(R.device)
    rc = _lib.load().gh_eval_hip(
        n, R.data_ptr(), z.data_ptr(), phi.data_ptr(), t.data_ptr(), table.ptr,
        *[o.data_ptr() for o in outs], stream.cuda_stream,
    )
    _lib.check(rc, "gh_eval_hip")
    return tuple(outs)


# ---------------------------------------------------------------------------
# integrate_dxdv: the orbit plus one phase-space deviation (rectangular)
# ---------------------------------------------------------------------------
def _dxdv_host(dim, pot, yo, dyo, t, int_method, dt, rtol, atol):
    """One call of gh_integrate_dxdv (or the single-orbit drop-in symbol) for
    rectangular states yo [N, 2 dim] and deviations dyo [N, 2 dim]."""
    lib = _lib.load()
    single = numpy.ndim(yo) == 1
    yo = numpy.atleast_2d(numpy.asarray(yo, dtype=numpy.float64))
    dyo = numpy.atleast_2d(numpy.asarray(dyo, dtype=numpy.float64))
    ncol = 2 * dim
    if yo.shape[-1] != ncol or dyo.shape != yo.shape:
        raise ValueError(f"integrate_dxdv takes states and deviations of {ncol} rectangular coordinates")
    nobj = len(yo)
    w = numpy.require(numpy.hstack((yo, dyo)), dtype=numpy.float64, requirements=["C", "W"])
    rtol, atol = _parse_tol(rtol, atol)
    npot, pot_type, pot_args, _ = (parse_pot if dim == 3 else parse_pot_planar)(pot)
    code = _parse_integrator(int_method)
    t = numpy.require(t, dtype=numpy.float64, requirements=["C", "W"])
    nt = t.shape[-1]
    result = numpy.empty((nobj, nt, 2 * ncol))
    err = numpy.zeros(nobj, dtype=numpy.int32)
    dtv = DT_SENTINEL if dt is None else float(dt)
    if single and t.ndim == 1:
        # the drop-in symbols: galpy's one-orbit signature
        if dim == 3:
            lib.integrateFullOrbit_dxdv(w[0], nt, t, npot, pot_type, pot_args, None, dtv, rtol, atol,
                                        result[0], err, code)
        else:
            lib.integratePlanarOrbit_dxdv(w[0], nt, t, npot, pot_type, pot_args, None, dtv, rtol,
                                          atol, result[0], err, code, None)
    else:
        lib.gh_integrate_dxdv(nobj, dim, w, nt, t, int(t.ndim > 1), npot, pot_type, pot_args, dtv,
                              rtol, atol, result, err, code)
    if numpy.any(err == -10):
        raise KeyboardInterrupt("Orbit integration interrupted by CTRL-C (SIGINT)")
    if numpy.any(err == -1):
        raise RuntimeError(f"HIP orbit integration failed: {_lib.last_error()}")
    if single:
        return (result[0], err[0])
    return (result, err)


def integrateFullOrbit_dxdv_hip(pot, yo, dyo, t, int_method, dt=None, rtol=None, atol=None):
    """galpy's ``integrateFullOrbit_dxdv_c`` (integrateFullOrbit.py:752-858):
    the orbit ``yo`` = (x,y,z,vx,vy,vz) and the deviation ``dyo`` =
    (dx,dy,dz,dvx,dvy,dvz), rectangular, integrated together; returns
    ``(result, err)`` with rows of 12.  ``yo``/``dyo`` may also be ``[N,6]``
    (one call for all orbits; galpy maps the one-orbit call)."""
    return _dxdv_host(3, pot, yo, dyo, t, int_method, dt, rtol, atol)


def integratePlanarOrbit_dxdv_hip(pot, yo, dyo, t, int_method, dt=None, rtol=None, atol=None):
    """galpy's ``integratePlanarOrbit_dxdv_c`` (integratePlanarOrbit.py:1124-1225):
    (x,y,vx,vy) and (dx,dy,dvx,dvy), rows of 8."""
    return _dxdv_host(2, pot, yo, dyo, t, int_method, dt, rtol, atol)


def integrate_dxdv_device(table, w, t, int_method, rtol=None, atol=None, dt=None, result=None,
                          err=None, stats=None, stream=None):
    """gh_integrate_dxdv_hip on device tensors: ``w`` is ``[N,12]`` (full
    orbits, ``table`` from the full packing) or ``[N,8]`` (planar, a planar
    table), rectangular (q, p | dq, dp); ``result`` ``[N,nt,12 or 8]``.
    Asynchronous on ``stream``; returns ``(result, err, stats)``."""
    import torch

    assert w.is_cuda and w.dtype == torch.float64 and w.is_contiguous() and w.dim() == 2
    ncol = w.shape[-1]
    assert ncol in (8, 12)
    dim = 3 if ncol == 12 else 2
    if (dim == 2) != getattr(table, "planar", False):
        raise ValueError("planar deviations need a PotentialTable(pot, planar=True), full ones a full one")
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    nobj, nt = w.shape[0], t.shape[-1]
    if result is None:
        result = torch.empty((nobj, nt, ncol), dtype=torch.float64, device=w.device)
    if err is None:
        err = torch.zeros(nobj, dtype=torch.int32, device=w.device)
    assert result.shape == (nobj, nt, ncol) and result.is_contiguous()
    rtol, atol = _parse_tol(rtol, atol)
    if stream is None:
        stream = torch.cuda.current_stream(w.device)
    code = METHOD_CODES[int_method] if isinstance(int_method, str) else int(int_method)
    rc = _lib.load().gh_integrate_dxdv_hip(
        nobj, dim, w.data_ptr(), nt, t.data_ptr(), int(t.dim() > 1), table.ptr,
        DT_SENTINEL if dt is None else float(dt), rtol, atol, result.data_ptr(), err.data_ptr(), code,
        _ptr(stats), stream.cuda_stream,
    )
    _lib.check(rc, "gh_integrate_dxdv_hip")
    return result, err, stats
