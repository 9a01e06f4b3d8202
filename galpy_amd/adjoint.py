"""Gradients of orbit integrations: a reverse-time adjoint behind
``torch.autograd.Function``.

``orbit_flow(w0, theta, spec, dt, nsteps, method)`` integrates the
rectangular states ``w0`` ([n, 6]: x, y, z, vx, vy, vz) for ``nsteps`` steps
of ``dt`` with galpy's leapfrog or symplec4 (galpy/util/bovy_symplecticode.c,
one step per output interval) in a potential whose selected packed arguments
are the entries of ``theta``, and returns the final states.  Backward runs
the HIP adjoint kernel: the trajectory is recomputed backwards by the
reversibility of the symmetric splitting (no trajectory is stored), the
adjoint is carried through every drift and kick with Hessian-vector products,
and dL/dtheta is accumulated per orbit and summed.

``ParamSpec`` names which packed arguments theta holds, e.g. NFW amp and a
and MiyamotoNagai a and b of MWPotential2014::

    spec = ParamSpec(MWPotential2014, [(2, 0), (2, 1), (1, 1), (1, 2)])
"""

import ctypes

import numpy

from . import _lib
from .integrate import METHOD_CODES, PotentialTable
from .potential import parse_pot

_SIGS = {
    "gh_flow_hip": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
         ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p],
    ),
    "gh_adjoint_hip": (
        ctypes.c_int,
        [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_longlong,
         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
         ctypes.c_void_p],
    ),
}
MAX_POT = 8
NPAR = 3
# packed arguments per component in galpy's order (integrateFullOrbit.py:65-365)
_NARGS = {0: 4, 1: 6, 3: 8, 4: 6, 5: 3, 6: 4, 16: 3, 25: 13, 7: 2, 8: 2, 9: 2, 10: 2, 14: 2, 15: 5, 17: 2, 18: 2, 19: 2,
          20: 2, 33: 2, 34: 3, 35: 3, 12: 4, 40: 0}
# the packed arguments the adjoint kernel accumulates dL/dtheta for:
# MiyamotoNagai (amp, a, b), NFW and Hernquist (amp, a), every other
# component with a device Hessian its amplitude only
_DIFFERENTIABLE = {5: (0, 1, 2), 8: (0, 1), 9: (0, 1), 0: (0,), 15: (0,), 7: (0,), 10: (0,),
                   14: (0,), 17: (0,), 18: (0,), 19: (0,), 20: (0,), 33: (0,), 34: (0,), 35: (0,),
                   12: (0,)}


def _lib_adj():
    lib = _lib.load()
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class _RawTable:
    """Device table from already-packed arrays (theta substituted)."""

    def __init__(self, pot_type, pot_args):
        lib = _lib.load()
        self.handle = ctypes.c_void_p()
        pt = numpy.ascontiguousarray(pot_type, dtype=numpy.int32)
        pa = numpy.ascontiguousarray(pot_args, dtype=numpy.float64)
        _lib.check(lib.gh_potential_upload(len(pt), pt.ctypes.data, pa.ctypes.data, ctypes.byref(self.handle)),
                   "gh_potential_upload")

    @property
    def ptr(self):
        return self.handle.value

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                _lib.load().gh_potential_free(self.handle)
                self.handle = None
        except Exception:
            pass


class ParamSpec:
    """Which packed potential arguments are differentiable parameters.

    params: list of (component index, packed-argument index) pairs; the
    argument order is galpy's (integrateFullOrbit.py:65-181), e.g.
    MiyamotoNagai (amp, a, b), NFW (amp, a)."""

    def __init__(self, pot, params):
        self.npot, self.pot_type, self.pot_args, _ = parse_pot(pot)
        if any(int(c) < 0 for c in self.pot_type):
            raise ValueError("the adjoint does not differentiate wrapped (time-dependent) potentials")
        self.offsets = numpy.cumsum([0] + [_NARGS[int(c)] for c in self.pot_type])[:-1]
        self.params = [(int(c), int(j)) for c, j in params]
        for c, j in self.params:
            if not 0 <= c < self.npot:
                raise ValueError(f"component index {c} out of range for {self.npot} components")
            code = int(self.pot_type[c])
            if j not in _DIFFERENTIABLE.get(code, ()):
                raise ValueError(
                    f"packed argument {j} of component {c} (pot_type {code}) is not differentiable; "
                    f"the adjoint differentiates arguments {sorted(_DIFFERENTIABLE.get(code, ()))}"
                )

    def theta0(self):
        return numpy.array([self.pot_args[self.offsets[c] + j] for c, j in self.params])

    def packed(self, theta):
        args = self.pot_args.copy()
        for v, (c, j) in zip(numpy.asarray(theta, dtype=float), self.params):
            args[self.offsets[c] + j] = v
        return self.pot_type, args


def flow_raw(table, w0, dt, nsteps, method="symplec4_hip", t0=0.0):
    """Forward flow only (no autograd): final states on the device."""
    import torch

    wT = torch.empty_like(w0)
    code = METHOD_CODES[method] if isinstance(method, str) else int(method)
    rc = _lib_adj().gh_flow_hip(w0.shape[0], w0.data_ptr(), wT.data_ptr(), float(t0), float(dt),
                                int(nsteps), table.ptr, code, torch.cuda.current_stream(w0.device).cuda_stream)
    _lib.check(rc, "gh_flow_hip")
    return wT


def adjoint_raw(table, wT, gT, dt, nsteps, method="symplec4_hip"):
    """Reverse pass: (dL/dw0, recomputed w0, per-orbit dL/dargs [n, 8, 3])."""
    import torch

    n = wT.shape[0]
    lam0 = torch.empty_like(wT)
    w0rec = torch.empty_like(wT)
    dth = torch.empty((n, MAX_POT, NPAR), dtype=torch.float64, device=wT.device)
    code = METHOD_CODES[method] if isinstance(method, str) else int(method)
    rc = _lib_adj().gh_adjoint_hip(n, wT.data_ptr(), gT.contiguous().data_ptr(), float(dt), int(nsteps),
                                   table.ptr, code, lam0.data_ptr(), w0rec.data_ptr(), dth.data_ptr(),
                                   torch.cuda.current_stream(wT.device).cuda_stream)
    _lib.check(rc, "gh_adjoint_hip")
    return lam0, w0rec, dth


def _make_function():
    import torch

    class OrbitFlow(torch.autograd.Function):
        """w_T = Phi_{N dt}(w_0; theta); backward by the reverse-time adjoint."""

        @staticmethod
        def forward(ctx, w0, theta, spec, dt, nsteps, method):
            pt, pa = spec.packed(theta.detach().cpu().numpy())
            table = _RawTable(pt, pa)
            wT = flow_raw(table, w0.detach().contiguous(), dt, nsteps, method)
            ctx.table = table
            ctx.spec = spec
            ctx.args = (dt, nsteps, method)
            ctx.theta_dtype = theta.dtype
            ctx.save_for_backward(wT)
            return wT

        @staticmethod
        def backward(ctx, gT):
            (wT,) = ctx.saved_tensors
            dt, nsteps, method = ctx.args
            lam0, _, dth = adjoint_raw(ctx.table, wT, gT.contiguous(), dt, nsteps, method)
            tot = dth.sum(dim=0)
            gtheta = torch.stack([tot[c, j] for c, j in ctx.spec.params]).to(ctx.theta_dtype)
            return lam0, gtheta, None, None, None, None

    return OrbitFlow


_OrbitFlow = None


def orbit_flow(w0, theta, spec, dt, nsteps, method="symplec4_hip"):
    """Differentiable final states: w0 [n,6] float64 cuda, theta [len(spec.params)]."""
    global _OrbitFlow
    if _OrbitFlow is None:
        _OrbitFlow = _make_function()
    return _OrbitFlow.apply(w0, theta, spec, float(dt), int(nsteps), method)


def cyl_to_rect_torch(v):
    """(R,vR,vT,z,vz,phi) -> (x,y,z,vx,vy,vz), differentiable (bovy_coords.c:101-115)."""
    import torch

    R, vR, vT, z, vz, phi = v.unbind(-1)
    c, s = torch.cos(phi), torch.sin(phi)
    return torch.stack([R * c, R * s, z, vR * c - vT * s, vR * s + vT * c, vz], dim=-1)
