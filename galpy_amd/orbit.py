"""``Orbit``: the user-facing container, mirroring galpy's ``Orbit`` for the
HIP integration methods.

Mirrors galpy/orbit/Orbits.py: construction from ``[R,vR,vT,z,vz,phi]``
rows, or planar ``[R,vR,vT,phi]`` rows (one or many orbits, :293),
``integrate(t, pot, method, dt, rtol, atol)`` (:1987-2061) dispatching to the
C-ABI through ``integrateFullOrbit_hip`` / ``integratePlanarOrbit_hip`` (the
``integrateFullOrbit_c`` / ``integratePlanarOrbit_c`` calls at :1806 and
:1795), and the phase-space and
integral accessors used by galpy's tests: ``R, vR, vT, z, vz, phi, x, y,
vx, vy`` (:2868), ``getOrbit``, ``E`` (:2902), ``Jacobi`` (:3298), ``L``,
``Lz``, ``ER``, ``Ez``, ``rperi``, ``rap``, ``zmax``, ``e``.

Only the HIP methods (``leapfrog_hip``, ``rk4_hip``, ``rk6_hip``,
``symplec4_hip``, ``symplec6_hip``, ``dopr54_hip``, ``dop853_hip``,
``ias15_hip``) integrate; there is no silent fallback to another backend.  As in galpy,
``numpy.copy`` of the initial conditions is passed (Orbits.py:1793) because
the C-ABI overwrites its ``yo`` argument.
"""

import numpy

from .integrate import (
    METHOD_CODES,
    integrateFullOrbit_dxdv_hip,
    integrateFullOrbit_hip,
    integratePlanarOrbit_dxdv_hip,
    integratePlanarOrbit_hip,
)
from .potential import evaluatePotentials, flatten

_METHODS = tuple(METHOD_CODES)


def _check_integrate_dt(t, dt):
    """Orbits.py:9456-9462: dt must divide the output spacing (rounding
    check, tolerance 1e-10); the C code then truncates init_dt/dt."""
    if dt is None:
        return True
    dt = numpy.fabs(dt)
    time_step = numpy.fabs(t[1] - t[0])
    return numpy.fabs(round(time_step / dt) * dt - time_step) <= 10.0**-10.0


class Orbit:
    """One or many orbits in a potential; rows are (R,vR,vT,z,vz,phi), or
    (R,vR,vT,phi) for planar orbits."""

    def __init__(self, vxvv):
        vxvv = numpy.array(vxvv, dtype=float)
        self._single = vxvv.ndim == 1
        vxvv = numpy.atleast_2d(vxvv)
        if vxvv.shape[-1] not in (4, 6):
            raise ValueError(
                "this Orbit supports full 3D orbits, rows of (R,vR,vT,z,vz,phi), and planar "
                "orbits, rows of (R,vR,vT,phi)"
            )
        self.vxvv = vxvv
        self._planar = vxvv.shape[-1] == 4
        self.t = None
        self.orbit = None
        self._pot = None
        self._err = None

    # --- shape ---------------------------------------------------------------
    @property
    def shape(self):
        return () if self._single else (len(self.vxvv),)

    def __len__(self):
        return 1 if self._single else len(self.vxvv)

    @property
    def size(self):
        return len(self.vxvv)

    def dim(self):
        return 2 if self._planar else 3

    def phasedim(self):
        return self.vxvv.shape[-1]

    # --- integration --------------------------------------------------------
    def integrate(self, t, pot, method="symplec4_hip", progressbar=False, dt=None, rtol=None,
                  atol=None):
        """Integrate every orbit over the times ``t`` (1D, shared by all
        orbits, or [N, nt] per orbit), galpy's Orbits.py:1624-1956 path for a
        C method: checks, one integration call, and storage of the rows."""
        if method.lower() not in _METHODS:
            raise ValueError(
                f"{method} is not a method of this engine; use one of {', '.join(_METHODS)}"
            )
        t = numpy.array(t, dtype=float)
        tt = t if t.ndim == 1 else t[0]
        if len(tt) < 2:
            raise ValueError("integrate needs at least two output times")
        steps = numpy.diff(tt)
        if not numpy.allclose(steps, steps[0], rtol=1e-10, atol=1e-14):
            raise ValueError("Input time array must be equally spaced")
        if not _check_integrate_dt(tt, dt):
            raise ValueError("dt input (integrator stepsize) for Orbit.integrate must be an integer divisor of the output stepsize")
        flatten(pot)  # validates the potential
        integrator = integratePlanarOrbit_hip if self._planar else integrateFullOrbit_hip
        out, err = integrator(pot, numpy.copy(self.vxvv), t, method.lower(), rtol=rtol, atol=atol,
                              progressbar=progressbar, dt=dt)
        self.orbit = out
        self._err = err
        self.t = t
        self._pot = pot
        return None

    def integrate_dxdv(self, dxdv, t, pot, method="dop853_hip", rectIn=False, rectOut=False,
                       dt=None, rtol=None, atol=None):
        """Integrate the orbits together with one phase-space deviation each
        (Orbits.py:2276-2470 for the HIP methods): ``dxdv`` is ``[N, phasedim]``
        (or ``[phasedim]``), cylindrical (dR,dvR,dvT,dz,dvz,dphi) /
        (dR,dvR,dvT,dphi) unless ``rectIn``; the deviation along the orbit
        comes back from ``getOrbit_dxdv`` cylindrical unless ``rectOut``
        (integrateFullOrbit.py:926-1037, integratePlanarOrbit.py:1452-1540)."""
        if method.lower() not in _METHODS:
            raise ValueError(
                f"{method} is not a method of this engine; use one of {', '.join(_METHODS)}"
            )
        t = numpy.array(t, dtype=float)
        if t.ndim != 1 or len(t) < 2:
            raise ValueError("integrate_dxdv needs a 1D array of at least two output times")
        steps = numpy.diff(t)
        if not numpy.allclose(steps, steps[0], rtol=1e-10, atol=1e-14):
            raise ValueError("Input time array must be equally spaced")
        if not _check_integrate_dt(t, dt):
            raise ValueError("dt input (integrator stepsize) for Orbit.integrate must be an integer divisor of the output stepsize")
        flatten(pot)
        d = numpy.atleast_2d(numpy.array(dxdv, dtype=float))
        if d.shape[-1] != self.phasedim():
            raise ValueError(f"dxdv must have {self.phasedim()} components per orbit")
        if len(d) == 1 and len(self.vxvv) > 1:
            d = numpy.repeat(d, len(self.vxvv), axis=0)
        y = self.vxvv
        if self._planar:
            R, vR, vT, phi = y.T
            yr = numpy.stack(_cyl_to_rect_state(R, vR, vT, phi), axis=1)
            dr = d if rectIn else numpy.stack(_cyl_to_rect_dev(R, vR, vT, phi, *d.T), axis=1)
            out, err = integratePlanarOrbit_dxdv_hip(pot, yr, dr, t, method.lower(), dt=dt,
                                                     rtol=rtol, atol=atol)
            x, yy, vx, vy = numpy.moveaxis(out[..., :4], -1, 0)
            R, vR, vT, phi = _rect_to_cyl_state(x, yy, vx, vy, wrap=True)
            base = numpy.stack([R, vR, vT, phi], axis=-1)
            dev = out[..., 4:]
            if not rectOut:
                dR, dvR, dvT, dphi = _rect_to_cyl_dev(R, vR, vT, phi, *numpy.moveaxis(dev, -1, 0))
                dev = numpy.stack([dR, dvR, dvT, dphi], axis=-1)
        else:
            R, vR, vT, z, vz, phi = y.T
            x, yy, vx, vy = _cyl_to_rect_state(R, vR, vT, phi)
            yr = numpy.stack([x, yy, z, vx, vy, vz], axis=1)
            if rectIn:
                dr = d
            else:
                dR, dvR, dvT, dz, dvz, dphi = d.T
                dx, dy, dvx, dvy = _cyl_to_rect_dev(R, vR, vT, phi, dR, dvR, dvT, dphi)
                dr = numpy.stack([dx, dy, dz, dvx, dvy, dvz], axis=1)
            out, err = integrateFullOrbit_dxdv_hip(pot, yr, dr, t, method.lower(), dt=dt,
                                                   rtol=rtol, atol=atol)
            x, yy, z, vx, vy, vz = numpy.moveaxis(out[..., :6], -1, 0)
            R, vR, vT, phi = _rect_to_cyl_state(x, yy, vx, vy, wrap=False)
            base = numpy.stack([R, vR, vT, z, vz, phi], axis=-1)
            dev = out[..., 6:]
            if not rectOut:
                dx, dy, dz, dvx, dvy, dvz = numpy.moveaxis(dev, -1, 0)
                dR, dvR, dvT, dphi = _rect_to_cyl_dev(R, vR, vT, phi, dx, dy, dvx, dvy)
                dev = numpy.stack([dR, dvR, dvT, dz, dvz, dphi], axis=-1)
        self.orbit = base
        self.orbit_dxdv = dev
        self._err = err
        self.t = t
        self._pot = pot
        return None

    def getOrbit_dxdv(self):
        """The deviation along the orbit from ``integrate_dxdv`` (Orbits.py getOrbit_dxdv)."""
        if getattr(self, "orbit_dxdv", None) is None:
            raise AttributeError("Integrate orbit with integrate_dxdv first")
        return self._out(self.orbit_dxdv.copy())

    def _check_integrated(self):
        if self.orbit is None:
            raise AttributeError("Integrate orbit first")

    def _rows(self, t=None):
        """Rows at times t (None: all output times); t must lie on the grid."""
        self._check_integrated()
        if t is None:
            return self.orbit
        grid = self.t if self.t.ndim == 1 else self.t[0]
        ts = numpy.atleast_1d(numpy.asarray(t, dtype=float))
        idx = numpy.searchsorted(grid, ts) if grid[-1] > grid[0] else numpy.searchsorted(-grid, -ts)
        idx = numpy.clip(idx, 0, len(grid) - 1)
        for k, tv in enumerate(ts):
            for cand in (idx[k] - 1, idx[k], idx[k] + 1):
                if 0 <= cand < len(grid) and abs(grid[cand] - tv) <= 1e-10 * max(1.0, abs(tv)):
                    idx[k] = cand
                    break
            else:
                raise ValueError(f"time {tv} is not one of the output times of the integration")
        rows = self.orbit[:, idx]
        return rows[:, 0] if numpy.ndim(t) == 0 else rows

    def _out(self, a):
        return a[0] if self._single else a

    def getOrbit(self):
        self._check_integrated()
        return self._out(self.orbit.copy())

    def time(self):
        self._check_integrated()
        return self.t

    # --- coordinates --------------------------------------------------------
    def _col(self, k, t):
        if self.orbit is None and t is None:
            return self._out(self.vxvv[:, k])
        return self._out(self._rows(t)[..., k])

    def R(self, t=None):
        return self._col(0, t)

    def vR(self, t=None):
        return self._col(1, t)

    def vT(self, t=None):
        return self._col(2, t)

    def z(self, t=None):
        if self._planar:
            raise AttributeError("linear and planar orbits do not have z()")  # Orbits.py:4726
        return self._col(3, t)

    def vz(self, t=None):
        if self._planar:
            raise AttributeError("linear and planar orbits do not have vz()")
        return self._col(4, t)

    def phi(self, t=None):
        return self._col(self.phasedim() - 1, t)

    def x(self, t=None):
        return self.R(t) * numpy.cos(self.phi(t))

    def y(self, t=None):
        return self.R(t) * numpy.sin(self.phi(t))

    def vx(self, t=None):
        phi = self.phi(t)
        return self.vR(t) * numpy.cos(phi) - self.vT(t) * numpy.sin(phi)

    def vy(self, t=None):
        phi = self.phi(t)
        return self.vR(t) * numpy.sin(phi) + self.vT(t) * numpy.cos(phi)

    # --- integrals ------------------------------------------------------------
    def _times_like(self, t, rows):
        if t is not None:
            return numpy.broadcast_to(numpy.asarray(t, dtype=float), rows.shape[:-1])
        tt = self.t if self.t.ndim > 1 else self.t[None, :]
        return numpy.broadcast_to(tt, rows.shape[:-1])

    def E(self, t=None, pot=None):
        """Energy (Orbits.py:2902): potential + kinetic, per orbit and time."""
        pot = self._pot if pot is None else pot
        rows = self._rows(t) if self.orbit is not None else self.vxvv
        tt = self._times_like(t, rows) if self.orbit is not None else 0.0
        if self._planar:
            R, vR, vT, phi = numpy.moveaxis(rows, -1, 0)
            return self._out(evaluatePotentials(pot, R, 0.0, phi, tt) + 0.5 * (vR**2 + vT**2))
        R, vR, vT, z, vz, phi = numpy.moveaxis(rows, -1, 0)
        return self._out(evaluatePotentials(pot, R, z, phi, tt) + 0.5 * (vR**2 + vT**2 + vz**2))

    def Jacobi(self, t=None, OmegaP=None, pot=None):
        """Jacobi integral E - OmegaP Lz (Orbits.py:3298); OmegaP defaults
        to the pattern speed of a bar in the potential."""
        pot = self._pot if pot is None else pot
        if OmegaP is None:
            OmegaP = next((p.OmegaP() for p in flatten(pot) if hasattr(p, "OmegaP")), None)
            if OmegaP is None:
                raise ValueError("OmegaP must be given for a potential without a pattern speed")
        return self.E(t=t, pot=pot) - OmegaP * self.Lz(t)

    def Lz(self, t=None):
        return self.R(t) * self.vT(t)

    def L(self, t=None):
        """Angular momentum vector (x,y,z components) from the rectangular
        state; R vT for planar orbits (Orbits.py:3147-3149)."""
        if self._planar:
            return self.Lz(t)
        x, y, z = self.x(t), self.y(t), self.z(t)
        vx, vy, vz = self.vx(t), self.vy(t), self.vz(t)
        return numpy.stack([y * vz - z * vy, z * vx - x * vz, x * vy - y * vx], axis=-1)

    def ER(self, t=None, pot=None):
        pot = self._pot if pot is None else pot
        R, vR, vT = self.R(t), self.vR(t), self.vT(t)
        return evaluatePotentials(pot, R, 0.0) + 0.5 * (vR**2 + vT**2)

    def Ez(self, t=None, pot=None):
        pot = self._pot if pot is None else pot
        R, z, vz = self.R(t), self.z(t), self.vz(t)
        return evaluatePotentials(pot, R, z) - evaluatePotentials(pot, R, 0.0) + 0.5 * vz**2

    def _r(self):
        o = self._rows()
        if self._planar:
            return o[..., 0]
        return numpy.sqrt(o[..., 0] ** 2 + o[..., 3] ** 2)

    def rperi(self):
        return self._out(self._r().min(axis=-1))

    def rap(self):
        return self._out(self._r().max(axis=-1))

    def zmax(self):
        if self._planar:
            raise AttributeError("linear and planar orbits do not have zmax()")
        return self._out(numpy.fabs(self._rows()[..., 3]).max(axis=-1))

    def e(self):
        r = self._r()
        rap, rperi = r.max(axis=-1), r.min(axis=-1)
        return self._out((rap - rperi) / (rap + rperi))


Orbits = Orbit


# Coordinate maps of integrate_dxdv (integratePlanarOrbit.py:1452-1476 and
# 1517-1540; integrateFullOrbit.py:926-946 and 1003-1037 with
# coords.cyl_to_rect_jac, util/coords.py:1637-1690, inverted in closed form).
def _cyl_to_rect_state(R, vR, vT, phi):
    c, s = numpy.cos(phi), numpy.sin(phi)
    return R * c, R * s, vR * c - vT * s, vR * s + vT * c


def _rect_to_cyl_state(x, y, vx, vy, wrap):
    R = numpy.sqrt(x**2 + y**2)
    if wrap:  # planar: phi in [0, 2 pi) from arccos, as integratePlanarOrbit.py:1519-1520
        phi = numpy.arccos(x / R)
        phi = numpy.where(y < 0.0, 2.0 * numpy.pi - phi, phi)
    else:  # full: atan2, as coords.rect_to_cyl
        phi = numpy.arctan2(y, x)
    c, s = numpy.cos(phi), numpy.sin(phi)
    return R, vx * c + vy * s, vy * c - vx * s, phi


def _cyl_to_rect_dev(R, vR, vT, phi, dR, dvR, dvT, dphi):
    c, s = numpy.cos(phi), numpy.sin(phi)
    return (c * dR - R * s * dphi, s * dR + R * c * dphi,
            c * dvR - s * dvT - (vR * s + vT * c) * dphi,
            s * dvR + c * dvT + (vR * c - vT * s) * dphi)


def _rect_to_cyl_dev(R, vR, vT, phi, dx, dy, dvx, dvy):
    c, s = numpy.cos(phi), numpy.sin(phi)
    dR = c * dx + s * dy
    dphi = (c * dy - s * dx) / R
    return dR, c * dvx + s * dvy + vT * dphi, c * dvy - s * dvx - vR * dphi, dphi
