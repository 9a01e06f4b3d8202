"""Potentials for the MI355X orbit engine and galpy's C packing of them.

The classes mirror galpy's potential objects for the six force kernels the
engine implements on the GPU: their constructor arguments, amplitude
conventions and ``normalize`` behave as in galpy
(galpy/potential/Potential.py:902-920, PowerSphericalPotentialwCutoff.py:30-70,
MiyamotoNagaiPotential.py:24-66, TwoPowerSphericalPotential.py:526,776-851,
LogarithmicHaloPotential.py:12,32-75, DehnenBarPotential.py:55-146).
Host-side evaluation (normalization, energies of a few points) uses numpy and
scipy like galpy's Python layer; the integration itself runs in HIP.

``parse_pot`` produces galpy's ``(npot, pot_type, pot_args, pot_tfuncs)``
exactly as ``_parse_pot`` does (galpy/orbit/integrateFullOrbit.py:43-614) for
these classes: same integer codes and argument order.  ``parse_pot_planar``
is the planar packing (galpy/orbit/integratePlanarOrbit.py:33-663, for the
potentials galpy's ``toPlanarPotential`` turns into
planarPotentialFromRZPotential / planarPotentialFromFullPotential).
"""

import numpy
from scipy import special

_CORE = 10**-8  # LogarithmicHaloPotential.py:12


class Potential:
    """Base class: amplitude handling, normalization and force/energy calls."""

    isNonAxi = False
    hasC = True
    # integer code understood by the C-ABI (integrateFullOrbit.c:80-345)
    pot_code = None

    def __init__(self, amp=1.0):
        self._amp = float(amp)

    def normalize(self, norm):
        """Potential.py:902-920: scale amp so |Rforce(1,0)| = norm."""
        self._amp *= norm / numpy.fabs(self.Rforce(1.0, 0.0))

    def __call__(self, R, z=0.0, phi=0.0, t=0.0):
        return self._amp * self._evaluate(R, z, phi, t)

    def Rforce(self, R, z=0.0, phi=0.0, t=0.0):
        return self._amp * self._Rforce(R, z, phi, t)

    def zforce(self, R, z=0.0, phi=0.0, t=0.0):
        return self._amp * self._zforce(R, z, phi, t)

    def phitorque(self, R, z=0.0, phi=0.0, t=0.0):
        return self._amp * self._phitorque(R, z, phi, t)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def c_args(self):
        raise NotImplementedError

    def c_args_planar(self):
        """Arguments in the planar packing (integratePlanarOrbit.py:67-196);
        the same as the full packing except where noted."""
        return self.c_args()

    def __add__(self, other):
        return CompositePotential(*(_as_list(self) + _as_list(other)))

    def __radd__(self, other):
        return CompositePotential(*(_as_list(other) + _as_list(self)))


class CompositePotential(list):
    """A sum of potentials (galpy's CompositePotential is likewise a list)."""

    def __init__(self, *pots):
        super().__init__(pots)

    def __call__(self, R, z=0.0, phi=0.0, t=0.0):
        return sum(p(R, z, phi, t) for p in self)

    def Rforce(self, R, z=0.0, phi=0.0, t=0.0):
        return sum(p.Rforce(R, z, phi, t) for p in self)

    def zforce(self, R, z=0.0, phi=0.0, t=0.0):
        return sum(p.zforce(R, z, phi, t) for p in self)

    def phitorque(self, R, z=0.0, phi=0.0, t=0.0):
        return sum(p.phitorque(R, z, phi, t) for p in self)

    @property
    def isNonAxi(self):
        return any(p.isNonAxi for p in self)

    def __add__(self, other):
        return CompositePotential(*(list(self) + _as_list(other)))


def _as_list(p):
    if isinstance(p, (list, tuple)):
        return list(p)
    return [p]


class PowerSphericalPotentialwCutoff(Potential):
    """rho ~ amp r^-alpha exp(-(r/rc)^2); pot_type 15, args [amp, alpha, rc, 0, 0]."""

    pot_code = 15

    def __init__(self, amp=1.0, alpha=1.0, rc=1.0, normalize=False, r1=1.0):
        Potential.__init__(self, amp=amp)
        self.alpha = alpha
        self._amp *= r1**self.alpha
        self.rc = rc
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _mass(self, r):
        # PowerSphericalPotentialwCutoff.py:206-230 (hyp1f1 form)
        return (
            2.0 * numpy.pi * r ** (3.0 - self.alpha) / (1.5 - self.alpha / 2.0)
            * special.hyp1f1(1.5 - self.alpha / 2.0, 2.5 - self.alpha / 2.0, -((r / self.rc) ** 2.0))
        )

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R**2.0 + z**2.0)
        return (
            2.0 * numpy.pi * self.rc ** (3.0 - self.alpha) / r
            * (
                r / self.rc * special.gamma(1.0 - self.alpha / 2.0)
                * special.gammainc(1.0 - self.alpha / 2.0, (r / self.rc) ** 2.0)
                - special.gamma(1.5 - self.alpha / 2.0)
                * special.gammainc(1.5 - self.alpha / 2.0, (r / self.rc) ** 2.0)
            )
        )

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R * R + z * z)
        return -self._mass(r) * R / r**3.0

    def _zforce(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R * R + z * z)
        return -self._mass(r) * z / r**3.0

    def c_args(self):
        return [self._amp, self.alpha, self.rc, 0.0, 0.0]  # two cache slots

    def c_args_planar(self):
        return [self._amp, self.alpha, self.rc]  # integratePlanarOrbit.py:192-196: no cache


class MiyamotoNagaiPotential(Potential):
    """pot_type 5, args [amp, a, b]."""

    pot_code = 5

    def __init__(self, amp=1.0, a=1.0, b=0.1, normalize=False):
        Potential.__init__(self, amp=amp)
        self._a = a
        self._b = b
        self._b2 = b**2.0
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return -1.0 / numpy.sqrt(R**2.0 + (self._a + numpy.sqrt(z**2.0 + self._b2)) ** 2.0)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return -R / (R**2.0 + (self._a + numpy.sqrt(z**2.0 + self._b2)) ** 2.0) ** 1.5

    def _zforce(self, R, z, phi=0.0, t=0.0):
        sqrtbz = numpy.sqrt(self._b2 + z**2.0)
        asqrtbz = self._a + sqrtbz
        return -z * asqrtbz / sqrtbz / (R**2.0 + asqrtbz**2.0) ** 1.5

    def c_args(self):
        return [self._amp, self._a, self._b]


class NFWPotential(Potential):
    """pot_type 9, args [amp, a]."""

    pot_code = 9

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R**2.0 + z**2.0)
        return -special.xlogy(1.0 / r, 1.0 + r / self.a)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        Rz = R**2.0 + z**2.0
        sqrtRz = numpy.sqrt(Rz)
        return R * (1.0 / Rz / (self.a + sqrtRz) - numpy.log(1.0 + sqrtRz / self.a) / sqrtRz / Rz)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        Rz = R**2.0 + z**2.0
        sqrtRz = numpy.sqrt(Rz)
        return z * (1.0 / Rz / (self.a + sqrtRz) - numpy.log(1.0 + sqrtRz / self.a) / sqrtRz / Rz)

    def c_args(self):
        return [self._amp, self.a]


class HernquistPotential(Potential):
    """pot_type 8, args [amp, a]."""

    pot_code = 8

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return -1.0 / (1.0 + numpy.sqrt(R**2.0 + z**2.0) / self.a) / 2.0 / self.a

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        sqrtRz = numpy.sqrt(R**2.0 + z**2.0)
        return -R / self.a / sqrtRz * (1.0 + sqrtRz / self.a) ** -2.0 / 2.0 / self.a

    def _zforce(self, R, z, phi=0.0, t=0.0):
        sqrtRz = numpy.sqrt(R**2.0 + z**2.0)
        return -z / self.a / sqrtRz * (1.0 + sqrtRz / self.a) ** -2.0 / 2.0 / self.a

    def c_args(self):
        return [self._amp, self.a]


class LogarithmicHaloPotential(Potential):
    """pot_type 0, args [amp, q, core^2, 1-1/b^2 (2.0 when axisymmetric)]."""

    pot_code = 0

    def __init__(self, amp=1.0, core=_CORE, q=1.0, b=None, normalize=False):
        Potential.__init__(self, amp=amp)
        self._core2 = core**2.0
        self._q = q
        self._b = b
        if b is not None:
            self.isNonAxi = True
            self._1m1overb2 = 1.0 - 1.0 / b**2.0
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _Rt2(self, R, phi):
        if self.isNonAxi:
            return R**2.0 * (1.0 - self._1m1overb2 * numpy.sin(phi) ** 2.0)
        return R**2.0

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return 0.5 * numpy.log(self._Rt2(R, phi) + (z / self._q) ** 2.0 + self._core2)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        Rt2 = self._Rt2(R, phi)
        return -Rt2 / R / (Rt2 + (z / self._q) ** 2.0 + self._core2)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return -z / self._q**2.0 / (self._Rt2(R, phi) + (z / self._q) ** 2.0 + self._core2)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        if not self.isNonAxi:
            return 0.0 * numpy.asarray(R, dtype=float)
        return (
            R**2.0 / (self._Rt2(R, phi) + (z / self._q) ** 2.0 + self._core2)
            * numpy.sin(2.0 * phi) * self._1m1overb2 / 2.0
        )

    def c_args(self):
        if self.isNonAxi:
            return [self._amp, self._q, self._core2, self._1m1overb2]
        return [self._amp, self._q, self._core2, 2.0]


class DehnenBarPotential(Potential):
    """pot_type 1, args [amp*Af, tform, tsteady, rb, omegab, barphi]."""

    pot_code = 1
    isNonAxi = True

    def __init__(self, amp=1.0, omegab=None, rb=None, chi=0.8, rolr=0.9,
                 barphi=25.0 * numpy.pi / 180.0, tform=-4.0, tsteady=None, beta=0.0,
                 alpha=0.01, Af=None):
        Potential.__init__(self, amp=amp)
        self._barphi = barphi
        if omegab is None:
            self._rolr = rolr
            self._chi = chi
            self._beta = beta
            self._omegab = 1.0 / ((self._rolr ** (1.0 - self._beta)) / (1.0 + numpy.sqrt((1.0 + self._beta) / 2.0)))
            self._rb = self._chi * self._omegab ** (1.0 / (self._beta - 1.0))
            self._alpha = alpha
            self._af = self._alpha / 3.0 / self._rb**3.0
        else:
            self._omegab = omegab
            self._rb = rb
            self._af = Af
        self._tb = 2.0 * numpy.pi / self._omegab if self._omegab != 0.0 else 1.0
        self._tform = tform * self._tb
        if tsteady is None:
            self._tsteady = self._tform / 2.0
        else:
            self._tsteady = self._tform + tsteady * self._tb

    def OmegaP(self):
        return self._omegab

    def _smooth(self, t):
        t = numpy.asarray(t, dtype=float)
        deltat = t - self._tform
        xi = 2.0 * deltat / (self._tsteady - self._tform) - 1.0
        mid = 3.0 / 16.0 * xi**5.0 - 5.0 / 8 * xi**3.0 + 15.0 / 16.0 * xi + 0.5
        return numpy.where(t < self._tform, 0.0, numpy.where(t < self._tsteady, mid, 1.0))

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        r2 = R**2.0 + z**2.0
        r = numpy.sqrt(r2)
        inner = ((r / self._rb) ** 3.0 - 2.0) * R**2.0 / r2
        outer = -((self._rb / r) ** 3.0) / (1.0 + z**2.0 / R**2.0)
        out = numpy.where(r <= self._rb, inner, outer)
        return out * self._af * self._smooth(t) * numpy.cos(2.0 * (phi - self._omegab * t - self._barphi))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        r = numpy.sqrt(R**2.0 + z**2.0)
        c = self._af * self._smooth(t) * numpy.cos(2.0 * (phi - self._omegab * t - self._barphi))
        inner = -c * ((r / self._rb) ** 3.0 * R * (3.0 * R**2.0 + 2.0 * z**2.0) - 4.0 * R * z**2.0) / r**4.0
        outer = -c * (self._rb / r) ** 3.0 * R / r**4.0 * (3.0 * R**2.0 - 2.0 * z**2.0)
        return numpy.where(r <= self._rb, inner, outer)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        r = numpy.sqrt(R**2.0 + z**2.0)
        c = self._af * self._smooth(t) * numpy.cos(2.0 * (phi - self._omegab * t - self._barphi))
        inner = -c * ((r / self._rb) ** 3.0 + 4.0) * R**2.0 * z / r**4.0
        outer = -5.0 * c * (self._rb / r) ** 3.0 * R**2.0 * z / r**4.0
        return numpy.where(r <= self._rb, inner, outer)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        r2 = R**2.0 + z**2.0
        r = numpy.sqrt(r2)
        s = self._af * self._smooth(t) * numpy.sin(2.0 * (phi - self._omegab * t - self._barphi))
        inner = 2.0 * s * ((r / self._rb) ** 3.0 - 2.0) * R**2.0 / r2
        outer = -2.0 * s * (self._rb / r) ** 3.0 * R**2.0 / r2
        return numpy.where(r <= self._rb, inner, outer)

    def c_args(self):
        return [self._amp * self._af, self._tform, self._tsteady, self._rb, self._omegab, self._barphi]


def MWPotential2014_factory():
    """mwpotentials.py:25-29 (Bovy 2014, Table 1)."""
    return CompositePotential(
        PowerSphericalPotentialwCutoff(normalize=0.05, alpha=1.8, rc=1.9 / 8.0),
        MiyamotoNagaiPotential(a=3.0 / 8.0, b=0.28 / 8.0, normalize=0.6),
        NFWPotential(a=2.0, normalize=0.35),
    )


MWPotential2014 = MWPotential2014_factory()


def flatten(pot):
    """Return the potential as a flat list of components."""
    out = []
    for p in _as_list(pot):
        if isinstance(p, (list, tuple)):
            out.extend(flatten(p))
        else:
            out.append(p)
    return out


def parse_pot(pot):
    """galpy's ``_parse_pot`` (integrateFullOrbit.py:43-614) for the supported
    potentials: returns ``(npot, pot_type int32[], pot_args float64[], [])``.

    Raises NotImplementedError for a potential without a GPU kernel, as galpy
    does for a potential without a C implementation (:604-611)."""
    comps = flatten(pot)
    pot_type, pot_args = [], []
    for p in comps:
        if not isinstance(p, Potential) or p.pot_code is None:
            raise NotImplementedError(
                f"Potential {type(p).__name__} is not supported by the HIP orbit-integration backend"
            )
        pot_type.append(p.pot_code)
        pot_args.extend(p.c_args())
    return (
        len(comps),
        numpy.array(pot_type, dtype=numpy.int32, order="C"),
        numpy.array(pot_args, dtype=numpy.float64, order="C"),
        [],
    )


def parse_pot_planar(pot):
    """galpy's planar ``_parse_pot`` (integratePlanarOrbit.py:33-663) for the
    supported potentials, taken in the plane z = 0: returns
    ``(npot, pot_type int32[], pot_args float64[], [])``."""
    comps = flatten(pot)
    pot_type, pot_args = [], []
    for p in comps:
        if not isinstance(p, Potential) or p.pot_code is None:
            raise NotImplementedError(
                f"Potential {type(p).__name__} is not supported by the HIP orbit-integration backend"
            )
        pot_type.append(p.pot_code)
        pot_args.extend(p.c_args_planar())
    return (
        len(comps),
        numpy.array(pot_type, dtype=numpy.int32, order="C"),
        numpy.array(pot_args, dtype=numpy.float64, order="C"),
        [],
    )


def evaluatePotentials(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p(R, z, phi, t) for p in flatten(pot))


def evaluateRforces(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p.Rforce(R, z, phi, t) for p in flatten(pot))


def evaluatezforces(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p.zforce(R, z, phi, t) for p in flatten(pot))
