"""Potentials for the MI355X orbit engine and galpy's C packing of them.

The classes mirror galpy's potential objects for the force kernels the
engine implements on the GPU: their constructor arguments, amplitude
conventions and ``normalize`` behave as in galpy
(galpy/potential/Potential.py:902-920, PowerSphericalPotentialwCutoff.py:30-70,
MiyamotoNagaiPotential.py:24-66, TwoPowerSphericalPotential.py:301-851,
LogarithmicHaloPotential.py:12,32-75, DehnenBarPotential.py:55-146,
PowerSphericalPotential.py:25-60,400-440, IsochronePotential.py:24-53,
PlummerPotential.py:23-51, PseudoIsothermalPotential.py:20-53,
KuzminDiskPotential.py:24-52, BurkertPotential.py:20-53,
HomogeneousSpherePotential.py:23-52).
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
from scipy import interpolate, special

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


# Smith et al. (2015) fits used by MN3ExponentialDiskPotential.py:207-336:
# quartic polynomials in b/Rd of the three Miyamoto-Nagai masses and scale
# lengths (their Table 1: exponential-disk density matched everywhere; Table
# 2: positive density everywhere), and cubics from hz/Rd to b/Rd.
_MN3_TAB1 = {  # coefficients of brd^4, brd^3, brd^2, brd, 1
    "mass": ((-0.0090, 0.0640, -0.1653, 0.1164, 1.9487),
             (0.0173, -0.0903, 0.0877, 0.2029, -1.3077),
             (-0.0051, 0.0287, -0.0361, -0.0544, 0.2242)),
    "a": ((-0.0358, 0.2610, -0.6987, -0.1193, 2.0074),
          (-0.0830, 0.4992, -0.7967, -1.2966, 4.4441),
          (-0.0247, 0.1718, -0.4124, -0.5944, 0.7333)),
}
_MN3_TAB2 = {
    "mass": ((0.0036, -0.0330, 0.1117, -0.1335, 0.1749),
             (-0.0131, 0.1090, -0.3035, 0.2921, -5.7976),
             (-0.0048, 0.0454, -0.1425, 0.1012, 6.7120)),
    "a": ((-0.0158, 0.0993, -0.2070, -0.7089, 0.6445),
          (-0.0319, 0.1514, -0.1279, -0.9325, 2.6836),
          (-0.0326, 0.1816, -0.2943, -0.6329, 2.3193)),
}
_MN3_B_EXP = (-0.269, 1.080, 1.092)  # coefficients of hz^3, hz^2, hz
_MN3_B_SECH = (-0.033, 0.262, 0.659)


def _mn3_quartic(c, x):
    # the reference's form c4 x^4 + c3 x^3 + c2 x^2 + c1 x + c0, term by term
    return c[0] * x**4.0 + c[1] * x**3.0 + c[2] * x**2.0 + c[3] * x + c[4]


class MN3ExponentialDiskPotential(Potential):
    """Three Miyamoto-Nagai disks approximating an exponential (or sech^2)
    disk of scale length hr and height hz (MN3ExponentialDiskPotential.py:
    33-143); packed as three pot_type 5 components with amplitudes
    amp * mass_i (integrateFullOrbit.py:182-198, integratePlanarOrbit.py:
    197-215).  amp is the disk's central density, as in galpy."""

    pot_code = None

    def __init__(self, amp=1.0, hr=1.0 / 3.0, hz=1.0 / 16.0, sech=False, posdens=False,
                 normalize=False):
        Potential.__init__(self, amp=amp)
        self._hr = hr
        self._hz = hz
        self._amp *= 4.0 * numpy.pi * self._hr**2.0 * self._hz
        c = _MN3_B_SECH if sech else _MN3_B_EXP
        x = self._hz / self._hr
        self._brd = c[0] * x**3.0 + c[1] * x**2.0 + c[2] * x
        if self._brd < 0.0:
            raise OSError("MN3ExponentialDiskPotential's b/Rd is negative for the given hz")
        self._b = self._brd * self._hr
        tab = _MN3_TAB2 if posdens else _MN3_TAB1
        self._mn3 = [
            MiyamotoNagaiPotential(amp=_mn3_quartic(m, self._brd), a=_mn3_quartic(a, self._brd) * self._hr,
                                   b=self._b)
            for m, a in zip(tab["mass"], tab["a"])
        ]
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return sum(mn(R, z) for mn in self._mn3)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return sum(mn.Rforce(R, z) for mn in self._mn3)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return sum(mn.zforce(R, z) for mn in self._mn3)

    def expand(self):
        """The three packed components: [amp * mass_i, a_i, b]."""
        return [[self._amp * mn._amp, mn._a, mn._b] for mn in self._mn3]


class KuzminKutuzovStaeckelPotential(Potential):
    """pot_type 16, args [amp, ac, Delta] (KuzminKutuzovStaeckelPotential.py:
    16-140, integrateFullOrbit.py:199-201): Phi = -amp / (sqrt(l) + sqrt(n))
    with (l, n) the prolate spheroidal coordinates of focal distance Delta
    whose coordinate surfaces have axis ratio ac (ac > 1: gamma = Delta^2 /
    (1 - ac^2) < 0, alpha = gamma - Delta^2)."""

    pot_code = 16

    def __init__(self, amp=1.0, ac=5.0, Delta=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self._ac = ac
        self._Delta = Delta
        self._gamma = self._Delta**2.0 / (1.0 - self._ac**2.0)
        self._alpha = self._gamma - self._Delta**2.0
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _ln(self, R, z):
        # the roots l >= n of (t + alpha)(t + gamma) - R^2 (t + gamma) - z^2 (t + alpha)
        r2 = R**2.0 + z**2.0
        term = r2 - self._alpha - self._gamma
        sd = numpy.sqrt((r2 - self._Delta**2.0) ** 2.0 + 4.0 * self._Delta**2.0 * R**2.0)
        return 0.5 * (term + sd), 0.5 * (term - sd), sd

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        l, n, _ = self._ln(R, z)
        return -1.0 / (numpy.sqrt(l) + numpy.sqrt(numpy.maximum(n, 0.0)))

    def _dV(self, R, z):
        l, n, sd = self._ln(R, z)
        s2 = (numpy.sqrt(l) + numpy.sqrt(n)) ** 2.0
        return 0.5 / numpy.sqrt(l) / s2, 0.5 / numpy.sqrt(n) / s2, sd

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        dVdl, dVdn, sd = self._dV(R, z)
        u = (R**2.0 + z**2.0 + self._Delta**2.0) / sd
        return -R * ((1.0 + u) * dVdl + (1.0 - u) * dVdn)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        dVdl, dVdn, sd = self._dV(R, z)
        u = (R**2.0 + z**2.0 - self._Delta**2.0) / sd
        return -z * ((1.0 + u) * dVdl + (1.0 - u) * dVdn)

    def c_args(self):
        return [self._amp, self._ac, self._Delta]


class SoftenedNeedleBarPotential(Potential):
    """pot_type 25, args [amp, a, b, c^2, pa, omegab] and seven cache slots
    (SoftenedNeedleBarPotential.py:30-200, integrateFullOrbit.py:303-306):
    Long & Murali's softened needle of half-length a, rotating at omegab from
    angle pa, Phi = amp/(2a) log((x - a + T-) / (x + a + T+)) in the bar
    frame, T+- = sqrt((a +- x)^2 + y^2 + (b + sqrt(z^2 + c^2))^2)."""

    pot_code = 25

    def __init__(self, amp=1.0, a=4.0, b=0.0, c=1.0, normalize=False, pa=0.4, omegab=1.8):
        Potential.__init__(self, amp=amp)
        self._a = a
        self._b = b
        self._c2 = c**2.0
        self._pa = pa
        self._omegab = omegab
        if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
            self.normalize(normalize)

    def _frame(self, R, z, phi, t):
        ang = phi - self._pa - self._omegab * t
        x, y = R * numpy.cos(ang), R * numpy.sin(ang)
        zc = numpy.sqrt(self._c2 + z**2.0)
        second = y**2.0 + (self._b + zc) ** 2.0
        Tp = numpy.sqrt((self._a + x) ** 2.0 + second)
        Tm = numpy.sqrt((self._a - x) ** 2.0 + second)
        return x, y, zc, second, Tp, Tm

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        x, _, _, _, Tp, Tm = self._frame(R, z, phi, t)
        return numpy.log((x - self._a + Tm) / (x + self._a + Tp)) / 2.0 / self._a

    def _xyz(self, R, z, phi, t):
        x, y, zc, second, Tp, Tm = self._frame(R, z, phi, t)
        S = Tp + Tm
        Fx = -2.0 * x / Tp / Tm / S
        common = 0.5 / Tp / Tm * (S - 4.0 * x**2.0 / S) / second
        cp, sp = numpy.cos(self._pa + self._omegab * t), numpy.sin(self._pa + self._omegab * t)
        return cp * Fx + sp * y * common, sp * Fx - cp * y * common, -z * common * (self._b + zc) / zc

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        Fx, Fy, _ = self._xyz(R, z, phi, t)
        return numpy.cos(phi) * Fx + numpy.sin(phi) * Fy

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        Fx, Fy, _ = self._xyz(R, z, phi, t)
        return R * (-numpy.sin(phi) * Fx + numpy.cos(phi) * Fy)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._xyz(R, z, phi, t)[2]

    def c_args(self):
        return [self._amp, self._a, self._b, self._c2, self._pa, self._omegab] + [0.0] * 7


class EllipticalDiskPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 4, args [amp, tform,
    tsteady, twophio, p, phib] with tform = tsteady = NaN when tform is None
    (EllipticalDiskPotential.py:24-150, integratePlanarOrbit.py:127-135):
    Phi = amp smooth(t) twophio/2 (R/r1)^p cos(2 (phi - phib)), smooth
    Dehnen's growth polynomial from tform to tsteady (tform + 2 unless
    given, as tform + tsteady)."""

    pot_code = 4

    def __init__(self, amp=1.0, phib=25.0 * numpy.pi / 180.0, p=1.0, twophio=0.01, r1=1.0,
                 tform=None, tsteady=None):
        Potential.__init__(self, amp=amp)
        self._amp /= r1**p
        self._phib = phib
        self._twophio = twophio
        self._p = p
        self._tform = tform
        if tform is None:
            self._tsteady = None
        elif tsteady is None:
            self._tsteady = tform + 2.0
        else:
            self._tsteady = tform + tsteady

    def _smooth(self, t):
        if self._tform is None:
            return 1.0
        t = numpy.asarray(t, dtype=float)
        xi = 2.0 * (t - self._tform) / (self._tsteady - self._tform) - 1.0
        mid = 3.0 / 16.0 * xi**5.0 - 5.0 / 8.0 * xi**3.0 + 15.0 / 16.0 * xi + 0.5
        return numpy.where(t < self._tform, 0.0, numpy.where(t < self._tsteady, mid, 1.0))

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._smooth(t) * self._twophio / 2.0 * R**self._p * numpy.cos(2.0 * (phi - self._phib))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return (-self._smooth(t) * self._p * self._twophio / 2.0 * R ** (self._p - 1.0)
                * numpy.cos(2.0 * (phi - self._phib)))

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * R

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return self._smooth(t) * self._twophio * R**self._p * numpy.sin(2.0 * (phi - self._phib))

    def c_args(self):
        raise NotImplementedError("EllipticalDiskPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        tf = float("nan") if self._tform is None else self._tform
        ts = float("nan") if self._tform is None else self._tsteady
        return [self._amp, tf, ts, self._twophio, self._p, self._phib]


class LopsidedDiskPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 6, args [amp, phio, p,
    phib] (CosmphiDiskPotential.py:220-240 with m = 1,
    integratePlanarOrbit.py:142-144): Phi = amp phio R^p cos(phi - phib).
    As in galpy, r1 is accepted and not used (LopsidedDiskPotential does not
    pass it on)."""

    pot_code = 6

    def __init__(self, amp=1.0, phib=25.0 * numpy.pi / 180.0, p=1.0, phio=0.01, r1=1.0):
        Potential.__init__(self, amp=amp)
        self._phib = phib
        self._mphio = phio
        self._p = p

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._mphio * R**self._p * numpy.cos(phi - self._phib)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return -self._p * self._mphio * R ** (self._p - 1.0) * numpy.cos(phi - self._phib)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * R

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return self._mphio * R**self._p * numpy.sin(phi - self._phib)

    def c_args(self):
        raise NotImplementedError("LopsidedDiskPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        return [self._amp, self._mphio, self._p, self._phib]


class SteadyLogSpiralPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 3, args [amp, tform,
    tsteady, A, alpha, m, omegas, gamma] (NaN times when tform is None;
    SteadyLogSpiralPotential.py:20-150, integratePlanarOrbit.py:99-126):
    Phi = amp smooth(t) A/alpha cos(alpha ln R - m (phi - omegas t - gamma)),
    alpha = m / tan(p) when the pitch angle p is given; tform and tsteady in
    units of the pattern period 2 pi / omegas (tsteady: duration, default 2)."""

    pot_code = 3

    def __init__(self, amp=1.0, omegas=0.65, A=-0.035, alpha=-7.0, m=2, gamma=numpy.pi / 4.0, p=None,
                 tform=None, tsteady=None):
        Potential.__init__(self, amp=amp)
        self._omegas = omegas
        self._A = A
        self._m = m
        self._gamma = gamma
        self._alpha = self._m / numpy.tan(p) if p is not None else alpha
        ts = 2.0 * numpy.pi / self._omegas if self._omegas != 0.0 else 1.0
        self._tform = None if tform is None else tform * ts
        if self._tform is None:
            self._tsteady = None
        else:
            self._tsteady = self._tform + (2.0 if tsteady is None else tsteady) * ts

    def _smooth(self, t):
        if self._tform is None:
            return 1.0
        t = numpy.asarray(t, dtype=float)
        xi = 2.0 * (t - self._tform) / (self._tsteady - self._tform) - 1.0
        mid = 3.0 / 16.0 * xi**5.0 - 5.0 / 8.0 * xi**3.0 + 15.0 / 16.0 * xi + 0.5
        return numpy.where(t < self._tform, 0.0, numpy.where(t < self._tsteady, mid, 1.0))

    def _arg(self, R, phi, t):
        return self._alpha * numpy.log(R) - self._m * (phi - self._omegas * t - self._gamma)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._smooth(t) * self._A / self._alpha * numpy.cos(self._arg(R, phi, t))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._smooth(t) * self._A / R * numpy.sin(self._arg(R, phi, t))

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * R

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return -self._smooth(t) * self._A / self._alpha * self._m * numpy.sin(self._arg(R, phi, t))

    def c_args(self):
        raise NotImplementedError("SteadyLogSpiralPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        nan = float("nan")
        tf = nan if self._tform is None else self._tform
        ts = nan if self._tform is None else self._tsteady
        return [self._amp, tf, ts, self._A, self._alpha, float(self._m), self._omegas, self._gamma]


class TransientLogSpiralPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 2, args [amp, A, to,
    sigma^2, alpha, m, omegas, gamma] (TransientLogSpiralPotential.py:20-160,
    integratePlanarOrbit.py:94-98): Phi = amp A exp(-(t - to)^2 / (2
    sigma^2)) / alpha cos(alpha ln R - m (phi - omegas t - gamma)),
    alpha = m / tan(p) when the pitch angle p is given."""

    pot_code = 2
    isNonAxi = True

    def __init__(self, amp=1.0, omegas=0.65, A=-0.035, alpha=-7.0, m=2, gamma=numpy.pi / 4.0, p=None,
                 sigma=1.0, to=0.0):
        Potential.__init__(self, amp=amp)
        self._omegas = omegas
        self._A = A
        self._m = m
        self._gamma = gamma
        self._to = to
        self._sigma2 = sigma**2.0
        self._alpha = self._m / numpy.tan(p) if p is not None else alpha

    def _g(self, t):
        return numpy.exp(-((numpy.asarray(t, dtype=float) - self._to) ** 2.0) / 2.0 / self._sigma2)

    def _arg(self, R, phi, t):
        return self._alpha * numpy.log(R) - self._m * (phi - self._omegas * t - self._gamma)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._A * self._g(t) / self._alpha * numpy.cos(self._arg(R, phi, t))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._A * self._g(t) / R * numpy.sin(self._arg(R, phi, t))

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * R

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return -self._A * self._g(t) / self._alpha * self._m * numpy.sin(self._arg(R, phi, t))

    def c_args(self):
        raise NotImplementedError("TransientLogSpiralPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        return [self._amp, self._A, self._to, self._sigma2, self._alpha, float(self._m), self._omegas,
                self._gamma]


class CosmphiDiskPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 28, args [amp, mphio, p,
    mphib, m, rb, rbp, rb2p, r1p] (CosmphiDiskPotential.py:26-200,
    integratePlanarOrbit.py:410-414): Phi = amp mphio/m (R/r1)^p cos(m phi -
    mphib) outside the break radius rb and amp mphio/m rb^p (2 r1^p - rb^p
    R^-p) cos(m phi - mphib) / r1^p inside (F_R is continuous at rb, the
    potential only for r1 = 1); (cp, sp) = m phio (cos m phib, sin m phib)
    may be given instead of (phib, phio)."""

    pot_code = 28
    isNonAxi = True

    def __init__(self, amp=1.0, phib=25.0 * numpy.pi / 180.0, p=1.0, phio=0.01, m=4, r1=1.0, rb=None,
                 cp=None, sp=None):
        Potential.__init__(self, amp=amp)
        self._r1p = r1**p
        self._amp /= self._r1p
        self._m = int(m)
        if cp is None or sp is None:
            self._phib = phib
            self._mphio = phio * self._m
        else:
            self._mphio = numpy.sqrt(cp * cp + sp * sp)
            self._phib = numpy.arctan(sp / cp) / self._m
            if m < 2.0 and cp < 0.0:
                self._phib = numpy.pi + self._phib
        self._p = p
        if rb is None:
            self._rb, self._rbp, self._rb2p = 0.0, 1.0, 1.0
        else:
            self._rb = rb
            self._rbp = self._rb**self._p
            self._rb2p = self._rbp**2.0
        self._mphib = self._m * self._phib

    def _inner(self, R):
        return numpy.asarray(R, dtype=float) < self._rb

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        R = numpy.asarray(R, dtype=float)
        c = numpy.cos(self._m * phi - self._mphib) * self._mphio / self._m
        return c * numpy.where(self._inner(R), self._rbp * (2.0 * self._r1p - self._rbp / R**self._p),
                               R**self._p)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        R = numpy.asarray(R, dtype=float)
        c = -self._p * self._mphio / self._m * numpy.cos(self._m * phi - self._mphib)
        return c * numpy.where(self._inner(R), self._rb2p / R ** (self._p + 1.0), R ** (self._p - 1.0))

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        R = numpy.asarray(R, dtype=float)
        s = self._mphio * numpy.sin(self._m * phi - self._mphib)
        return s * numpy.where(self._inner(R), self._rbp * (2.0 * self._r1p - self._rbp / R**self._p),
                               R**self._p)

    def c_args(self):
        raise NotImplementedError("CosmphiDiskPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        return [self._amp, self._mphio, self._p, self._mphib, float(self._m), self._rb, self._rbp,
                self._rb2p, self._r1p]


class HenonHeilesPotential(Potential):
    """Planar only: integratePlanarOrbit's pot_type 29, args [amp]
    (HenonHeilesPotential.py:9-60): Phi = amp/2 (R^2 + 2/3 R^3 sin 3 phi)."""

    pot_code = 29
    isNonAxi = True

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return 0.5 * R * R * (1.0 + 2.0 / 3.0 * R * numpy.sin(3.0 * phi))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return -R * (1.0 + R * numpy.sin(3.0 * phi))

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return -(R**3.0) * numpy.cos(3.0 * phi)

    def c_args(self):
        raise NotImplementedError("HenonHeilesPotential is a planar potential: integrate planar orbits")

    def c_args_planar(self):
        return [self._amp]


class SpiralArmsPotential(Potential):
    """Cox & Gomez (2002) spiral arms, pot_type 27, args [nCs, amp, N,
    sin alpha, tan alpha, r_ref, phi_ref, Rs, H, omega, C_1..C_nCs] in both
    packings (SpiralArmsPotential.py:20-120, integrateFullOrbit.py:337-352,
    integratePlanarOrbit.py:391-409):
    Phi = -amp H exp(-(R - r_ref)/Rs) sum_n C_n / (K_n D_n) cos(n gamma)
          / cosh(K_n z / B_n)^B_n,
    gamma = N (phi - omega t - phi_ref - ln(R/r_ref) / tan alpha),
    K_n = n N / (R sin alpha), B_n = K_n H (1 + 0.4 K_n H),
    D_n = (1 + K_n H + 0.3 (K_n H)^2) / (1 + 0.3 K_n H).
    As in galpy, N and alpha are stored with flipped signs (a left-handed
    pattern), and amp is a density amplitude (rho_0 = 1/(4 pi))."""

    pot_code = 27
    isNonAxi = True

    def __init__(self, amp=1.0, N=2, alpha=0.2, r_ref=1.0, phi_ref=0.0, Rs=0.3, H=0.125, omega=0.0, Cs=(1.0,)):
        Potential.__init__(self, amp=amp)
        self._N = -N
        self._alpha = -alpha
        self._sin_alpha = numpy.sin(-alpha)
        self._tan_alpha = numpy.tan(-alpha)
        self._r_ref = r_ref
        self._phi_ref = phi_ref
        self._Rs = Rs
        self._H = H
        self._Cs = numpy.array(Cs, dtype=float)
        self._omega = omega

    def _terms(self, R, z, phi, t):
        R = numpy.asarray(R, dtype=float)[..., None]
        z = numpy.asarray(z, dtype=float)[..., None]
        phi = numpy.asarray(phi, dtype=float)[..., None]
        t = numpy.asarray(t, dtype=float)[..., None]
        n = numpy.arange(1, len(self._Cs) + 1, dtype=float)
        HNn = self._H * self._N * n
        K = n * self._N / R / self._sin_alpha
        B = HNn / R / self._sin_alpha * (0.4 * HNn / R / self._sin_alpha + 1.0)
        D = (0.3 * HNn**2 / self._sin_alpha / R + HNn + R * self._sin_alpha) / (0.3 * HNn + R * self._sin_alpha)
        g = self._N * (phi - self._omega * t - self._phi_ref - numpy.log(R / self._r_ref) / self._tan_alpha)
        pre = -self._H * numpy.exp(-(R[..., 0] - self._r_ref) / self._Rs)
        return R, z, n, K, B, D, g, pre

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        R, z, n, K, B, D, g, pre = self._terms(R, z, phi, t)
        return pre * numpy.sum(self._Cs / K / D * numpy.cos(n * g) / numpy.cosh(K * z / B) ** B, axis=-1)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        R, z, n, K, B, D, g, pre = self._terms(R, z, phi, t)
        u = K * z / B
        return pre * numpy.sum(self._Cs / D * numpy.cos(n * g) * numpy.tanh(u) / numpy.cosh(u) ** B, axis=-1)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        R, z, n, K, B, D, g, pre = self._terms(R, z, phi, t)
        return pre * numpy.sum(self._N * n * self._Cs / D / K / numpy.cosh(K * z / B) ** B * numpy.sin(n * g),
                               axis=-1)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        R, z, n, K, B, D, g, pre = self._terms(R, z, phi, t)
        HNn = self._H * self._N * n
        k = HNn / R / self._sin_alpha
        dg = -self._N / R / self._tan_alpha
        dK = -n * self._N / R**2 / self._sin_alpha
        dB = -HNn / R**3 / self._sin_alpha**2 * (0.8 * HNn + R * self._sin_alpha)
        dD = k * (0.3 * (k + 0.3 * k**2 + 1.0) / R / (0.3 * k + 1.0) ** 2 - (1.0 + 0.6 * k) / R / (0.3 * k + 1.0))
        u = z * K / B
        sech = 1.0 / numpy.cosh(u)
        cg, sg = numpy.cos(n * g), numpy.sin(n * g)
        terms = self._Cs * sech**B / D * ((n * dg / K * sg + cg * (z * numpy.tanh(u) * (dK / K - dB / B)
                                                                 - dB / K * numpy.log(sech) + dK / K**2
                                                                 + dD / D / K)) + cg / K / self._Rs)
        return pre * numpy.sum(terms, axis=-1)

    def c_args(self):
        return [float(len(self._Cs)), self._amp, self._N, self._sin_alpha, self._tan_alpha, self._r_ref,
                self._phi_ref, self._Rs, self._H, self._omega] + list(self._Cs)


def _de_psi(t):
    return t * numpy.tanh(numpy.pi / 2.0 * numpy.sinh(t))


def _de_psiprime(t):
    return (numpy.sinh(numpy.pi * numpy.sinh(t)) + numpy.pi * t * numpy.cosh(t)) / (
        numpy.cosh(numpy.pi * numpy.sinh(t)) + 1.0)


class DoubleExponentialDiskPotential(Potential):
    """rho = amp exp(-R/hr - |z|/hz); pot_type 11, args [amp, -4 pi alpha
    amp, alpha, beta, de_n, J0 nodes, J1 nodes, J0 weights, J1 weights]
    with alpha = 1/hr, beta = 1/hz (DoubleExponentialDiskPotential.py:20-140,
    integrateFullOrbit.py:93-107).  The disk's Hankel integrals are summed by
    Ogata's (2005) double-exponential formula for Bessel integrals: nodes
    pi/h psi(h xi_k) at the Bessel zeros xi_k pi, psi(t) = t tanh(pi/2
    sinh t), as galpy builds them (weights past the overflow of psi' are NaN
    there too, and the sums stop well before them)."""

    pot_code = 11

    def __init__(self, amp=1.0, hr=1.0 / 3.0, hz=1.0 / 16.0, normalize=False, de_h=1e-3, de_n=10000):
        Potential.__init__(self, amp=amp)
        self._hr, self._hz = hr, hz
        self._alpha = 1.0 / hr
        self._beta = 1.0 / hz
        self._de_h, self._de_n = de_h, de_n
        with numpy.errstate(over="ignore", invalid="ignore"):
            j0z = special.jn_zeros(0, de_n) / numpy.pi
            j1z = special.jn_zeros(1, de_n) / numpy.pi
            self._de_j0_xs = numpy.pi / de_h * _de_psi(de_h * j0z)
            self._de_j0_weights = (2.0 / (numpy.pi * j0z * special.j1(numpy.pi * j0z) ** 2.0)
                                   * special.j0(self._de_j0_xs) * _de_psiprime(de_h * j0z))
            self._de_j1_xs = numpy.pi / de_h * _de_psi(de_h * j1z)
            self._de_j1_weights = (2.0 / (numpy.pi * j1z * special.jv(2, numpy.pi * j1z) ** 2.0)
                                   * special.j1(self._de_j1_xs) * _de_psiprime(de_h * j1z))
        _set_normalize(self, normalize)

    def _sum(self, R, z, xs, ws, fun):
        shape = numpy.broadcast(numpy.asarray(R), numpy.asarray(z)).shape
        R = numpy.broadcast_to(numpy.asarray(R, dtype=float), shape).reshape(-1, 1)
        z = numpy.broadcast_to(numpy.asarray(z, dtype=float), shape).reshape(-1, 1)
        with numpy.errstate(over="ignore", invalid="ignore"):
            out = numpy.nansum(fun(xs[None, :] / R, numpy.fabs(z)) * ws[None, :], axis=1)
        return out.reshape(shape) if shape else float(out[0])

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        a2, b = self._alpha**2.0, self._beta
        f = lambda x, fz: (a2 + x * x) ** -1.5 * (b * numpy.exp(-x * fz) - x * numpy.exp(-b * fz)) / (b * b - x * x)
        return -4.0 * numpy.pi * self._alpha * self._sum(R, z, self._de_j0_xs, self._de_j0_weights, f) / R

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        a2, b = self._alpha**2.0, self._beta
        f = lambda x, fz: x * (a2 + x * x) ** -1.5 * (b * numpy.exp(-x * fz) - x * numpy.exp(-b * fz)) / (b * b - x * x)
        return -4.0 * numpy.pi * self._alpha * self._sum(R, z, self._de_j1_xs, self._de_j1_weights, f) / R

    def _zforce(self, R, z, phi=0.0, t=0.0):
        a2, b = self._alpha**2.0, self._beta
        f = lambda x, fz: x * (a2 + x * x) ** -1.5 * (numpy.exp(-x * fz) - numpy.exp(-b * fz)) / (b * b - x * x)
        out = -4.0 * numpy.pi * self._alpha * b * self._sum(R, z, self._de_j0_xs, self._de_j0_weights, f) / R
        return numpy.where(numpy.asarray(z) > 0.0, out, -out) if numpy.ndim(out) else (out if z > 0.0 else -out)

    def c_args(self):
        return ([self._amp, -4.0 * numpy.pi * self._alpha * self._amp, self._alpha, self._beta,
                 float(len(self._de_j1_xs))] + list(self._de_j0_xs) + list(self._de_j1_xs)
                + list(self._de_j0_weights) + list(self._de_j1_weights))


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


def _set_normalize(pot, normalize):
    if normalize or (isinstance(normalize, (int, float)) and not isinstance(normalize, bool)):
        pot.normalize(normalize)


class _SphericalPotential(Potential):
    """Spherical components given by Phi(r) / amp (``_phi``) and
    dPhi/dr / amp (``_dphi``); forces are -Phi'(r) (R, z) / r."""

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._phi(numpy.sqrt(R**2.0 + z**2.0))

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R**2.0 + z**2.0)
        return -self._dphi(r) * R / r

    def _zforce(self, R, z, phi=0.0, t=0.0):
        r = numpy.sqrt(R**2.0 + z**2.0)
        return -self._dphi(r) * z / r


class PowerSphericalPotential(_SphericalPotential):
    """rho ~ r^-alpha; pot_type 7, args [amp, alpha].  amp is the mass
    within r1 unless alpha = 3 (PowerSphericalPotential.py:25-60)."""

    pot_code = 7

    def __init__(self, amp=1.0, alpha=1.0, normalize=False, r1=1.0):
        Potential.__init__(self, amp=amp)
        self.alpha = alpha
        if self.alpha != 3.0:
            self._amp *= r1 ** (self.alpha - 3.0) * 4.0 * numpy.pi / (3.0 - self.alpha)
        _set_normalize(self, normalize)

    def _phi(self, r):
        if self.alpha == 2.0:
            return numpy.log(r)
        return -(r ** (2.0 - self.alpha)) / (self.alpha - 2.0)

    def _dphi(self, r):
        return r ** (1.0 - self.alpha)

    def c_args(self):
        return [self._amp, self.alpha]


class KeplerPotential(PowerSphericalPotential):
    """Phi = -amp / r: PowerSphericalPotential with alpha = 3
    (PowerSphericalPotential.py:400-440), packed as pot_type 7."""

    def __init__(self, amp=1.0, normalize=False):
        PowerSphericalPotential.__init__(self, amp=amp, alpha=3.0, normalize=normalize)


class JaffePotential(_SphericalPotential):
    """Phi = -amp log(1 + a/r) / a; pot_type 10, args [amp, a]."""

    pot_code = 10

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        _set_normalize(self, normalize)

    def _phi(self, r):
        return -numpy.log(1.0 + self.a / r) / self.a

    def _dphi(self, r):
        return 1.0 / (r * (r + self.a))

    def c_args(self):
        return [self._amp, self.a]


class IsochronePotential(_SphericalPotential):
    """Phi = -amp / (b + sqrt(r^2 + b^2)); pot_type 14, args [amp, b]."""

    pot_code = 14

    def __init__(self, amp=1.0, b=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self.b = b
        _set_normalize(self, normalize)

    def _phi(self, r):
        return -1.0 / (self.b + numpy.sqrt(r**2.0 + self.b**2.0))

    def _dphi(self, r):
        rb = numpy.sqrt(r**2.0 + self.b**2.0)
        return r / rb / (self.b + rb) ** 2.0

    def c_args(self):
        return [self._amp, self.b]


class PlummerPotential(_SphericalPotential):
    """Phi = -amp / sqrt(r^2 + b^2); pot_type 17, args [amp, b]."""

    pot_code = 17

    def __init__(self, amp=1.0, b=0.8, normalize=False):
        Potential.__init__(self, amp=amp)
        self._b = b
        self._b2 = b**2.0
        _set_normalize(self, normalize)

    def _phi(self, r):
        return -1.0 / numpy.sqrt(r**2.0 + self._b2)

    def _dphi(self, r):
        return r * (r**2.0 + self._b2) ** -1.5

    def c_args(self):
        return [self._amp, self._b]


class PseudoIsothermalPotential(_SphericalPotential):
    """rho ~ 1 / (1 + r^2/a^2); pot_type 18, args [amp, a]."""

    pot_code = 18

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self._a = a
        self._a2 = a**2.0
        _set_normalize(self, normalize)

    def _phi(self, r):
        r = numpy.asarray(r, dtype=float)
        with numpy.errstate(divide="ignore", invalid="ignore"):
            out = (0.5 * numpy.log(1.0 + r**2.0 / self._a2) + self._a / r * numpy.arctan(r / self._a)) / self._a
        return numpy.where(r == 0.0, 1.0 / self._a, out)

    def _dphi(self, r):
        return (1.0 / r - self._a / r**2.0 * numpy.arctan(r / self._a)) / self._a

    def c_args(self):
        return [self._amp, self._a]


class KuzminDiskPotential(Potential):
    """Phi = -amp / sqrt(R^2 + (a + |z|)^2); pot_type 19, args [amp, a]."""

    pot_code = 19

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self._a = a
        _set_normalize(self, normalize)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return -1.0 / numpy.sqrt(R**2.0 + (self._a + numpy.fabs(z)) ** 2.0)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return -R * (R**2.0 + (self._a + numpy.fabs(z)) ** 2.0) ** -1.5

    def _zforce(self, R, z, phi=0.0, t=0.0):
        az = self._a + numpy.fabs(z)
        return -numpy.sign(z) * az * (R**2.0 + az**2.0) ** -1.5

    def c_args(self):
        return [self._amp, self._a]


class BurkertPotential(_SphericalPotential):
    """rho = amp / ((1 + r/a)(1 + (r/a)^2)); pot_type 20, args [amp, a]."""

    pot_code = 20

    def __init__(self, amp=1.0, a=2.0, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        _set_normalize(self, normalize)

    def _phi(self, r):
        x = r / self.a
        return -(self.a**2.0) * numpy.pi / x * (
            -numpy.pi + 2.0 * (1.0 + x) * numpy.arctan(1.0 / x) + 2.0 * (1.0 + x) * numpy.log(1.0 + x)
            + (1.0 - x) * numpy.log(1.0 + x**2.0)
        )

    def _dphi(self, r):
        # enclosed mass / r^2
        x = r / self.a
        return (numpy.pi * self.a**3.0 / r**2.0
                * (2.0 * numpy.log(1.0 + x) + numpy.log(1.0 + x**2.0) - 2.0 * numpy.arctan(x)))

    def c_args(self):
        return [self._amp, self.a]


class DehnenSphericalPotential(_SphericalPotential):
    """Dehnen (1993) profile with inner slope alpha; pot_type 34, args
    [amp, a, alpha] (TwoPowerSphericalPotential.py:301-355)."""

    pot_code = 34

    def __init__(self, amp=1.0, a=1.0, alpha=1.5, normalize=False):
        if (alpha < 0.0) or (alpha >= 3.0):
            raise OSError("DehnenSphericalPotential requires 0 <= alpha < 3")
        Potential.__init__(self, amp=amp)
        self.a = a
        self.alpha = alpha
        _set_normalize(self, normalize)

    def _phi(self, r):
        if self.alpha == 2.0:  # the Jaffe limit
            return -numpy.log(1.0 + self.a / r) / self.a
        return -(1.0 - (r / (r + self.a)) ** (2.0 - self.alpha)) / (
            self.a * (2.0 - self.alpha) * (3.0 - self.alpha))

    def _dphi(self, r):
        return r ** (1.0 - self.alpha) * (r + self.a) ** (self.alpha - 3.0) / (3.0 - self.alpha)

    def c_args(self):
        return [self._amp, self.a, self.alpha]


class DehnenCoreSphericalPotential(DehnenSphericalPotential):
    """The cored (alpha = 0) Dehnen profile; pot_type 33, args [amp, a]
    (TwoPowerSphericalPotential.py:436-470)."""

    pot_code = 33

    def __init__(self, amp=1.0, a=1.0, normalize=False):
        DehnenSphericalPotential.__init__(self, amp=amp, a=a, alpha=0.0, normalize=normalize)

    def c_args(self):
        return [self._amp, self.a]


class HomogeneousSpherePotential(_SphericalPotential):
    """Uniform-density sphere of radius R; pot_type 35, args [amp, R^2, R^3]."""

    pot_code = 35

    def __init__(self, amp=1.0, R=1.1, normalize=False):
        Potential.__init__(self, amp=amp)
        self.R = R
        self._R2 = R**2.0
        self._R3 = R**3.0
        _set_normalize(self, normalize)

    def _phi(self, r):
        r = numpy.asarray(r, dtype=float)
        return numpy.where(r**2.0 < self._R2, r**2.0 - 3.0 * self._R2, -2.0 * self._R3 / numpy.maximum(r, 1e-300))

    def _dphi(self, r):
        r = numpy.asarray(r, dtype=float)
        return numpy.where(r**2.0 < self._R2, 2.0 * r, 2.0 * self._R3 / numpy.maximum(r, 1e-300) ** 2.0)

    def c_args(self):
        return [self._amp, self._R2, self._R3]


class EinastoPotential(_SphericalPotential):
    """rho = amp exp(-(r/h)^(1/n)); pot_type 41, args [amp, h, n]
    (EinastoPotential.py:20-140).  rs (the scale radius, with amp the
    density there) or rm2 (the radius where the logarithmic slope is -2)
    may be given instead of h."""

    pot_code = 41

    def __init__(self, amp=1.0, h=2.0, n=1, rs=None, rm2=None, normalize=False):
        Potential.__init__(self, amp=amp)
        if rs is not None:
            from scipy.optimize import fsolve

            est = (3 * n - 1 / 3 + 8 / (1215 * n) + 184 / (229635 * n**2) + 1048 / (31000725 * n**3)
                   - 17557576 / (1242974068875 * n**4))
            dn = fsolve(lambda x: 2 * special.gammaincc(3 * n, x) * special.gamma(3 * n)
                        - special.gamma(3 * n), est)[0]
            self._amp = amp * numpy.e**dn
            h = rs / (dn**n)
        elif rm2 is not None:
            self._amp = amp * numpy.e ** (2 * n)
            h = rm2 / ((2 * n) ** n)
        self.h = h
        self.n = n
        _set_normalize(self, normalize)

    def _phi(self, r):
        s = numpy.asarray(r, dtype=float) / self.h
        x = s ** (1.0 / self.n)
        g3 = special.gamma(3 * self.n)
        return -4.0 * numpy.pi * self.h**2 * self.n * (
            g3 * special.gammainc(3 * self.n, x) / s
            + special.gamma(2 * self.n) * special.gammaincc(2 * self.n, x))

    def _dphi(self, r):
        s = numpy.asarray(r, dtype=float) / self.h
        return (4.0 * numpy.pi * self.h * self.n * special.gamma(3 * self.n) * s**-2.0
                * special.gammainc(3 * self.n, s ** (1.0 / self.n)))

    def c_args(self):
        return [self._amp, self.h, self.n]


class ExpTruncNFWPotential(_SphericalPotential):
    """NFW density times exp(-r/rc): rho = amp / (4 pi a^3) exp(-r/rc) /
    ((r/a) (1 + r/a)^2); pot_type 46, args [amp, a, rc]
    (ExpTruncNFWPotential.py:20-340).  mass sets amp from the total mass."""

    pot_code = 46

    def __init__(self, amp=1.0, a=1.0, rc=2.0, mass=None, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        self.rc = rc
        self._alpha = a / rc
        self._Ftot = numpy.exp(self._alpha) * (1.0 + self._alpha) * special.exp1(self._alpha) - 1.0
        if mass is not None:
            self._amp = mass / self._Ftot
        _set_normalize(self, normalize)

    def _F(self, r):
        r = numpy.asarray(r, dtype=float)
        beta = (self.a + r) / self.rc
        return (numpy.exp(self._alpha) * (1.0 + self._alpha) * (special.exp1(self._alpha) - special.exp1(beta))
                - 1.0 + self.a * numpy.exp(-r / self.rc) / (self.a + r))

    def _phi(self, r):
        r = numpy.asarray(r, dtype=float)
        G = (numpy.exp(-r / self.rc) / (self.a + r)
             - numpy.exp(self._alpha) * special.exp1((self.a + r) / self.rc) / self.rc)
        return -(self._F(r) / r + G)

    def _dphi(self, r):
        r = numpy.asarray(r, dtype=float)
        return self._F(r) / r**2.0

    def c_args(self):
        return [self._amp, self.a, self.rc]


class TwoPowerSphericalPotential(_SphericalPotential):
    """rho = amp / (4 pi a^3) / ((r/a)^alpha (1 + r/a)^(beta - alpha));
    pot_type 42, args [amp, a, alpha, beta] (TwoPowerSphericalPotential.py:
    31-180; galpy packs this code for any alpha, beta, the Hernquist, Jaffe
    and NFW classes having codes of their own)."""

    pot_code = 42

    def __init__(self, amp=1.0, a=5.0, alpha=1.5, beta=3.5, normalize=False):
        Potential.__init__(self, amp=amp)
        self.a = a
        self.alpha = alpha
        self.beta = beta
        _set_normalize(self, normalize)

    def _dphi(self, r):
        r = numpy.asarray(r, dtype=float)
        return (r ** (1.0 - self.alpha) * self.a ** (self.alpha - 3.0) / (3.0 - self.alpha)
                * special.hyp2f1(3.0 - self.alpha, self.beta - self.alpha, 4.0 - self.alpha, -r / self.a))

    def _phi(self, r):
        r = numpy.asarray(r, dtype=float)
        a, al, be = self.a, self.alpha, self.beta
        if be == 3.0:
            return (1.0 / a * (1.0 - (r / a) ** (2.0 - al) / (3.0 - al)
                               * special.hyp2f1(3.0 - al, 2.0 - al, 4.0 - al, -r / a)) / (al - 2.0))
        r = r + 1e-11
        return (special.gamma(be - 3.0) * ((r / a) ** (3.0 - be) / special.gamma(be - 1.0)
                                          * special.hyp2f1(be - 3.0, be - al, be - 1.0, -a / r)
                                          - special.gamma(3.0 - al) / special.gamma(be - al)) / r)

    def c_args(self):
        return [self._amp, self.a, self.alpha, self.beta]


class interpSphericalPotential(_SphericalPotential):
    """A spherical potential from its tabulated radial force (interpSphericalPotential.py
    :15-130); pot_type 36, args [nr, rgrid, rforce on rgrid, amp, rmin, rmax,
    M(<rmax), Phi0, Phimax] (integrateFullOrbit.py:366-373).  ``rforce`` is a
    function of r or a potential (then Phi0 is its value at rgrid[0]).  As in
    galpy, the host evaluation interpolates with scipy's spline through the
    grid, while the C (and GPU) evaluation uses GSL's natural cubic spline of
    the same table; beyond rgrid[-1] the mass inside it acts as a point
    mass, inside rgrid[0] the force is 0."""

    pot_code = 36

    def __init__(self, rforce=None, rgrid=numpy.geomspace(0.01, 20, 101), Phi0=None):
        Potential.__init__(self, amp=1.0)
        self._rgrid = numpy.array(rgrid, dtype=float)
        if isinstance(rforce, (Potential, list, tuple)):
            _rforce = lambda r: evaluateRforces(rforce, r, 0.0)
            Phi0 = evaluatePotentials(rforce, self._rgrid[0], 0.0)
        else:
            _rforce = rforce
            Phi0 = 0.0 if Phi0 is None else Phi0
        self._rforce_grid = numpy.array([float(_rforce(r)) for r in self._rgrid])
        self._force_spline = interpolate.InterpolatedUnivariateSpline(self._rgrid, self._rforce_grid, k=3, ext=0)
        self._pot_spline = self._force_spline.antiderivative()
        self._Phi0 = Phi0 + self._pot_spline(self._rgrid[0])
        self._rmin = self._rgrid[0]
        self._rmax = self._rgrid[-1]
        self._total_mass = -(self._rmax**2.0) * self._force_spline(self._rmax)
        self._Phimax = -self._pot_spline(self._rmax) + self._Phi0 + self._total_mass / self._rmax

    def _dphi(self, r):
        r = numpy.asarray(r, dtype=float)
        return numpy.where(r >= self._rmax, self._total_mass / r**2.0, -self._force_spline(numpy.minimum(r, self._rmax)))

    def _phi(self, r):
        r = numpy.asarray(r, dtype=float)
        return numpy.where(r >= self._rmax, -self._total_mass / r + self._Phimax,
                           -self._pot_spline(numpy.minimum(r, self._rmax)) + self._Phi0)

    def c_args(self):
        return ([float(len(self._rgrid))] + list(self._rgrid) + list(self._rforce_grid)
                + [self._amp, self._rmin, self._rmax, float(self._total_mass), float(self._Phi0),
                   float(self._Phimax)])


def sph_harm_normalization(L, M):
    """beta_lm = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) (2 - delta_m0) for
    0 <= l < L, 0 <= m < M, zero where m > l (galpy/util/special.py:122-160)."""
    l = numpy.arange(L)[:, None]
    m = numpy.arange(M)[None, :]
    with numpy.errstate(invalid="ignore"):
        ln = 0.5 * (numpy.log(2 * l + 1.0) - numpy.log(4.0 * numpy.pi) + special.gammaln(l - m + 1.0)
                    - special.gammaln(l + m + 1.0))
    out = numpy.exp(ln) * numpy.where(m == 0, 1.0, 2.0)
    out[m > l] = 0.0
    return out


class SCFPotential(Potential):
    """Hernquist & Ostriker (1992) basis-function expansion (SCFPotential.py
    :20-330): Phi = sum_nlm N_lm P_l^m(cos theta) Phi_nl(r) (Acos cos m phi +
    Asin sin m phi), Phi_nl = -sqrt(4 pi) (a r)^l / (a + r)^(2l+1)
    C_n^(2l+3/2)((r - a)/(r + a)); amp is halved so that Acos = [[[1]]] is
    galpy's Hernquist potential of the same amp.  pot_type 24, args [a,
    isNonAxi, N, L, M, Nt, amp Acos N_lm (and Asin), 11 cache slots] or, with
    a time grid, tgrid and the coefficients' cubic splines in time as
    [Nt-1][NLM][4] blocks (integratePlanarOrbit.py:726-762).  ``Acos``/
    ``Asin``: (N, L, M) arrays, or with ``tgrid`` (Nt, N, L, M) arrays or
    callables of t."""

    pot_code = 24

    def __init__(self, amp=1.0, Acos=numpy.array([[[1.0]]]), Asin=None, a=1.0, tgrid=None, normalize=False):
        Potential.__init__(self, amp=amp / 2.0)
        self._a = a
        self._tdep = tgrid is not None
        if self._tdep:
            from scipy.interpolate import CubicSpline

            self._tgrid = numpy.array(tgrid, dtype=float)
            Ac = numpy.array([Acos(t) for t in self._tgrid]) if callable(Acos) else numpy.array(Acos, dtype=float)
            As = None if Asin is None else (numpy.array([Asin(t) for t in self._tgrid]) if callable(Asin)
                                            else numpy.array(Asin, dtype=float))
            Nt, N, L, M = Ac.shape
            self.isNonAxi = not (As is None or L == 1 or (numpy.all(Ac[..., 1:] == 0) and numpy.all(As == 0)))
            NN = sph_harm_normalization(L, M)
            self._Acos_all = Ac * NN[None, None]
            self._Asin_all = numpy.zeros_like(self._Acos_all) if As is None else As * NN[None, None]
            self._coeff_shape = (N, L, M)
            self._Acos_interp = CubicSpline(self._tgrid, self._Acos_all.reshape(Nt, -1))
            self._Asin_interp = CubicSpline(self._tgrid, self._Asin_all.reshape(Nt, -1))
        else:
            Ac = numpy.array(Acos, dtype=float)
            N, L, M = Ac.shape
            self.isNonAxi = not (Asin is None or L == 1 or (numpy.all(Ac[:, :, 1:] == 0)
                                                            and numpy.all(numpy.asarray(Asin) == 0)))
            NN = sph_harm_normalization(L, M)
            self._Acos = Ac * NN[None]
            self._Asin = numpy.zeros_like(self._Acos) if Asin is None else numpy.asarray(Asin, dtype=float) * NN[None]
            self._coeff_shape = (N, L, M)
        _set_normalize(self, normalize)

    def _coeffs(self, t):
        if not self._tdep:
            return self._Acos, self._Asin
        N, L, M = self._coeff_shape
        return (self._Acos_interp(float(t)).reshape(N, L, M), self._Asin_interp(float(t)).reshape(N, L, M))

    def _spher(self, R, z, phi, t):
        """Phi and its derivatives in (r, theta, phi) at scalar points."""
        a = self._a
        r = numpy.sqrt(R * R + z * z)
        th = numpy.arctan2(R, z)
        x = numpy.cos(th)
        xi = (r - a) / (r + a)
        Ac, As = self._coeffs(t)
        N, L, M = self._coeff_shape
        pot = dr = dth = dph = 0.0
        for l in range(L):
            base = -((a * r) ** l) / (a + r) ** (2 * l + 1)
            dbase = base * (l / r - (2 * l + 1) / (a + r))
            for m in range(min(l + 1, M) if self.isNonAxi else 1):
                P = special.lpmv(m, l, x)
                Pm1 = special.lpmv(m, l - 1, x) if l - 1 >= m else 0.0
                dPdx = (l * x * P - (l + m) * Pm1) / (x * x - 1.0) if l > 0 else 0.0
                c, sn = numpy.cos(m * phi), numpy.sin(m * phi)
                for n in range(N):
                    A = Ac[n, l, m] * c + As[n, l, m] * sn
                    dA = (-Ac[n, l, m] * sn + As[n, l, m] * c) * m
                    al = 2 * l + 1.5
                    Cn = special.eval_gegenbauer(n, al, xi)
                    dCn = 2 * al * special.eval_gegenbauer(n - 1, al + 1, xi) if n > 0 else 0.0
                    ph = base * Cn
                    dph_r = dbase * Cn + base * dCn * 2 * a / (a + r) ** 2
                    pot += A * P * ph
                    dr += A * P * dph_r
                    dth += A * dPdx * (-numpy.sin(th)) * ph
                    dph += dA * P * ph
        s4 = numpy.sqrt(4 * numpy.pi)
        return s4 * pot, s4 * dr, s4 * dth, s4 * dph, r, th

    def _vec(self, f, R, z, phi, t):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        out = numpy.array([f(*v) for v in zip(R.ravel(), z.ravel(), phi.ravel(), t.ravel())])
        return out.reshape(R.shape) if R.shape else float(out[0])

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._vec(lambda *v: self._spher(*v)[0], R, z, phi, t)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        def f(*v):
            _, dr, dth, _, r, th = self._spher(*v)
            return -(dr * numpy.sin(th) + dth * numpy.cos(th) / r)
        return self._vec(f, R, z, phi, t)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        def f(*v):
            _, dr, dth, _, r, th = self._spher(*v)
            return -(dr * numpy.cos(th) - dth * numpy.sin(th) / r)
        return self._vec(f, R, z, phi, t)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return self._vec(lambda *v: -self._spher(*v)[3], R, z, phi, t)

    def c_args(self):
        N, L, M = self._coeff_shape
        cache = [-1.0] + [0.0] * 10
        if self._tdep:
            out = [self._a, float(self.isNonAxi), N, L, M, float(len(self._tgrid))] + list(self._tgrid)
            out += list(self._amp * numpy.ascontiguousarray(self._Acos_interp.c.transpose(1, 2, 0)).ravel())
            if self.isNonAxi:
                out += list(self._amp * numpy.ascontiguousarray(self._Asin_interp.c.transpose(1, 2, 0)).ravel())
            return out + cache
        out = [self._a, float(self.isNonAxi), N, L, M, 0.0] + list(self._amp * self._Acos.flatten(order="C"))
        if self.isNonAxi:
            out += list(self._amp * self._Asin.flatten(order="C"))
        return out + cache


def scf_compute_coeffs_axi(dens, N, L, a=1.0, radial_order=None, costheta_order=None):
    """Expansion coefficients (N, L, 1) of an axisymmetric density dens(R, z)
    in SCFPotential's basis (the m = 0 case of galpy's scf_compute_coeffs,
    SCFPotential.py): product Gauss-Legendre quadrature over xi = (r - a)/(r
    + a) and cos(theta), the basis normalisation I_nl of Hernquist &
    Ostriker (1992)."""
    nxi = radial_order if radial_order is not None else max(N + 3 * L // 2 + 1, 20)
    nct = costheta_order if costheta_order is not None else max(L + 1, 20)
    xg, wg = numpy.polynomial.legendre.leggauss(nxi)
    cg, cw = numpy.polynomial.legendre.leggauss(nct)
    xi = xg[:, None]
    ct = cg[None, :]
    r = a * (1.0 + xi) / (1.0 - xi)
    R, z = r * numpy.sqrt(1.0 - ct**2), r * ct
    rho = dens(R, z)
    dV = (1.0 + xi) ** 2 * (1.0 - xi) ** -4.0
    n = numpy.arange(N)[:, None]
    l = numpy.arange(L)[None, :]
    out = numpy.zeros((N, L))
    for ll in range(L):
        Pl = special.eval_legendre(ll, ct)
        for nn in range(N):
            phi_nl = -(a**3) * (1.0 + xi) ** ll * (1.0 - xi) ** (ll + 1.0) * special.eval_gegenbauer(nn, 2 * ll + 1.5, xi)
            out[nn, ll] = 2.0 * numpy.pi * numpy.sum(wg[:, None] * cw[None, :] * rho * phi_nl * Pl * dV)
    K = 0.5 * n * (n + 4 * l + 3) + (l + 1) * (2 * l + 1)
    NN = numpy.exp(-(2 * l) * numpy.log(2.0))
    lnI = (-(8 * l + 6) * numpy.log(2.0) + special.gammaln(n + 4 * l + 3) - special.gammaln(n + 1)
           - numpy.log(n + 2 * l + 1.5) - 2 * special.gammaln(2 * l + 1.5))
    I = -K * 4.0 * numpy.pi * numpy.exp(lnI)
    return (2.0 / I * out * NN * numpy.sqrt(2 * l + 1.0))[:, :, None]


class DiskSCFPotential(Potential):
    """A disk-like density as an SCF expansion plus the Kuijken & Dubinski
    approximation pairs Sigma_i(r) H_i(z) (DiskSCFPotential.py and
    KuijkenDubinskiDiskExpansionPotential.py): the pairs carry the disk's
    thinness, the SCF expansion (N, L) the axisymmetric remainder of the
    density.  Sigma: dict (or list of dicts) {"type": "exp" or "expwhole",
    "amp", "h"[, "Rhole"]}; hz: {"type": "exp" or "sech2", "h"}.  galpy packs
    it as the SCF (pot_type 24, amp folded in) followed by one pot_type 26
    per pair, [nsigma_args, type, 4 pi amp, h[, Rhole], hz type, hz h]
    (integrateFullOrbit.py:323-336, integratePlanarOrbit.py:778-812), each
    counted in npot."""

    pot_code = None

    def __init__(self, amp=1.0, dens=None, Sigma=None, hz=None, N=10, L=10, a=1.0, radial_order=None,
                 costheta_order=None):
        Potential.__init__(self, amp=amp)
        self._Sigma_dict = list(Sigma) if isinstance(Sigma, (list, tuple)) else [Sigma]
        self._hz_dict = list(hz) if isinstance(hz, (list, tuple)) else [hz]
        if len(self._hz_dict) == 1 and len(self._Sigma_dict) > 1:
            self._hz_dict = self._hz_dict * len(self._Sigma_dict)
        self._dens = dens
        me = scf_compute_coeffs_axi(self._me_dens, N, L, a=a, radial_order=radial_order,
                                    costheta_order=costheta_order)
        self._me = SCFPotential(Acos=me, a=a)

    def _pair_terms(self, Sg, hz, r, z):
        amp, h = 4.0 * numpy.pi * Sg.get("amp", 1.0), Sg.get("h", 1.0 / 3.0)
        if Sg.get("type", "exp") == "exp" and "Rhole" not in Sg:
            S = amp * numpy.exp(-r / h)
            dS, d2S = -S / h, S / h**2
        else:
            Rh = Sg.get("Rhole", 0.5)
            S = amp * numpy.exp(-Rh / r - r / h)
            dS = S * (Rh / r**2 - 1.0 / h)
            d2S = S * ((Rh / r**2 - 1.0 / h) ** 2 - 2.0 * Rh / r**3)
        hh, fz = hz.get("h", 0.0375), numpy.fabs(z)
        if hz.get("type", "exp") == "exp":
            hzv = 0.5 * numpy.exp(-fz / hh) / hh
            H = 0.5 * (numpy.exp(-fz / hh) - 1.0 + fz / hh) * hh
            dH = 0.5 * numpy.sign(z) * (1.0 - numpy.exp(-fz / hh))
        else:
            ex = numpy.exp(-fz / hh)  # 1/cosh^2(z/2h) = 4 e^(-|z|/h) / (1 + e^(-|z|/h))^2
            hzv = ex / (1.0 + ex) ** 2 / hh
            H = hh * (numpy.log1p(numpy.exp(-fz / hh)) + 0.5 * fz / hh - numpy.log(2.0))
            dH = 0.5 * numpy.tanh(0.5 * z / hh)
        return S, dS, d2S, hzv, H, dH

    def _me_dens(self, R, z, phi=None):
        """The density left to the expansion: dens (of (R, z), or of (R, z,
        phi) when phi is given) minus the pairs' densities
        (DiskSCFPotential.c:221-231, per pair)."""
        r = numpy.sqrt(R * R + z * z)
        out = self._dens(R, z) if phi is None else self._dens(R, z, phi)
        for Sg, hz in zip(self._Sigma_dict, self._hz_dict):
            S, dS, d2S, hzv, H, dH = self._pair_terms(Sg, hz, r, z)
            out = out - (S * hzv + d2S * H + 2.0 / r * dS * (H + z * dH)) / (4.0 * numpy.pi)
        return out

    def _pairs(self, fun, R, z):
        R, z = numpy.asarray(R, dtype=float), numpy.asarray(z, dtype=float)
        r = numpy.sqrt(R * R + z * z)
        return sum(fun(*self._pair_terms(Sg, hz, r, z), R, z, r) for Sg, hz in zip(self._Sigma_dict, self._hz_dict))

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._me._amp * self._me._evaluate(R, z, phi, t) + self._pairs(lambda S, dS, d2S, hzv, H, dH, R, z, r: S * H, R, z)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._me._amp * self._me._Rforce(R, z, phi, t) + self._pairs(
            lambda S, dS, d2S, hzv, H, dH, R, z, r: -dS * H * R / r, R, z)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._me._amp * self._me._zforce(R, z, phi, t) + self._pairs(
            lambda S, dS, d2S, hzv, H, dH, R, z, r: -dS * H * z / r - S * dH, R, z)

    def pack(self):
        """(pot_type, pot_args, npot) in galpy's packing."""
        me_args = self._me.c_args()
        N, L, M = self._me._coeff_shape
        me_args = me_args[:6] + [v * self._amp for v in me_args[6:6 + N * L * M]] + me_args[6 + N * L * M:]
        pt, pa = self._pair_args()
        return [24] + pt, list(me_args) + pa, 1 + len(pt)

    def _pair_args(self):
        """pot_type 26 codes and arguments of the approximation pairs
        (integratePlanarOrbit.py:778-812)."""
        pt, pa = [], []
        for Sg, hz in zip(self._Sigma_dict, self._hz_dict):
            pt.append(26)
            famp = 4.0 * numpy.pi * Sg.get("amp", 1.0) * self._amp
            if Sg.get("type", "exp") == "exp" and "Rhole" not in Sg:
                pa.extend([3, 0, famp, Sg.get("h", 1.0 / 3.0)])
            else:
                pa.extend([4, 1, famp, Sg.get("h", 1.0 / 3.0), Sg.get("Rhole", 0.5)])
            pa.extend([0 if hz.get("type", "exp") == "exp" else 1, hz.get("h", 0.0375)])
        return pt, pa


def _quintic_hermite(v, d1, d2, x):
    """Power-basis coefficients (highest power first, [..., Nr-1, 6]) of the
    C2 piecewise quintic that takes the value, first and second derivative
    given at every node of ``x`` -- the polynomial scipy's
    BPoly.from_derivatives builds with three constraints per node
    (MultipoleExpansionPotential.py:1015-1079).  With D, D', D'' the end
    conditions left after the left node's Taylor terms, h the interval,
    a3 h^3 = 10 D - 4 D' h + D'' h^2 / 2, a4 h^4 = -15 D + 7 D' h - D'' h^2,
    a5 h^5 = 6 D - 3 D' h + D'' h^2 / 2."""
    h = numpy.diff(x)
    a0, a1, a2 = v[..., :-1], d1[..., :-1], 0.5 * d2[..., :-1]
    D = v[..., 1:] - (a0 + h * (a1 + h * a2))
    Dp = (d1[..., 1:] - (a1 + 2.0 * a2 * h)) * h
    Dpp = (d2[..., 1:] - 2.0 * a2) * h * h
    A = 10.0 * D - 4.0 * Dp + 0.5 * Dpp
    B = -15.0 * D + 7.0 * Dp - Dpp
    C = 6.0 * D - 3.0 * Dp + 0.5 * Dpp
    return numpy.stack([C / h**5, B / h**4, A / h**3, a2, a1, a0], axis=-1)


def _legendre_PdP(x, L, M):
    """P_l^m(x) (Condon-Shortley phase, as GSL's unnormalised arrays) and
    dP_l^m/dx for l < L, m < min(l + 1, M): arrays [..., L, M]."""
    x = numpy.asarray(x, dtype=float)
    P = numpy.zeros(x.shape + (L, M))
    dP = numpy.zeros(x.shape + (L, M))
    for m in range(min(L, M)):
        for l in range(m, L):
            P[..., l, m] = special.lpmv(m, l, x)
            if l > 0:
                prev = P[..., l - 1, m] if l > m else 0.0
                dP[..., l, m] = (l * x * P[..., l, m] - (l + m) * prev) / (x * x - 1.0)
    return P, dP


class MultipoleExpansionPotential(Potential):
    """Multipole expansion of a density on a radial grid
    (MultipoleExpansionPotential.py:44-2216): Phi = amp sum_lm P_l^m(cos theta)
    [R^c_lm(r) cos m phi + R^s_lm(r) sin m phi], R_lm = r^-(l+1) I_inner(r) +
    r^l I_outer(r), I_inner = -4 pi/(2l+1) int_0^r a^(l+2) rho_lm,
    I_outer = -4 pi/(2l+1) int_r^inf a^(1-l) rho_lm, stored as C2 quintic
    piecewise polynomials on the grid; the density's multipoles rho_lm carry
    the real spherical-harmonic normalisation beta_lm.  pot_type 44 with
    galpy's packing (MultipoleExpansionPotential.py:2101-2216).

    ``rho_cos_splines`` / ``rho_sin_splines``: [L][M] splines of r (callable
    as s(r) and s(r, 1), e.g. scipy's InterpolatedUnivariateSpline), or with
    ``tgrid`` callables f(r, t) returning the values on ``rgrid``; default a
    Hernquist monopole.  ``from_density`` builds them from a density."""

    pot_code = 44

    def __init__(self, amp=1.0, rho_cos_splines=None, rho_sin_splines=None,
                 rgrid=numpy.geomspace(1e-3, 30, 1_001), tgrid=None, normalize=False):
        from scipy.interpolate import CubicSpline, InterpolatedUnivariateSpline

        Potential.__init__(self, amp=amp)
        rg = numpy.asarray(rgrid, dtype=float)
        self._rgrid = rg
        self._tdep = tgrid is not None
        if rho_cos_splines is None:  # a Hernquist monopole (MultipoleExpansionPotential.py:311-324)
            b00 = sph_harm_normalization(1, 1)[0, 0]
            vals = b00 * b00 * 4.0 * numpy.pi / (2.0 * numpy.pi) / rg / (1.0 + rg) ** 3
            rho_cos_splines, rho_sin_splines = [[InterpolatedUnivariateSpline(rg, vals, k=3)]], None
        L, M = len(rho_cos_splines), len(rho_cos_splines[0])

        def values(splines, t=None):
            if splines is None:
                return numpy.zeros((L, M, rg.size))
            if t is None:
                return numpy.array([[numpy.asarray(s(rg), dtype=float) for s in row] for row in splines])
            return numpy.array([[numpy.asarray(f(rg, t), dtype=float) for f in row] for row in splines])

        if self._tdep:
            self._tgrid = numpy.asarray(tgrid, dtype=float)
            rc = numpy.array([values(rho_cos_splines, t) for t in self._tgrid])  # (Nt, L, M, Nr)
            rs = numpy.array([values(rho_sin_splines, t) for t in self._tgrid])
        else:
            rc, rs = values(rho_cos_splines)[None], values(rho_sin_splines)[None]
        # isNonAxi: some m > 0 multipole above 1e-12 of the largest m = 0 one
        # (MultipoleExpansionPotential.py:156-186, 228-252); else M = 1
        tol = 1e-12 * max(numpy.abs(rc[:, :, 0]).max(), 1e-16)
        self.isNonAxi = bool(M > 1 and (numpy.abs(rc[:, :, 1:]).max() > tol or numpy.abs(rs[:, :, 1:]).max() > tol))
        if not self.isNonAxi:
            M = 1
            rc, rs = rc[:, :, :1], rs[:, :, :1]
        self._L, self._M = L, M

        def deriv(arr, splines, l, m):
            # the density's radial derivative: the given spline's own when
            # static (s(r, 1)), else that of a cubic spline through the values
            if not self._tdep and splines is not None:
                try:
                    return numpy.asarray(splines[l][m](rg, 1), dtype=float)[None]
                except TypeError:
                    pass
            return numpy.array([InterpolatedUnivariateSpline(rg, a, k=3).derivative()(rg) for a in arr])

        blocks = []  # (l, m, trig) l-major, cos before sin: (Nt, Nr-1, 6) x 2, rho (Nt, Nr)
        for l in range(L):
            pref = -4.0 * numpy.pi / (2 * l + 1)
            for m in range(min(l + 1, M)):
                for arr, spl in ((rc[:, l, m], rho_cos_splines), (rs[:, l, m], rho_sin_splines))[: 1 + (m > 0)]:
                    drho = deriv(arr, spl, l, m)
                    f_in, f_out = rg ** (l + 2) * arr, rg ** (1.0 - l) * arr
                    I_in = numpy.array([InterpolatedUnivariateSpline(rg, f, k=3).antiderivative()(rg) for f in f_in])
                    I_in -= I_in[:, :1]
                    cum = numpy.array([InterpolatedUnivariateSpline(rg, f, k=3).antiderivative()(rg) for f in f_out])
                    I_out = cum[:, -1:] - cum
                    d2_in = (l + 2) * rg ** (l + 1) * arr + rg ** (l + 2) * drho
                    d2_out = -(1.0 - l) * rg ** (-1.0 * l) * arr - rg ** (1.0 - l) * drho
                    blocks.append((_quintic_hermite(pref * I_in, pref * f_in, pref * d2_in, rg),
                                   _quintic_hermite(pref * I_out, -pref * f_out, pref * d2_out, rg), arr))
        self._blocks = blocks
        if self._tdep:
            self._tblocks = [tuple(CubicSpline(self._tgrid, b.reshape(b.shape[0], -1)) for b in blk) for blk in blocks]
        _set_normalize(self, normalize)

    @staticmethod
    def _density_arity(dens):
        import inspect

        pars = [p for p in inspect.signature(dens).parameters.values()
                if p.name != "t" and p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD)]
        return len(pars), "t" in inspect.signature(dens).parameters

    @staticmethod
    def _rho_lm(dens, rgrid, L, M, nct, nphi):
        """The density's multipoles alpha_lm int rho P_l^m cos|sin(m phi)
        dOmega on the grid, (Nr, L, M) each: Gauss-Legendre in cos theta,
        the trapezoidal rule in phi (MultipoleExpansionPotential.py:726-842);
        ``dens(R, z, phi)`` vectorised."""
        beta = sph_harm_normalization(L, M)
        alpha = beta.copy()
        alpha[:, 1:] /= 2.0
        Nr = rgrid.size
        if L == 1 and M == 1:
            rc = numpy.zeros((Nr, 1, 1))
            rc[:, 0, 0] = alpha[0, 0] * 4.0 * numpy.pi * dens(rgrid, 0.0 * rgrid, 0.0 * rgrid)
            return rc, numpy.zeros_like(rc)
        ct, w = numpy.polynomial.legendre.leggauss(nct)
        st = numpy.sqrt(1.0 - ct**2)
        P, _ = _legendre_PdP(ct, L, M)  # (nct, L, M)
        if M == 1:
            Rr, zz = rgrid[:, None] * st[None], rgrid[:, None] * ct[None]
            d = dens(Rr, zz, 0.0 * Rr) * 2.0 * numpy.pi  # (Nr, nct)
            rc = numpy.einsum("rc,c,cl->rl", d, w, P[:, :, 0])[:, :, None] * alpha[None]
            return rc, numpy.zeros_like(rc)
        ph = numpy.linspace(0.0, 2.0 * numpy.pi, nphi, endpoint=False)
        Rr = rgrid[:, None, None] * st[None, :, None] * numpy.ones(nphi)[None, None]
        zz = rgrid[:, None, None] * ct[None, :, None] * numpy.ones(nphi)[None, None]
        d = numpy.asarray(dens(Rr, zz, numpy.broadcast_to(ph, Rr.shape)), dtype=float)  # (Nr, nct, nphi)
        cm, sm = numpy.cos(numpy.outer(ph, numpy.arange(M))), numpy.sin(numpy.outer(ph, numpy.arange(M)))
        dphi = 2.0 * numpy.pi / nphi
        rc = numpy.einsum("rcp,pm,c,clm->rlm", d, cm, w, P) * dphi * alpha[None]
        rs = numpy.einsum("rcp,pm,c,clm->rlm", d, sm, w, P) * dphi * alpha[None]
        return rc, rs

    @classmethod
    def from_density(cls, dens, L=6, rgrid=numpy.geomspace(1e-3, 30, 1_001), tgrid=None, symmetry=None,
                     costheta_order=None, phi_order=None, amp=1.0, normalize=False):
        """MultipoleExpansionPotential.py:326-492: the expansion of
        ``dens`` (a function of r, of (R, z) or of (R, z, phi), with a ``t``
        keyword when time-dependent, then ``tgrid`` is required) to degree
        L - 1; ``symmetry`` 'spherical' (L = M = 1), 'axisymmetric' (M = 1)
        or None (M = L)."""
        from scipy.interpolate import InterpolatedUnivariateSpline, make_interp_spline

        rgrid = numpy.asarray(rgrid, dtype=float)
        nargs, has_t = cls._density_arity(dens)
        if has_t and tgrid is None:
            raise ValueError("tgrid is required when the density function depends on time")

        def at(t):
            if nargs == 1:
                return lambda R, z, phi: dens(numpy.sqrt(R * R + z * z), **({"t": t} if has_t else {}))
            if nargs == 2:
                return lambda R, z, phi: dens(R, z, **({"t": t} if has_t else {}))
            return lambda R, z, phi: dens(R, z, phi, **({"t": t} if has_t else {}))

        if symmetry is not None and symmetry.startswith("spher"):
            L = M = 1
        elif symmetry is not None and symmetry.startswith("axi"):
            M = 1
        else:
            M = L
        nct = max(20, L + 1) if costheta_order is None else costheta_order
        nphi = max(20, 2 * L + 1) if phi_order is None else phi_order
        beta = sph_harm_normalization(L, M)
        if has_t:
            tgrid = numpy.asarray(tgrid, dtype=float)
            both = [cls._rho_lm(at(t), rgrid, L, M, nct, nphi) for t in tgrid]
            rc = numpy.array([b[0] for b in both]) * beta[None, None]  # (Nt, Nr, L, M)
            rs = numpy.array([b[1] for b in both]) * beta[None, None]

            def fn(arr, l, m):
                sp = make_interp_spline(tgrid, arr[:, :, l, m], k=3)
                return lambda r, t: sp(t)

            cos_f = [[fn(rc, l, m) for m in range(M)] for l in range(L)]
            sin_f = [[fn(rs, l, m) for m in range(M)] for l in range(L)]
            return cls(amp=amp, rho_cos_splines=cos_f, rho_sin_splines=sin_f, rgrid=rgrid, tgrid=tgrid,
                       normalize=normalize)
        rc, rs = cls._rho_lm(at(None), rgrid, L, M, nct, nphi)
        cos_s = [[InterpolatedUnivariateSpline(rgrid, beta[l, m] * rc[:, l, m], k=3) for m in range(M)]
                 for l in range(L)]
        sin_s = [[InterpolatedUnivariateSpline(rgrid, beta[l, m] * rs[:, l, m], k=3) for m in range(M)]
                 for l in range(L)]
        return cls(amp=amp, rho_cos_splines=cos_s, rho_sin_splines=sin_s, rgrid=rgrid, normalize=normalize)

    def _coeff_blocks(self, t):
        if not self._tdep:
            return [(ci[0], co[0]) for ci, co, _ in self._blocks]
        Nr = self._rgrid.size
        return [(si(t).reshape(Nr - 1, 6), so(t).reshape(Nr - 1, 6)) for si, so, _ in self._tblocks]

    def _radial(self, ci, co, r, l):
        """R_lm and dR_lm/dr at radii r (eval_radial_lm,
        MultipoleExpansionPotential.c:438-507)."""
        rg = self._rgrid
        rmin, rmax = rg[0], rg[-1]
        ir = numpy.clip(numpy.searchsorted(rg, r, side="right") - 1, 0, rg.size - 2)
        below, above = r < rmin, r > rmax
        ir = numpy.where(above, rg.size - 2, ir)
        d = numpy.where(above, rmax - rg[-2], r - rg[ir])
        hor = lambda c: ((((c[:, 0] * d + c[:, 1]) * d + c[:, 2]) * d + c[:, 3]) * d + c[:, 4]) * d + c[:, 5]
        dhor = lambda c: (((5 * c[:, 0] * d + 4 * c[:, 1]) * d + 3 * c[:, 2]) * d + 2 * c[:, 3]) * d + c[:, 4]
        Ii, Io = hor(ci[ir]), hor(co[ir])
        rl, rnl = r**l, r ** (-(l + 1.0))
        Rv = rnl * Ii + numpy.where(above, 0.0, rl * Io)
        dR = -(l + 1) * rnl / r * Ii + numpy.where(above, 0.0, rnl * dhor(ci[ir]) + l * rl / r * Io + rl * dhor(co[ir]))
        if numpy.any(below):
            rb = r[below]
            P0 = ci[0, 4] / rmin ** (l + 2)
            Iie = P0 / (l + 3) * rb ** (l + 3)
            extra = P0 * numpy.log(rmin / rb) if l == 2 else P0 / (2 - l) * (rmin ** (2.0 - l) - rb ** (2.0 - l))
            Ioe = co[0, 5] + extra
            a, b = rb ** (-(l + 1.0)), rb**l
            Rv[below] = a * Iie + b * Ioe
            dR[below] = -(l + 1) * a / rb * Iie + l * b / rb * Ioe
        return Rv, dR

    def _spher(self, R, z, phi, t):
        """(Phi, dPhi/dr, dPhi/dtheta, dPhi/dphi) / amp at points (one time)."""
        R, z, phi = (numpy.atleast_1d(numpy.asarray(v, dtype=float)) for v in numpy.broadcast_arrays(R, z, phi))
        r = numpy.sqrt(R * R + z * z)
        x, st = z / r, R / r
        L, M = self._L, self._M
        P, dP = _legendre_PdP(x, L, M)
        out = [numpy.zeros_like(r) for _ in range(4)]
        k = 0
        blocks = self._coeff_blocks(t)
        for l in range(L):
            for m in range(min(l + 1, M)):
                for trig in range(1 + (m > 0)):
                    Rv, dR = self._radial(*blocks[k], r, l)
                    k += 1
                    c = numpy.cos(m * phi) if trig == 0 else numpy.sin(m * phi)
                    dc = -m * numpy.sin(m * phi) if trig == 0 else m * numpy.cos(m * phi)
                    out[0] += P[:, l, m] * Rv * c
                    out[1] += P[:, l, m] * dR * c
                    out[2] += -st * dP[:, l, m] * Rv * c
                    out[3] += P[:, l, m] * Rv * dc
        return out

    def _at_times(self, fn, R, z, phi, t):
        R, z, phi, t = numpy.broadcast_arrays(*(numpy.asarray(v, dtype=float) for v in (R, z, phi, t)))
        if not self._tdep or t.size == 1 or numpy.all(t == t.flat[0]):
            out = fn(self._spher(R.ravel(), z.ravel(), phi.ravel(), float(t.flat[0])), R.ravel(), z.ravel())
            return out.reshape(R.shape) if R.ndim else out[0]
        res = numpy.empty(R.size)
        for tv in numpy.unique(t):
            sel = (t == tv).ravel()
            res[sel] = fn(self._spher(R.ravel()[sel], z.ravel()[sel], phi.ravel()[sel], tv), R.ravel()[sel],
                          z.ravel()[sel])
        return res.reshape(R.shape)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._at_times(lambda o, R, z: o[0], R, z, phi, t)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._at_times(lambda o, R, z: -(o[1] * R + o[2] * z / numpy.sqrt(R * R + z * z))
                              / numpy.sqrt(R * R + z * z), R, z, phi, t)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._at_times(lambda o, R, z: -(o[1] * z - o[2] * R / numpy.sqrt(R * R + z * z))
                              / numpy.sqrt(R * R + z * z), R, z, phi, t)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return self._at_times(lambda o, R, z: -o[3], R, z, phi, t)

    def c_args(self):
        rg = self._rgrid
        Nr = rg.size
        head = [float(Nr), float(self._L), float(self._M), float(int(self.isNonAxi))]
        if not self._tdep:
            parts = [numpy.array(head), rg, numpy.array([self._amp, 0.0])]
            for ci, co, rho in self._blocks:
                parts.extend([ci[0].ravel(), co[0].ravel(), rho[0]])
        else:
            parts = [numpy.array(head), rg, numpy.array([self._amp, float(self._tgrid.size)]), self._tgrid]
            for si, so, sr in self._tblocks:
                parts.extend(numpy.ascontiguousarray(sp.c.transpose(1, 2, 0)).ravel() for sp in (si, so, sr))
        return numpy.concatenate(parts)


class DiskMultipoleExpansionPotential(DiskSCFPotential):
    """A disk-like density as a multipole expansion plus the Kuijken &
    Dubinski approximation pairs Sigma_i(r) H_i(z)
    (DiskMultipoleExpansionPotential.py:12-127): the pairs carry the disk's
    thinness, the expansion (degree L - 1 on ``rgrid``) the remainder of the
    density (of (R, z), or of (R, z, phi) when not axisymmetric).  Sigma and
    hz as for DiskSCFPotential.  galpy packs the expansion (pot_type 44, amp
    folded in) followed by one pot_type 26 per pair
    (integrateFullOrbit.py:307-322)."""

    def __init__(self, amp=1.0, dens=None, Sigma=None, hz=None, L=10, rgrid=numpy.geomspace(1e-3, 30, 1001),
                 symmetry=None, costheta_order=None, phi_order=None):
        Potential.__init__(self, amp=amp)
        if dens is None:
            dens = lambda R, z: 13.5 * numpy.exp(-3.0 * R) * numpy.exp(-27.0 * numpy.fabs(z))
        Sigma = {"type": "exp", "h": 1.0 / 3.0, "amp": 1.0} if Sigma is None else Sigma
        hz = {"type": "exp", "h": 1.0 / 27.0} if hz is None else hz
        self._Sigma_dict = list(Sigma) if isinstance(Sigma, (list, tuple)) else [Sigma]
        self._hz_dict = list(hz) if isinstance(hz, (list, tuple)) else [hz]
        if len(self._hz_dict) == 1 and len(self._Sigma_dict) > 1:
            self._hz_dict = self._hz_dict * len(self._Sigma_dict)
        self._dens = dens
        self._dens_nargs = MultipoleExpansionPotential._density_arity(dens)[0]
        if symmetry is None and self._dens_nargs < 3:
            symmetry = "axisymmetric"
        self._me = MultipoleExpansionPotential.from_density(
            self._me_dens3, L=L, rgrid=rgrid, symmetry=symmetry, costheta_order=costheta_order,
            phi_order=phi_order)

    def _me_dens3(self, R, z, phi):
        return self._me_dens(R, z, phi if self._dens_nargs >= 3 else None)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return self._me._phitorque(R, z, phi, t)

    def pack(self):
        """(pot_type, pot_args, npot) in galpy's packing."""
        me_args = numpy.array(self._me.c_args())
        me_args[4 + int(me_args[0])] *= self._amp  # integratePlanarOrbit.py:765-775
        pt, pa = self._pair_args()
        return [44] + pt, list(me_args) + pa, 1 + len(pt)

class interpRZPotential(Potential):
    """An axisymmetric potential interpolated on an (R, z) grid
    (interpRZPotential.py:30-460) with cubic B-splines (mirror boundaries),
    as galpy's C evaluates it; pot_type 13, args [nR, nz, 0, 0, 0, the R
    grid (log R when logR), the z grid, the B-spline coefficients of the
    potential, R force and z force, amp, logR] (integrateFullOrbit.py:111-173;
    the second-derivative grids are not built here).  rgrid and zgrid are
    (start, stop, n) for numpy.linspace; z >= 0 (the potential is symmetric
    in z).  Full orbits only, as in galpy."""

    pot_code = 13

    def __init__(self, RZPot=None, rgrid=(numpy.log(0.01), numpy.log(20.0), 101), zgrid=(0.0, 1.0, 101),
                 logR=False):
        from scipy import ndimage

        Potential.__init__(self, amp=1.0)
        self._logR = bool(logR)
        self._rg = numpy.linspace(*rgrid)
        self._zgrid = numpy.linspace(*zgrid)
        self._rgrid = numpy.exp(self._rg) if self._logR else self._rg
        RR, ZZ = numpy.meshgrid(self._rgrid, self._zgrid, indexing="ij")
        grids = [evaluatePotentials(RZPot, RR, ZZ), evaluateRforces(RZPot, RR, ZZ), evaluatezforces(RZPot, RR, ZZ)]
        self._coeffs = [ndimage.spline_filter(numpy.asarray(g, dtype=float), order=3, mode="mirror") for g in grids]

    def _interp(self, k, R, z):
        from scipy import ndimage

        R, z = numpy.broadcast_arrays(numpy.asarray(R, dtype=float), numpy.asarray(z, dtype=float))
        x = numpy.where(R > 0.0, numpy.log(numpy.where(R > 0.0, R, 1.0)), -20.72326583694641) if self._logR else R
        x = numpy.clip(x, self._rg[0], self._rg[-1])
        y = numpy.clip(numpy.fabs(z), self._zgrid[0], self._zgrid[-1])
        ix = numpy.clip(numpy.searchsorted(self._rg, x, side="right") - 1, 0, len(self._rg) - 2)
        iy = numpy.clip(numpy.searchsorted(self._zgrid, y, side="right") - 1, 0, len(self._zgrid) - 2)
        xn = ix + (x - self._rg[ix]) / (self._rg[ix + 1] - self._rg[ix])
        yn = iy + (y - self._zgrid[iy]) / (self._zgrid[iy + 1] - self._zgrid[iy])
        out = ndimage.map_coordinates(self._coeffs[k], [xn.ravel(), yn.ravel()], order=3, mode="mirror",
                                      prefilter=False).reshape(R.shape)
        return out if out.shape else float(out)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return self._interp(0, R, z)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._interp(1, R, z)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        f = self._interp(2, R, z)
        return numpy.where(numpy.asarray(z) < 0.0, -f, f) if numpy.ndim(f) else (-f if z < 0.0 else f)

    def c_args(self):
        return ([float(len(self._rg)), float(len(self._zgrid)), 0.0, 0.0, 0.0] + list(self._rg) + list(self._zgrid)
                + [v for c in self._coeffs for v in c.flatten(order="C")] + [self._amp, float(self._logR)])

    def c_args_planar(self):
        raise NotImplementedError("interpRZPotential has no planar C implementation")


class FlattenedPowerPotential(Potential):
    """Phi = -amp / (alpha m^alpha), m^2 = R^2 + z^2/q^2 + core^2 (amp/2 log m^2
    at alpha = 0); pot_type 12, args [amp, alpha, q^2, core^2], planar
    [amp, alpha, core^2] (FlattenedPowerPotential.py:30-80,
    integratePlanarOrbit.py:183-186)."""

    pot_code = 12

    def __init__(self, amp=1.0, alpha=0.5, q=0.9, core=_CORE, normalize=False, r1=1.0):
        Potential.__init__(self, amp=amp)
        self.alpha = alpha
        self.q2 = q**2.0
        self.core2 = core**2.0
        self._amp *= r1**self.alpha
        _set_normalize(self, normalize)

    def _m2(self, R, z):
        return self.core2 + R**2.0 + z**2.0 / self.q2

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        if self.alpha == 0.0:
            return 0.5 * numpy.log(self._m2(R, z))
        return -self._m2(R, z) ** (-0.5 * self.alpha) / self.alpha

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return -self._m2(R, z) ** (-0.5 * self.alpha - 1.0) * R

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return -self._m2(R, z) ** (-0.5 * self.alpha - 1.0) * z / self.q2

    def c_args(self):
        return [self._amp, self.alpha, self.q2, self.core2]

    def c_args_planar(self):
        return [self._amp, self.alpha, self.core2]


class NullPotential(Potential):
    """No force; pot_type 40 without arguments.  The packing drops it unless
    every component is one (integrateFullOrbit.py:51-56)."""

    pot_code = 40

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def c_args(self):
        return []


def _zvec_rotation(zvec):
    """The matrix that turns the unit vector zvec into the z axis
    (galpy/util/__init__.py:169-195 with inv=True): the transpose of
    Rodrigues' rotation taking e_z to zvec about e_z x zvec; the identity when
    zvec is e_z and minus the identity when it is -e_z, as galpy sets them."""
    zvec = numpy.asarray(zvec, dtype=float)
    zvec = zvec / numpy.sqrt(numpy.sum(zvec**2.0))
    cth = zvec[2]
    if numpy.fabs(cth - 1.0) < 1e-10:
        return numpy.eye(3)
    if numpy.fabs(cth + 1.0) < 1e-10:
        return -numpy.eye(3)
    k = numpy.cross([0.0, 0.0, 1.0], zvec)
    k /= numpy.sqrt(numpy.sum(k**2.0))
    K = numpy.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    sth = numpy.sqrt(1.0 - cth**2.0)
    M = cth * numpy.eye(3) + sth * K + (1.0 - cth) * numpy.outer(k, k)
    return M.T


class EllipsoidalPotential(Potential):
    """Densities constant on ellipsoids m^2 = x^2 + y^2/b^2 + z^2/c^2 of a
    frame turned by pa about z and tilted so that zvec is its z axis
    (EllipsoidalPotential.py:39-135).  With s in (0, 1), tau = 1/s^2 - 1 and
    D(s) = sqrt((1 + (b^2-1) s^2)(1 + (c^2-1) s^2)), the potential and the
    forces are Gauss-Legendre sums over s (EllipsoidalPotential.py:376-444):
      Phi = 2 pi b c sum w psi(m_s) / D,  m_s^2 = x^2/(1+tau) + y^2/(b^2+tau)
            + z^2/(c^2+tau),
      F_x = -4 pi b c sum w rho(m_s) x / (1+tau) / D  (y, z alike),
    with psi(m) = -int_m^inf dm'^2 rho(m').  galpy's packing: amp, 15 cache
    slots, the number of psi arguments and those arguments, b^2, c^2,
    aligned, the 3x3 rotation, glorder, the nodes and the weights with
    -4 pi b c / D folded in (integrateFullOrbit.py:221-300)."""

    isNonAxi = False

    def __init__(self, amp=1.0, b=1.0, c=1.0, zvec=None, pa=None, glorder=50):
        Potential.__init__(self, amp=amp)
        self._b = float(b)
        self._c = float(c)
        self._b2 = self._b**2.0
        self._c2 = self._c**2.0
        if zvec is None and (pa is None or numpy.fabs(pa) < 1e-10):
            self._aligned = True
            self._rot = numpy.eye(3)
        else:
            self._aligned = False
            pa_rot = numpy.eye(3)
            if pa is not None:
                pa_rot = numpy.array([[numpy.cos(pa), numpy.sin(pa), 0.0],
                                      [-numpy.sin(pa), numpy.cos(pa), 0.0],
                                      [0.0, 0.0, 1.0]])
            zvec_rot = numpy.eye(3) if zvec is None else _zvec_rotation(zvec)
            self._rot = numpy.dot(pa_rot, zvec_rot)
        if glorder is None:
            raise NotImplementedError("glorder=None (adaptive quadrature) has no HIP kernel")
        self._glorder = int(glorder)
        glx, glw = numpy.polynomial.legendre.leggauss(self._glorder)
        self._glx = 0.5 * glx + 0.5
        self._glw = 0.5 * glw
        self.isNonAxi = (not self._aligned) or numpy.fabs(self._b - 1.0) > 1e-10

    def _frame(self, R, z, phi):
        x, y = R * numpy.cos(phi), R * numpy.sin(phi)
        if self._aligned:
            return x, y, z
        v = numpy.dot(self._rot, numpy.array([x, y, z], dtype=float))
        return v[0], v[1], v[2]

    def _nodes(self, x, y, z):
        s = self._glx
        tau = 1.0 / s**2.0 - 1.0
        D = numpy.sqrt((1.0 + (self._b2 - 1.0) * s**2.0) * (1.0 + (self._c2 - 1.0) * s**2.0))
        m = numpy.sqrt(x**2.0 / (1.0 + tau) + y**2.0 / (self._b2 + tau) + z**2.0 / (self._c2 + tau))
        return tau, D, m

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        x, y, zz = self._frame(float(R), float(z), float(phi))
        _, D, m = self._nodes(x, y, zz)
        return 2.0 * numpy.pi * self._b * self._c * numpy.sum(self._glw * self._psi(m) / D)

    def _forces_xyz(self, R, z, phi):
        x, y, zz = self._frame(float(R), float(z), float(phi))
        tau, D, m = self._nodes(x, y, zz)
        w = -4.0 * numpy.pi * self._b * self._c * self._glw * self._mdens(m) / D
        F = numpy.array([numpy.sum(w * x / (1.0 + tau)), numpy.sum(w * y / (self._b2 + tau)),
                         numpy.sum(w * zz / (self._c2 + tau))])
        return F if self._aligned else numpy.dot(self._rot.T, F)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        F = self._forces_xyz(R, z, phi)
        return numpy.cos(phi) * F[0] + numpy.sin(phi) * F[1]

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._forces_xyz(R, z, phi)[2]

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        F = self._forces_xyz(R, z, phi)
        return R * (-numpy.sin(phi) * F[0] + numpy.cos(phi) * F[1])

    def _psi_args(self):
        raise NotImplementedError

    def c_args(self):
        s = self._glx
        D = numpy.sqrt((1.0 + (self._b2 - 1.0) * s**2.0) * (1.0 + (self._c2 - 1.0) * s**2.0))
        pa = list(self._psi_args())
        return ([self._amp] + [0.0] * 15 + [float(len(pa))] + pa
                + [self._b2, self._c2, float(int(self._aligned))] + list(self._rot.flatten())
                + [float(self._glorder)] + list(self._glx)
                + list(-4.0 * numpy.pi * self._glw * self._b * self._c / D))


class TriaxialHernquistPotential(EllipsoidalPotential):
    """rho = amp / (4 pi a^3) / ((m/a) (1 + m/a)^3); pot_type 21, psi
    arguments [a, a^4] (TwoPowerTriaxialPotential.py:193-310)."""

    pot_code = 21

    def __init__(self, amp=1.0, a=2.0, normalize=False, b=1.0, c=1.0, zvec=None, pa=None,
                 glorder=50):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self.a = a
        self.a4 = a**4
        self._amp /= 4.0 * numpy.pi * a**3
        _set_normalize(self, normalize)

    def _psi(self, m):
        return -self.a4 / (m + self.a) ** 2.0

    def _mdens(self, m):
        return self.a4 / m / (m + self.a) ** 3.0

    def _psi_args(self):
        return [self.a, self.a4]


class TriaxialNFWPotential(EllipsoidalPotential):
    """rho = amp / (4 pi a^3) / ((m/a) (1 + m/a)^2); pot_type 22, psi
    arguments [a, a^3] (TwoPowerTriaxialPotential.py:429-575; the conc/mvir
    set-up is not offered)."""

    pot_code = 22

    def __init__(self, amp=1.0, a=2.0, b=1.0, c=1.0, zvec=None, pa=None, normalize=False,
                 glorder=50):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self.a = a
        self.a3 = a**3
        self._amp /= 4.0 * numpy.pi * self.a3
        _set_normalize(self, normalize)

    def _psi(self, m):
        return -2.0 * self.a3 / (m + self.a)

    def _mdens(self, m):
        return self.a3 / m / (m + self.a) ** 2.0

    def _psi_args(self):
        return [self.a, self.a3]


class TriaxialJaffePotential(EllipsoidalPotential):
    """rho = amp / (4 pi a^3) / ((m/a)^2 (1 + m/a)^2); pot_type 23, psi
    arguments [a, a^2] (TwoPowerTriaxialPotential.py:313-426)."""

    pot_code = 23

    def __init__(self, amp=1.0, a=2.0, b=1.0, c=1.0, zvec=None, pa=None, normalize=False,
                 glorder=50):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self.a = a
        self.a2 = a**2
        self._amp /= 4.0 * numpy.pi * self.a2 * self.a
        _set_normalize(self, normalize)

    def _psi(self, m):
        return 2.0 * self.a2 * (self.a / (m + self.a) + numpy.log(m / (m + self.a)))

    def _mdens(self, m):
        return self.a2**2.0 / m**2.0 / (m + self.a) ** 2.0

    def _psi_args(self):
        return [self.a, self.a2]


class TwoPowerTriaxialPotential(EllipsoidalPotential):
    """rho = amp / (4 pi a^3) / ((m/a)^alpha (1 + m/a)^(beta - alpha)),
    0 <= alpha < 3 < ... beta > 2; pot_type 43, psi arguments [a, alpha,
    beta, beta-alpha, 2-alpha, 3-alpha, psi_inf] with psi_inf =
    Gamma(beta-2) Gamma(3-alpha) / Gamma(beta-alpha) (0 at alpha = 2)
    (TwoPowerTriaxialPotential.py:20-190)."""

    pot_code = 43

    def __init__(self, amp=1.0, a=5.0, alpha=1.5, beta=3.5, b=1.0, c=1.0, zvec=None, pa=None,
                 glorder=50, normalize=False):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        if beta <= 2.0 or alpha >= 3.0:
            raise OSError("TwoPowerTriaxialPotential requires 0 <= alpha < 3 and beta > 2")
        self.a = a
        self.alpha = alpha
        self.beta = beta
        self.betaminusalpha = beta - alpha
        self.twominusalpha = 2.0 - alpha
        self.threeminusalpha = 3.0 - alpha
        self.psi_inf = 0.0
        if self.twominusalpha != 0.0:
            self.psi_inf = (special.gamma(beta - 2.0) * special.gamma(3.0 - alpha)
                            / special.gamma(self.betaminusalpha))
        self._amp /= 4.0 * numpy.pi * a**3
        _set_normalize(self, normalize)

    def _psi(self, m):
        a2 = self.a**2.0
        if self.twominusalpha == 0.0:
            aom = self.a / m
            bma = self.betaminusalpha
            return -2.0 * a2 * aom**bma / bma * special.hyp2f1(bma, bma, bma + 1.0, -aom)
        moa = m / self.a
        return -2.0 * a2 * (self.psi_inf - moa**self.twominusalpha / self.twominusalpha
                            * special.hyp2f1(self.twominusalpha, self.betaminusalpha,
                                             self.threeminusalpha, -moa))

    def _mdens(self, m):
        return (self.a / m) ** self.alpha / (1.0 + m / self.a) ** self.betaminusalpha

    def _psi_args(self):
        return [self.a, self.alpha, self.beta, self.betaminusalpha, self.twominusalpha,
                self.threeminusalpha, self.psi_inf if self.twominusalpha != 0.0 else 0.0]


class PerfectEllipsoidPotential(EllipsoidalPotential):
    """rho = amp a / (pi^2 b c) / (a^2 + m^2)^2; pot_type 30, psi argument
    [a^2] (PerfectEllipsoidPotential.py:16-110)."""

    pot_code = 30

    def __init__(self, amp=1.0, a=5.0, b=1.0, c=1.0, zvec=None, pa=None, glorder=50,
                 normalize=False):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self.a = a
        self.a2 = a**2
        self._amp *= a / (numpy.pi**2 * self._b * self._c)
        _set_normalize(self, normalize)

    def _psi(self, m):
        return -1.0 / (self.a2 + m**2.0)

    def _mdens(self, m):
        return 1.0 / (self.a2 + m**2.0) ** 2.0

    def _psi_args(self):
        return [self.a2]


class TriaxialGaussianPotential(EllipsoidalPotential):
    """rho = amp / ((2 pi)^1.5 sigma^3 b c) exp(-m^2 / (2 sigma^2));
    pot_type 37, psi argument [-2 sigma^2] (TriaxialGaussianPotential.py:17-115)."""

    pot_code = 37

    def __init__(self, amp=1.0, sigma=5.0, b=1.0, c=1.0, zvec=None, pa=None, glorder=50,
                 normalize=False):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self._sigma = sigma
        self._twosigma2 = 2.0 * sigma**2
        self._amp /= (2.0 * numpy.pi) ** 1.5 * sigma**3.0 * self._b * self._c
        _set_normalize(self, normalize)

    def _psi(self, m):
        return -self._twosigma2 * numpy.exp(-(m**2.0) / self._twosigma2)

    def _mdens(self, m):
        return numpy.exp(-(m**2) / self._twosigma2)

    def _psi_args(self):
        return [-self._twosigma2]


class PowerTriaxialPotential(EllipsoidalPotential):
    """rho = amp (3 - alpha) / (4 pi) m^-alpha (amp the mass within r1 in the
    spherical case); pot_type 38, psi argument [alpha]
    (PowerTriaxialPotential.py:17-115)."""

    pot_code = 38

    def __init__(self, amp=1.0, alpha=1.0, r1=1.0, b=1.0, c=1.0, zvec=None, pa=None, glorder=50,
                 normalize=False):
        EllipsoidalPotential.__init__(self, amp=amp, b=b, c=c, zvec=zvec, pa=pa, glorder=glorder)
        self.alpha = alpha
        if alpha != 3.0:
            self._amp *= r1 ** (alpha - 3.0) * 4.0 * numpy.pi / (3.0 - alpha)
        self._amp *= (3.0 - alpha) / 4.0 / numpy.pi
        _set_normalize(self, normalize)

    def _psi(self, m):
        return 2.0 / (2.0 - self.alpha) * m ** (2.0 - self.alpha)

    def _mdens(self, m):
        return m**-self.alpha

    def _psi_args(self):
        return [self.alpha]


class WrapperPotential(Potential):
    """A potential that modifies the potential it wraps (galpy's
    WrapperPotential.py:91-150): the wrapped components are evaluated at
    phi - offset(R, t) and scaled by amp * factor(t).  galpy packs the
    wrapper's code followed by the wrapped codes, and as arguments the number
    of wrapped components, their arguments, then the wrapper's own
    (integrateFullOrbit.py:396-448)."""

    def __init__(self, amp=1.0, pot=None):
        Potential.__init__(self, amp=amp)
        if pot is None:
            raise ValueError("a wrapper needs the potential it wraps")
        self._pot = pot
        self.isNonAxi = any(q.isNonAxi for q in flatten(pot))

    def _factor(self, t):
        return 1.0

    def _offset(self, R, t):
        """(offset, d offset / dR)"""
        return 0.0, 0.0

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        off, _ = self._offset(R, t)
        return self._factor(t) * evaluatePotentials(self._pot, R, z, phi - off, t)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        off, doff = self._offset(R, t)
        return self._factor(t) * (evaluateRforces(self._pot, R, z, phi - off, t)
                                  - evaluatephitorques(self._pot, R, z, phi - off, t) * doff)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        off, _ = self._offset(R, t)
        return self._factor(t) * evaluatezforces(self._pot, R, z, phi - off, t)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        off, _ = self._offset(R, t)
        return self._factor(t) * evaluatephitorques(self._pot, R, z, phi - off, t)

    def c_args(self):
        raise TypeError("wrappers are packed by parse_pot")


class DehnenSmoothWrapperPotential(WrapperPotential):
    """Grows (or, with decay, removes) the wrapped potential between tform
    and tsteady with Dehnen's quintic; pot_type -1, own arguments
    [amp, tform, tsteady, grow] (DehnenSmoothWrapperPotential.py:28-56)."""

    pot_code = -1

    def __init__(self, amp=1.0, pot=None, tform=-4.0, tsteady=None, decay=False):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._tform = tform
        self._tsteady = self._tform / 2.0 if tsteady is None else self._tform + tsteady
        self._grow = not decay

    def _factor(self, t):
        t = numpy.asarray(t, dtype=float)
        xi = 2.0 * (t - self._tform) / (self._tsteady - self._tform) - 1.0
        mid = 3.0 / 16.0 * xi**5.0 - 5.0 / 8.0 * xi**3.0 + 15.0 / 16.0 * xi + 0.5
        smooth = numpy.where(t < self._tform, 0.0, numpy.where(t < self._tsteady, mid, 1.0))
        return smooth if self._grow else 1.0 - smooth

    def wrapper_args(self):
        return [self._amp, self._tform, self._tsteady, float(int(self._grow))]


class SolidBodyRotationWrapperPotential(WrapperPotential):
    """Rotates the wrapped potential at pattern speed omega from angle pa;
    pot_type -2, own arguments [amp, omega, pa]."""

    pot_code = -2

    def __init__(self, amp=1.0, pot=None, omega=1.0, pa=0.0):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._omega = omega
        self._pa = pa

    def OmegaP(self):
        return self._omega

    def _offset(self, R, t):
        return self._omega * numpy.asarray(t, dtype=float) + self._pa, 0.0

    def wrapper_args(self):
        return [self._amp, self._omega, self._pa]


class CorotatingRotationWrapperPotential(WrapperPotential):
    """Winds the wrapped potential up with the rotation curve
    vpo R^beta: offset vpo R^(beta-1) (t - to) + pa; pot_type -4, own
    arguments [amp, vpo, beta, pa, to]."""

    pot_code = -4

    def __init__(self, amp=1.0, pot=None, vpo=1.0, beta=0.0, to=0.0, pa=0.0):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._vpo = vpo
        self._beta = beta
        self._pa = pa
        self._to = to

    def _offset(self, R, t):
        R = numpy.asarray(R, dtype=float)
        dt = numpy.asarray(t, dtype=float) - self._to
        off = self._vpo * R ** (self._beta - 1.0) * dt + self._pa
        doff = self._vpo * (self._beta - 1.0) * R ** (self._beta - 2.0) * dt
        return off, doff

    def wrapper_args(self):
        return [self._amp, self._vpo, self._beta, self._pa, self._to]


class GaussianAmplitudeWrapperPotential(WrapperPotential):
    """Scales the wrapped potential by exp(-(t - to)^2 / (2 sigma^2));
    pot_type -5, own arguments [amp, to, sigma^2]."""

    pot_code = -5

    def __init__(self, amp=1.0, pot=None, to=0.0, sigma=1.0):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._to = to
        self._sigma2 = sigma**2.0

    def _factor(self, t):
        return numpy.exp(-0.5 * (numpy.asarray(t, dtype=float) - self._to) ** 2.0 / self._sigma2)

    def wrapper_args(self):
        return [self._amp, self._to, self._sigma2]


class RotateAndTiltWrapperPotential(WrapperPotential):
    """Rotates and tilts the wrapped potential (RotateAndTiltWrapperPotential.py
    :20-210): it is evaluated at x' = rot x + offset, its Cartesian force
    turned back by rot^T; rot turns by galaxy_pa about z and takes zvec to
    the z axis, or is set from (inclination, sky_pa, galaxy_pa); pot_type -8,
    own arguments [amp, 7 cache slots, rot (row-major), rotSet, offsetSet,
    offset, 4 NaN and 6 zero cache slots] (integrateFullOrbit.py:546-568).
    Full orbits only, as in galpy."""

    pot_code = -8

    def __init__(self, amp=1.0, inclination=None, galaxy_pa=None, sky_pa=None, zvec=None, offset=None, pot=None):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        zvec, galaxy_pa = self._parse_inclination(inclination, sky_pa, zvec, galaxy_pa)
        self._offset_xyz = None if offset is None else numpy.array(offset, dtype=float)
        pa_rot = numpy.eye(3) if galaxy_pa is None else numpy.array(
            [[numpy.cos(galaxy_pa), numpy.sin(galaxy_pa), 0.0], [-numpy.sin(galaxy_pa), numpy.cos(galaxy_pa), 0.0],
             [0.0, 0.0, 1.0]])
        zvec_rot = numpy.eye(3) if zvec is None else _zvec_rotation(zvec)
        self._rot = numpy.dot(pa_rot, zvec_rot)
        self._norot = bool((self._rot == numpy.eye(3)).all())
        self.isNonAxi = True

    @staticmethod
    def _parse_inclination(inclination, sky_pa, zvec, galaxy_pa):
        """RotateAndTiltWrapperPotential.py:112-137."""
        if inclination is None:
            return zvec, galaxy_pa
        sky_pa = 0.0 if sky_pa is None else sky_pa
        zvec_rot = numpy.dot(
            numpy.array([[numpy.sin(sky_pa), numpy.cos(sky_pa), 0.0], [-numpy.cos(sky_pa), numpy.sin(sky_pa), 0.0],
                         [0.0, 0.0, 1.0]]),
            numpy.array([[1.0, 0.0, 0.0], [0.0, -numpy.cos(inclination), -numpy.sin(inclination)],
                         [0.0, -numpy.sin(inclination), numpy.cos(inclination)]]))
        zvec = numpy.dot(zvec_rot, numpy.array([0.0, 0.0, 1.0]))
        int_rot = _zvec_rotation(zvec).T  # the rotation taking e_z to zvec
        pa = numpy.dot(int_rot, numpy.dot(zvec_rot, [1.0, 0.0, 0.0]))
        return zvec, (0.0 if galaxy_pa is None else galaxy_pa) + numpy.arctan2(pa[1], pa[0])

    def _aligned(self, R, z, phi):
        R, z, phi = (numpy.asarray(v, dtype=float) for v in numpy.broadcast_arrays(R, z, phi))
        xyz = numpy.stack([R * numpy.cos(phi), R * numpy.sin(phi), z])
        if not self._norot:
            xyz = numpy.tensordot(self._rot, xyz, axes=1)
        if self._offset_xyz is not None:
            xyz = xyz + self._offset_xyz.reshape((3,) + (1,) * (xyz.ndim - 1))
        return xyz

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        x, y, zp = self._aligned(R, z, phi)
        return evaluatePotentials(self._pot, numpy.sqrt(x * x + y * y), zp, numpy.arctan2(y, x), t)

    def _cart_forces(self, R, z, phi, t):
        x, y, zp = self._aligned(R, z, phi)
        Rp, php = numpy.sqrt(x * x + y * y), numpy.arctan2(y, x)
        fr = evaluateRforces(self._pot, Rp, zp, php, t)
        fp = evaluatephitorques(self._pot, Rp, zp, php, t)
        F = numpy.stack([numpy.cos(php) * fr - numpy.sin(php) * fp / Rp,
                         numpy.sin(php) * fr + numpy.cos(php) * fp / Rp,
                         evaluatezforces(self._pot, Rp, zp, php, t) + 0.0 * Rp])
        return F if self._norot else numpy.tensordot(self._rot.T, F, axes=1)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        Fx, Fy, _ = self._cart_forces(R, z, phi, t)
        return numpy.cos(phi) * Fx + numpy.sin(phi) * Fy

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        Fx, Fy, _ = self._cart_forces(R, z, phi, t)
        return R * (-numpy.sin(phi) * Fx + numpy.cos(phi) * Fy)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._cart_forces(R, z, phi, t)[2]

    def wrapper_args(self):
        off = [0.0, 0.0, 0.0] if self._offset_xyz is None else list(self._offset_xyz)
        return ([self._amp] + [0.0] * 7 + list(self._rot.flatten()) + [float(not self._norot),
                float(self._offset_xyz is not None)] + off + [numpy.nan] * 4 + [0.0] * 6)


class KuzminLikeWrapperPotential(WrapperPotential):
    """Phi(R, z) = amp Phi_pot(xi, 0), xi = sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)
    for an axisymmetric wrapped potential (KuzminLikeWrapperPotential.py
    :20-120); pot_type -10, own arguments [amp, a, b^2]."""

    pot_code = -10

    def __init__(self, amp=1.0, a=1.1, b=0.0, pot=None):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        if any(q.isNonAxi for q in flatten(pot)):
            raise RuntimeError("KuzminLikeWrapperPotential only works for spherical or axisymmetric potentials")
        self._a = a
        self._b2 = b**2.0
        self.isNonAxi = False

    def _xi(self, R, z):
        return numpy.sqrt(R**2.0 + (self._a + numpy.sqrt(z**2.0 + self._b2)) ** 2.0)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return evaluatePotentials(self._pot, self._xi(R, z), 0.0, phi, t)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return evaluateRforces(self._pot, self._xi(R, z), 0.0, phi, t) * R / self._xi(R, z)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        s = numpy.sqrt(z**2.0 + self._b2)
        return evaluateRforces(self._pot, self._xi(R, z), 0.0, phi, t) * (self._a + s) * z / self._xi(R, z) / s

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def wrapper_args(self):
        return [self._amp, self._a, self._b2]


class CylindricallySeparablePotentialWrapper(WrapperPotential):
    """Phi(R, z) = amp [Phi_pot(R, 0) + Phi_pot(Rp, z) - Phi_pot(Rp, 0)]
    (CylindricallySeparablePotentialWrapper.py:15-160); pot_type -12, own
    arguments [amp, Rp, Phi_pot(Rp, 0)]."""

    pot_code = -12

    def __init__(self, amp=1.0, pot=None, Rp=1.0):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._Rp = Rp
        self._refpot = evaluatePotentials(self._pot, self._Rp, 0.0)
        self.isNonAxi = False

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        return evaluatePotentials(self._pot, R, 0.0) + evaluatePotentials(self._pot, self._Rp, z) - self._refpot

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return evaluateRforces(self._pot, R, 0.0) + 0.0 * numpy.asarray(z, dtype=float)

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return evaluatezforces(self._pot, self._Rp, z) + 0.0 * numpy.asarray(R, dtype=float)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def wrapper_args(self):
        return [self._amp, self._Rp, self._refpot]


def _on_grid(f, tgrid):
    """A time function on a grid: one vectorised call when it takes arrays,
    else a loop (scalars broadcast)."""
    try:
        v = numpy.asarray(f(tgrid), dtype=float)
        if v.shape == tgrid.shape:
            return v
        if v.shape == ():
            return numpy.full(tgrid.shape, float(v))
    except Exception:
        pass
    return numpy.array([float(f(t)) for t in tgrid])


class NonInertialFrameForce(Potential):
    """The fictitious forces of integrating in a non-inertial frame
    (NonInertialFrameForce.py:35-360): the frame's origin accelerates by
    a0(t) (three functions or constants) and the frame rotates at Omega
    (a number for rotation about z, or three; constants or functions of t,
    with Omegadot); x0(t), v0(t) the origin's position and velocity when it
    both rotates and accelerates.  A velocity-dependent force: integrated
    with RK4, RK6, DOPR54 and DOP853 (galpy reroutes the symplectic methods).
    Packed as pot_type 39 (the functions passed as pot_tfuncs) or, with
    cinterp, 45 (natural splines of the functions over the integration's
    times, cinterp_n points; integratePlanarOrbit.py:835-965)."""

    pot_code = 39
    dissipative = True

    def __init__(self, amp=1.0, Omega=None, Omegadot=None, x0=None, v0=None, a0=None, cinterp=True,
                 cinterp_n=3000):
        Potential.__init__(self, amp=amp)
        self.isNonAxi = True
        self._cinterp, self._cinterp_n = bool(cinterp), int(cinterp_n)
        self._rot_acc = Omega is not None
        self._omegaz_only = numpy.size(Omega) == 1 if not callable(Omega) else True
        self._const_freq = Omegadot is None
        self._Omega_as_func = self._rot_acc and (callable(Omega) if self._omegaz_only else callable(Omega[0]))
        self._Omega, self._Omegadot = Omega, Omegadot
        if self._Omega_as_func and Omegadot is None:  # central differences (NonInertialFrameForce.py:28-32)
            self._const_freq = False
            fd = lambda f: (lambda t: (f(t + 1e-6) - f(t - 1e-6)) / 2e-6)
            self._Omegadot = fd(Omega) if self._omegaz_only else [fd(o) for o in Omega]
        self._lin_acc = a0 is not None
        self._a0_as_func = self._lin_acc and callable(a0[0])
        if self._lin_acc:
            self._a0 = list(a0) if self._a0_as_func else [(lambda t, c=c: c) for c in a0]
        if self._lin_acc and self._rot_acc:
            if x0 is None or v0 is None:
                raise ValueError("x0 and v0 are needed when the frame both rotates and accelerates")
            self._x0, self._v0 = list(x0), list(v0)
        self._Omega2 = (float(numpy.sum(numpy.asarray(Omega, dtype=float) ** 2))
                        if self._rot_acc and not self._Omega_as_func else 0.0)

    def _base_args(self):
        a = [self._amp] + [0.0] * 10 + [float(self._rot_acc), float(self._lin_acc), float(self._omegaz_only),
                                        float(self._const_freq), float(self._Omega_as_func)]
        if self._Omega_as_func or not self._rot_acc:
            return a + [0.0] * 7
        a += [0.0, 0.0, float(self._Omega)] if self._omegaz_only else [float(o) for o in self._Omega]
        a.append(self._Omega2)
        if self._const_freq:
            return a + [0.0, 0.0, 0.0]
        return a + ([0.0, 0.0, float(self._Omegadot)] if self._omegaz_only else [float(o) for o in self._Omegadot])

    def _functions(self):
        """The primitives in galpy's index order: a0, x0, v0, Omega."""
        f = []
        if self._lin_acc:
            f += self._a0
            if self._rot_acc:
                f += self._x0 + self._v0
        if self._Omega_as_func:
            f += [self._Omega] if self._omegaz_only else list(self._Omega)
        return f

    def pack(self, t=None):
        """(pot_type, pot_args, pot_tfuncs) as galpy's
        _parse_noninertial_frame_force packs them."""
        genuine = (self._lin_acc and self._a0_as_func) or (self._lin_acc and self._rot_acc) or self._Omega_as_func
        if self._cinterp and genuine:
            if t is None:
                raise NotImplementedError("NonInertialFrameForce with cinterp=True needs the integration times")
            tmin, tmax = float(numpy.amin(t)), float(numpy.amax(t))
            tgrid = numpy.linspace(tmin, tmax, self._cinterp_n)
            arrs = [_on_grid(f, tgrid) for f in self._functions()]
            args = [float(len(arrs)), float(self._cinterp_n)] + list(tgrid)
            for a in arrs:
                args += list(a)
            return 45, args + self._base_args() + [tmin, tmax], []
        tf = []
        if self._lin_acc:
            tf += self._a0
            if self._rot_acc:
                tf += self._x0 + self._v0
        if self._Omega_as_func:
            tf += [self._Omega, self._Omegadot] if self._omegaz_only else list(self._Omega) + list(self._Omegadot)
        return 39, self._base_args(), tf

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        raise NotImplementedError("a NonInertialFrameForce has no potential")


class ChandrasekharDynamicalFrictionForce(Potential):
    """Chandrasekhar dynamical friction on a satellite of mass GMs moving
    through the density of ``dens`` (ChandrasekharDynamicalFrictionForce.py
    :20-330): F = -4 pi GMs amp rho lnLambda [erf X - 2X e^-X^2/sqrt(pi)] v /
    v^3, X = v / (sqrt 2 sigma_r(r)), lnLambda constant (const_lnLambda) or
    0.5 ln(1 + r^2 / (gamma^2 max(rhm, GMs/v^2)^2)), nothing inside minr.
    sigma_r: a function of r, sampled on nr points over [minr, maxr]; the
    default density is a LogarithmicHalo (normalize=1, q=1) with sigma_r =
    1/sqrt(2).  Velocity-dependent: integrated with RK4, RK6, DOPR54 and
    DOP853; full orbits only.  Packed as pot_type -7 followed by the density's
    codes; arguments: npot of the density, its arguments, nr, the radii,
    sigma_r there, then amp, 8 cache slots, GMs, rhm, gamma^2, lnLambda (-1:
    variable), minr^2, minr, maxr (integrateFullOrbit.py:517-545)."""

    pot_code = -7
    dissipative = True

    def __init__(self, amp=1.0, GMs=0.1, gamma=1.0, rhm=0.0, dens=None, sigmar=None, const_lnLambda=False,
                 minr=0.0001, maxr=25.0, nr=501):
        Potential.__init__(self, amp=amp * GMs)
        self.isNonAxi = True
        self._ms, self._gamma, self._rhm = GMs, gamma, rhm
        self._minr, self._maxr = minr, maxr
        if dens is None:
            dens = LogarithmicHaloPotential(normalize=1.0, q=1.0)
            if sigmar is None:
                sigmar = lambda r: 1.0 / numpy.sqrt(2.0)
        if sigmar is None:
            raise NotImplementedError("give sigmar, the radial velocity dispersion as a function of r")
        self._dens_pot = dens
        self._rs = numpy.linspace(minr, maxr, nr)
        self._srs = numpy.array([sigmar(r) for r in self._rs], dtype=float)
        self._lnLambda = const_lnLambda if const_lnLambda else False
        self._amp *= 4.0 * numpy.pi

    def own_args(self):
        return ([float(len(self._rs))] + list(self._rs) + list(self._srs) + [self._amp, -1.0] + [0.0] * 7
                + [self._ms, self._rhm, self._gamma**2.0, -1.0 if not self._lnLambda else self._lnLambda,
                   self._minr**2.0, self._rs[0], self._rs[-1]])

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        raise NotImplementedError("dynamical friction has no potential")


class FDMDynamicalFrictionForce(ChandrasekharDynamicalFrictionForce):
    """Dynamical friction in fuzzy dark matter (FDMDynamicalFrictionForce.py
    :8-260): the Chandrasekhar force scaled by FDMfactor / CDMfactor where
    that is below one (always, for a constant const_FDMfactor), FDMfactor
    from kr = 2 (m/hbar) v r and the Mach number v / sigma_r.  ``mhbar``: the
    particle mass over hbar in the internal units (galpy's m converted).
    Packed as -11 wrapping the Chandrasekhar force (-7) of the same
    satellite, own arguments its sampled sigma_r, the 16, m/hbar and the
    constant factor or -1 (integrateFullOrbit.py:468-515)."""

    pot_code = -11

    def __init__(self, amp=1.0, GMs=0.1, gamma=1.0, rhm=0.0, mhbar=1.0, dens=None, sigmar=None,
                 const_lnLambda=False, const_FDMfactor=False, minr=0.0001, maxr=25.0, nr=501):
        ChandrasekharDynamicalFrictionForce.__init__(self, amp=amp, GMs=GMs, gamma=gamma, rhm=rhm, dens=dens,
                                                     sigmar=sigmar, const_lnLambda=const_lnLambda, minr=minr,
                                                     maxr=maxr, nr=nr)
        self._mhbar = float(mhbar)
        self._const_FDMfactor = const_FDMfactor


def _os_Rz_to_uv(R, z, delta):
    """Prolate spheroidal (u, v) of focal length delta (galpy's
    coords.Rz_to_uv, as OblateStaeckelWrapperPotential.c:4-13 forms it)."""
    d1 = numpy.sqrt((z + delta) ** 2 + R**2)
    d2 = numpy.sqrt((z - delta) ** 2 + R**2)
    return numpy.arccosh(0.5 / delta * (d1 + d2)), numpy.arccos(0.5 / delta * (d1 - d2))


class OblateStaeckelWrapperPotential(WrapperPotential):
    """The wrapped potential in Staeckel form (OblateStaeckelWrapperPotential.py
    :20-380): Phi = amp (U(u) - V(v)) / (sinh^2 u + sin^2 v) in prolate
    spheroidal coordinates of focal length delta, U(u) = cosh^2 u Phi(u, pi/2),
    V(v) = cosh^2 u0 Phi(u0, pi/2) - (sinh^2 u0 + sin^2 v) Phi(u0, v), the
    wrapped potential taken at phi = t = 0; exact for a Staeckel potential of
    the same delta.  u0: a number, or (R, z) of the reference point.
    pot_type -3, own arguments [amp, delta, u0, v0 = pi/2, refpot]."""

    pot_code = -3

    def __init__(self, amp=1.0, pot=None, delta=0.5, u0=0.0):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._delta = float(delta)
        if isinstance(u0, (tuple, list, numpy.ndarray)):
            u0 = _os_Rz_to_uv(u0[0], u0[1], self._delta)[0]
        self._u0 = float(u0)
        self._v0 = numpy.pi / 2.0
        R0, z0 = self._uv_to_Rz(self._u0, self._v0)
        self._refpot = float(evaluatePotentials(self._pot, R0, z0)) * numpy.cosh(self._u0) ** 2
        self.isNonAxi = False

    def _uv_to_Rz(self, u, v):
        return self._delta * numpy.sinh(u) * numpy.sin(v), self._delta * numpy.cosh(u) * numpy.cos(v)

    def _terms(self, R, z):
        """u, v, U, dU/du, V, dV/dv (OblateStaeckelWrapperPotential.c:31-137)."""
        u, v = _os_Rz_to_uv(R, z, self._delta)
        Ra, za = self._uv_to_Rz(u, self._v0)
        pa = evaluatePotentials(self._pot, Ra, za)
        U = numpy.cosh(u) ** 2 * pa
        dU = 2 * numpy.cosh(u) * numpy.sinh(u) * pa - numpy.cosh(u) ** 2 * (
            evaluateRforces(self._pot, Ra, za) * Ra / (numpy.tanh(u) + 1e-12)
            + evaluatezforces(self._pot, Ra, za) * za * numpy.tanh(u))
        Rb, zb = self._uv_to_Rz(self._u0, v)
        pb = evaluatePotentials(self._pot, Rb, zb)
        pre0 = numpy.sinh(self._u0) ** 2 + numpy.sin(v) ** 2
        V = self._refpot - pre0 * pb
        dV = -2 * numpy.sin(v) * numpy.cos(v) * pb + pre0 * (
            evaluateRforces(self._pot, Rb, zb) * Rb / numpy.tan(v) - evaluatezforces(self._pot, Rb, zb) * zb * numpy.tan(v))
        return u, v, U, dU, V, dV

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        u, v, U, _, V, _ = self._terms(R, z)
        return (U - V) / (numpy.sinh(u) ** 2 + numpy.sin(v) ** 2)

    def _forces(self, R, z):
        u, v, U, dU, V, dV = self._terms(R, z)
        d = self._delta
        pre = numpy.sinh(u) ** 2 + numpy.sin(v) ** 2
        dpu, dpv = 2 * numpy.sinh(u) * numpy.cosh(u), 2 * numpy.sin(v) * numpy.cos(v)
        FR = (-dU * d * numpy.sin(v) * numpy.cosh(u) + dV * numpy.tanh(u) * z
              + (U - V) * (dpu * d * numpy.sin(v) * numpy.cosh(u) + dpv * numpy.tanh(u) * z) / pre) / (d * pre) ** 2
        Fz = (-dU * R / numpy.tan(v) - dV * d * numpy.sin(v) * numpy.cosh(u)
              + (U - V) * (dpu / numpy.tan(v) * R - dpv * d * numpy.sin(v) * numpy.cosh(u)) / pre) / (d * pre) ** 2
        return FR, Fz

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        return self._forces(R, z)[0]

    def _zforce(self, R, z, phi=0.0, t=0.0):
        return self._forces(R, z)[1]

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        return 0.0 * numpy.asarray(R, dtype=float)

    def wrapper_args(self):
        return [self._amp, self._delta, self._u0, self._v0, self._refpot]


class TimeDependentAmplitudeWrapperPotential(WrapperPotential):
    """Multiplies the wrapped potential by a function of time A(t)
    (TimeDependentAmplitudeWrapperPotential.py:15-90); pot_type -9, own
    arguments [amp] and one time function, A (integrateFullOrbit.py:569-580).
    The GPU evaluates A from a table built by calling it on the host."""

    pot_code = -9

    def __init__(self, amp=1.0, A=None, pot=None):
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        if A is None or not callable(A):
            raise ValueError("TimeDependentAmplitudeWrapperPotential needs a function A(t)")
        self._A = A

    def _factor(self, t):
        t = numpy.asarray(t, dtype=float)
        return numpy.vectorize(lambda s: float(self._A(float(s))))(t) if t.ndim else float(self._A(float(t)))

    def wrapper_args(self):
        return [self._amp]

    def wrapper_tfuncs(self):
        return [self._A]


class MovingObjectPotential(WrapperPotential):
    """The potential of an object moving on a given orbit
    (MovingObjectPotential.py:20-150): the axisymmetric potential ``pot``
    (default a Plummer sphere, amp 0.06, b 0.01) centred on the orbit's
    position at each time; pot_type -6, own arguments [n, t[n], x[n], y[n],
    z[n], amp, t[0], t[-1]] (planar: without z; integrateFullOrbit.py:449-463,
    integratePlanarOrbit.py:570-588).  ``orbit``: an integrated Orbit (its
    stored times and positions are the spline's knots).  galpy's C, and so
    the GPU, interpolates the position with natural cubic splines over the
    normalised time, clamped at the ends; the host evaluation here uses the
    same splines."""

    pot_code = -6

    def __init__(self, orbit, pot=None, amp=1.0):
        if pot is None:
            pot = PlummerPotential(amp=0.06, b=0.01)
        if any(q.isNonAxi for q in flatten(pot)):
            raise NotImplementedError("MovingObjectPotential for non-axisymmetric potentials is not supported")
        WrapperPotential.__init__(self, amp=amp, pot=pot)
        self._t = numpy.array(orbit.t, dtype=float)
        self._x = numpy.array(orbit.x(self._t), dtype=float).reshape(-1)
        self._y = numpy.array(orbit.y(self._t), dtype=float).reshape(-1)
        self._z = numpy.array(orbit.z(self._t), dtype=float).reshape(-1)
        tt = (self._t - self._t[0]) / (self._t[-1] - self._t[0])
        self._spl = [interpolate.CubicSpline(tt, v, bc_type="natural") for v in (self._x, self._y, self._z)]
        self.isNonAxi = True

    def _pos(self, t):
        d = numpy.clip((numpy.asarray(t, dtype=float) - self._t[0]) / (self._t[-1] - self._t[0]), 0.0, 1.0)
        return [s(d) for s in self._spl]

    def _rel(self, R, z, phi, t):
        ox, oy, oz = self._pos(t)
        x, y = R * numpy.cos(phi), R * numpy.sin(phi)
        return x, y, ox, oy, oz, numpy.sqrt((x - ox) ** 2 + (y - oy) ** 2)

    def _evaluate(self, R, z, phi=0.0, t=0.0):
        x, y, ox, oy, oz, Rd = self._rel(R, z, phi, t)
        return evaluatePotentials(self._pot, Rd, oz - z, phi, t)

    def _Rforce(self, R, z, phi=0.0, t=0.0):
        x, y, ox, oy, oz, Rd = self._rel(R, z, phi, t)
        RF = evaluateRforces(self._pot, Rd, oz - z, phi, t)
        return -RF * (numpy.cos(phi) * (ox - x) + numpy.sin(phi) * (oy - y)) / Rd

    def _zforce(self, R, z, phi=0.0, t=0.0):
        x, y, ox, oy, oz, Rd = self._rel(R, z, phi, t)
        return -evaluatezforces(self._pot, Rd, oz - z, phi, t)

    def _phitorque(self, R, z, phi=0.0, t=0.0):
        x, y, ox, oy, oz, Rd = self._rel(R, z, phi, t)
        RF = evaluateRforces(self._pot, Rd, oz - z, phi, t)
        return -RF * R * (numpy.cos(phi) * (oy - y) - numpy.sin(phi) * (ox - x)) / Rd

    def wrapper_args(self):
        return ([float(len(self._t))] + list(self._t) + list(self._x) + list(self._y) + list(self._z)
                + [self._amp, self._t[0], self._t[-1]])

    def wrapper_args_planar(self):
        return ([float(len(self._t))] + list(self._t) + list(self._x) + list(self._y)
                + [self._amp, self._t[0], self._t[-1]])


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


def _pack(comps, planar, tfuncs=None, t=None):
    """Codes and arguments of a list of components in galpy's order; a
    wrapper's code is followed by its wrapped codes, its arguments are
    nwrapped, the wrapped arguments, then its own
    (integrateFullOrbit.py:396-448, integratePlanarOrbit.py:507-565); time
    functions are appended to ``tfuncs`` in the same order (wrapped
    components first, integrateFullOrbit.py:569-580)."""
    pot_type, pot_args = [], []
    if tfuncs is None:
        tfuncs = []
    for p in comps:
        if isinstance(p, WrapperPotential):
            if planar and isinstance(p, RotateAndTiltWrapperPotential):  # integratePlanarOrbit.py:592-599
                raise NotImplementedError(
                    "Planar orbit integration for RotateAndTiltWrapperPotential is not implemented; "
                    "integrate an orbit with (z, vz) = (0, 0) instead")
            inner = flatten(p._pot)
            wt, wa = _pack(inner, planar, tfuncs, t)
            pot_type.append(p.pot_code)
            pot_type.extend(wt)
            pot_args.append(float(_npot(inner)))
            pot_args.extend(wa)
            pot_args.extend(p.wrapper_args_planar() if planar and hasattr(p, "wrapper_args_planar")
                            else p.wrapper_args())
            tfuncs.extend(getattr(p, "wrapper_tfuncs", lambda: [])())
            continue
        if isinstance(p, MN3ExponentialDiskPotential):
            for args in p.expand():
                pot_type.append(MiyamotoNagaiPotential.pot_code)
                pot_args.extend(args)
            continue
        if isinstance(p, DiskSCFPotential):
            pt, pa, _ = p.pack()
            pot_type.extend(pt)
            pot_args.extend(pa)
            continue
        if isinstance(p, ChandrasekharDynamicalFrictionForce):
            if planar:  # integratePlanarOrbit.py: no planar dynamical friction in C
                raise NotImplementedError("ChandrasekharDynamicalFrictionForce has no planar version")
            inner = flatten(p._dens_pot)
            wt, wa = _pack(inner, planar, tfuncs, t)
            if isinstance(p, FDMDynamicalFrictionForce):  # -11 wrapping the -7 of the same satellite
                pot_type.append(-11)
                pot_args.append(1.0)
            pot_type.append(-7)
            pot_type.extend(wt)
            pot_args.append(float(_npot(inner)))
            pot_args.extend(wa)
            pot_args.extend(p.own_args())
            if isinstance(p, FDMDynamicalFrictionForce):
                pot_args.extend(p.own_args())
                pot_args.extend([p._mhbar, -1.0 if not p._const_FDMfactor else p._const_FDMfactor])
            continue
        if isinstance(p, NonInertialFrameForce):
            code, pa, tf = p.pack(t)
            pot_type.append(code)
            pot_args.extend(pa)
            tfuncs.extend(tf)
            continue
        if not isinstance(p, Potential) or p.pot_code is None:
            raise NotImplementedError(
                f"Potential {type(p).__name__} is not supported by the HIP orbit-integration backend"
            )
        pot_type.append(p.pot_code)
        pot_args.extend(p.c_args_planar() if planar else p.c_args())
    return pot_type, pot_args


def _npot(comps):
    """galpy's npot: an MN3ExponentialDiskPotential counts as its three
    Miyamoto-Nagai disks (integrateFullOrbit.py:184)."""
    return sum(3 if isinstance(p, MN3ExponentialDiskPotential)
               else (1 + len(p._Sigma_dict) if isinstance(p, DiskSCFPotential) else 1) for p in comps)


def _purge_null(comps):
    """integrateFullOrbit.py:51-56: NullPotentials are dropped unless nothing
    else is left."""
    kept = [p for p in comps if not isinstance(p, NullPotential)]
    return kept if kept else comps


def parse_pot(pot, t=None):
    """galpy's ``_parse_pot`` (integrateFullOrbit.py:43-614) for the supported
    potentials: returns ``(npot, pot_type int32[], pot_args float64[],
    pot_tfuncs)``, pot_tfuncs the components' time functions (callables);
    npot counts the top-level components (a wrapper is one).

    Raises NotImplementedError for a potential without a GPU kernel, as galpy
    does for a potential without a C implementation (:604-611).  ``t``: the
    integration's times (a NonInertialFrameForce with cinterp tabulates its
    functions over their range)."""
    comps = _purge_null(flatten(pot))
    tfuncs = []
    pot_type, pot_args = _pack(comps, False, tfuncs, t)
    return (
        _npot(comps),
        numpy.array(pot_type, dtype=numpy.int32, order="C"),
        numpy.array(pot_args, dtype=numpy.float64, order="C"),
        tfuncs,
    )


def parse_pot_planar(pot, t=None):
    """galpy's planar ``_parse_pot`` (integratePlanarOrbit.py:33-663) for the
    supported potentials, taken in the plane z = 0: returns
    ``(npot, pot_type int32[], pot_args float64[], [])``."""
    comps = _purge_null(flatten(pot))
    tfuncs = []
    pot_type, pot_args = _pack(comps, True, tfuncs, t)
    return (
        _npot(comps),
        numpy.array(pot_type, dtype=numpy.int32, order="C"),
        numpy.array(pot_args, dtype=numpy.float64, order="C"),
        tfuncs,
    )


def evaluatePotentials(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p(R, z, phi, t) for p in flatten(pot))


def evaluateRforces(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p.Rforce(R, z, phi, t) for p in flatten(pot))


def evaluatezforces(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p.zforce(R, z, phi, t) for p in flatten(pot))


def evaluatephitorques(pot, R, z=0.0, phi=0.0, t=0.0):
    return sum(p.phitorque(R, z, phi, t) for p in flatten(pot))
