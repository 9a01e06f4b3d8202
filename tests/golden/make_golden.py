"""Writes the golden fixtures under tests/golden/ from the mathematics alone
(mpmath at 50 digits and closed forms); no reference code is imported.

    python tests/golden/make_golden.py

incgamma.json    lower/upper incomplete gamma gamma(a,x), Gamma(a,x) for the
                 exponents of PowerSphericalPotentialwCutoff (a = 1.5 - alpha/2,
                 1 - alpha/2) over x in [1e-8, 1e3].
potentials.json  forces and potential values of the supported potentials at
                 seeded points, from closed forms in mpmath; forces are the
                 50-digit numerical derivatives of the closed-form potential;
                 second derivatives R2/z2/Rz for the axisymmetric ones.  The
                 spherical and Kuzmin closed forms are checked against
                 Poisson's equation with their documented densities when the
                 file is written.
"""

import json
import os

import mpmath as mp
import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 50


def incgamma():
    out = []
    for a in (0.1, 0.6, 1.0, 1.5, 0.35):
        for x in numpy.concatenate([numpy.geomspace(1e-8, 1e3, 61), [a + 1.0 - 1e-9, a + 1.0 + 1e-9]]):
            x = float(x)
            lo = mp.gammainc(a, 0, x)
            up = mp.gammainc(a, x, mp.inf)
            out.append({"a": a, "x": x, "lower": float(lo), "upper": float(up)})
    return out


# --- closed-form potentials (natural units, amplitude conventions of galpy) ---
def psc_phi(R, z, amp, alpha, rc):
    r2 = R * R + z * z
    r = mp.sqrt(r2)
    x = r2 / rc**2
    g1 = mp.gammainc(1 - alpha / 2, 0, x)
    g2 = mp.gammainc(mp.mpf(3) / 2 - alpha / 2, 0, x)
    return amp * 2 * mp.pi * rc ** (3 - alpha) / r * (r / rc * g1 - g2)


def mn_phi(R, z, amp, a, b):
    return -amp / mp.sqrt(R * R + (a + mp.sqrt(z * z + b * b)) ** 2)


def nfw_phi(R, z, amp, a):
    r = mp.sqrt(R * R + z * z)
    return -amp * mp.log(1 + r / a) / r


def hq_phi(R, z, amp, a):
    r = mp.sqrt(R * R + z * z)
    return -amp / (1 + r / a) / 2 / a


def lh_phi(R, z, phi, amp, q, core2, onem1overb2):
    Rt2 = R * R * (1 - onem1overb2 * mp.sin(phi) ** 2) if onem1overb2 < 1 else R * R
    return amp / 2 * mp.log(Rt2 + (z / q) ** 2 + core2)


def bar_phi(R, z, phi, t, ampaf, tform, tsteady, rb, omegab, barphi):
    if t < tform:
        smooth = 0
    elif t < tsteady:
        xi = 2 * (t - tform) / (tsteady - tform) - 1
        smooth = mp.mpf(3) / 16 * xi**5 - mp.mpf(5) / 8 * xi**3 + mp.mpf(15) / 16 * xi + mp.mpf(1) / 2
    else:
        smooth = 1
    r2 = R * R + z * z
    r = mp.sqrt(r2)
    if r <= rb:
        out = ((r / rb) ** 3 - 2) * R * R / r2
    else:
        out = -((rb / r) ** 3) * R * R / r2
    return ampaf * smooth * mp.cos(2 * (phi - omegab * t - barphi)) * out


def _r(R, z):
    return mp.sqrt(R * R + z * z)


def powersph_phi(R, z, amp, alpha):
    r = _r(R, z)
    if alpha == 2:
        return amp * mp.log(r)
    return -amp * r ** (2 - alpha) / (alpha - 2)


def jaffe_phi(R, z, amp, a):
    return -amp * mp.log(1 + a / _r(R, z)) / a


def isochrone_phi(R, z, amp, b):
    return -amp / (b + mp.sqrt(R * R + z * z + b * b))


def plummer_phi(R, z, amp, b):
    return -amp / mp.sqrt(R * R + z * z + b * b)


def pseudoiso_phi(R, z, amp, a):
    r = _r(R, z)
    return amp * (mp.log(1 + r * r / (a * a)) / 2 + a / r * mp.atan(r / a)) / a


def kuzmin_phi(R, z, amp, a):
    return -amp / mp.sqrt(R * R + (a + abs(z)) ** 2)


def burkert_phi(R, z, amp, a):
    x = _r(R, z) / a
    return -amp * a * a * mp.pi / x * (
        -mp.pi + 2 * (1 + x) * mp.atan(1 / x) + 2 * (1 + x) * mp.log(1 + x) + (1 - x) * mp.log(1 + x * x))


def dehnen_phi(R, z, amp, a, alpha):
    r = _r(R, z)
    return -amp * (1 - (r / (r + a)) ** (2 - alpha)) / (a * (2 - alpha) * (3 - alpha))


def flatpower_phi(R, z, amp, alpha, q2, core2):
    m2 = core2 + R * R + z * z / q2
    if alpha == 0:
        return amp * mp.log(m2) / 2
    return -amp * m2 ** (-alpha / 2) / alpha


def homog_phi(R, z, amp, R2, R3):
    r2 = R * R + z * z
    if r2 < R2:
        return amp * (r2 - 3 * R2)
    return -2 * amp * R3 / mp.sqrt(r2)


# 4 pi rho of each closed form above, from the documented densities
# (natural units, G = 1): Poisson's equation pins normalisations and
# zero-free constants that a retyped potential would not
def four_pi_rho(name, R, z, args):
    r = _r(R, z)
    A = mp.mpf(args[0])
    if name.startswith("powersph") or name == "kepler":
        alpha = mp.mpf(args[1])
        return A * (3 - alpha) * r ** (-alpha)
    if name == "jaffe":
        a = mp.mpf(args[1])
        return A * a / (r * r * (r + a) ** 2)
    if name == "isochrone":
        b = mp.mpf(args[1])
        s = mp.sqrt(r * r + b * b)
        return A * (3 * (b + s) * s * s - r * r * (b + 3 * s)) / ((b + s) ** 3 * s**3)
    if name == "plummer":
        b = mp.mpf(args[1])
        return 3 * A * b * b * (r * r + b * b) ** mp.mpf(-2.5)
    if name == "pseudoiso":
        a = mp.mpf(args[1])
        return A / (a**3 * (1 + r * r / (a * a)))
    if name == "kuzmin":  # vacuum off the plane
        return mp.mpf(0)
    if name == "burkert":
        x = r / mp.mpf(args[1])
        return 4 * mp.pi * A / ((1 + x) * (1 + x * x))
    if name.startswith("dehnen") and name != "dehnenbar":
        a = mp.mpf(args[1])
        alpha = mp.mpf(args[2]) if len(args) > 2 else mp.mpf(0)
        return A * a * r ** (-alpha) * (r + a) ** (alpha - 4)
    if name == "homogsphere":
        return 6 * A if r * r < mp.mpf(args[1]) else mp.mpf(0)
    if name.startswith("flatpower"):
        # the Laplacian of f(m^2), m^2 = R^2 + z^2/q^2 + core^2, derived by hand:
        # amp m^(-alpha-4) / q^2 [R^2 (1 - alpha q^2) + z^2 (2 - (1+alpha)/q^2)
        # + core^2 (2 q^2 + 1)]
        alpha, q2, c2 = (mp.mpf(v) for v in args[1:4])
        m2 = c2 + R * R + z * z / q2
        return A / q2 * (R * R * (1 - alpha * q2) + z * z * (2 - (1 + alpha) / q2) + c2 * (2 * q2 + 1)) * \
            m2 ** (-alpha / 2 - 2)
    raise KeyError(name)


def forces(fn, R, z, phi, t):
    """(-dPhi/dR, -dPhi/dz, -dPhi/dphi) by 50-digit numerical differentiation."""
    FR = -mp.diff(lambda v: fn(v, z, phi, t), R)
    Fz = -mp.diff(lambda v: fn(R, v, phi, t), z)
    Fp = -mp.diff(lambda v: fn(R, z, v, t), phi)
    return FR, Fz, Fp


def potentials():
    # parameters in galpy's packed order (pot_type codes of integrateFullOrbit.c)
    amp_b = 0.05 / (2 * mp.pi / (mp.mpf(3) / 2 - mp.mpf("0.9")) * mp.hyp1f1(mp.mpf("0.6"), mp.mpf("1.6"), -1 / (mp.mpf("1.9") / 8) ** 2))
    amp_d = mp.mpf("0.6") * (1 + (mp.mpf(3) / 8 + mp.mpf("0.28") / 8) ** 2) ** mp.mpf(1.5)
    amp_h = mp.mpf("0.35") / abs(1 / (1 + mp.mpf(2)) - mp.log(1 + 1 / mp.mpf(2)))
    omegab = mp.mpf("1.85")
    tb = 2 * mp.pi / omegab
    cases = {
        "psc": (15, [float(amp_b), 1.8, 1.9 / 8.0, 0.0, 0.0],
                lambda R, z, phi, t: psc_phi(R, z, amp_b, mp.mpf("1.8"), mp.mpf("1.9") / 8)),
        "mn": (5, [float(amp_d), 3.0 / 8.0, 0.28 / 8.0],
               lambda R, z, phi, t: mn_phi(R, z, amp_d, mp.mpf(3) / 8, mp.mpf("0.28") / 8)),
        "nfw": (9, [float(amp_h), 2.0], lambda R, z, phi, t: nfw_phi(R, z, amp_h, mp.mpf(2))),
        "hernquist": (8, [1.0, 1.3], lambda R, z, phi, t: hq_phi(R, z, 1, mp.mpf("1.3"))),
        "loghalo_q09": (0, [1.0, 0.9, 1e-16, 2.0],
                        lambda R, z, phi, t: lh_phi(R, z, phi, 1, mp.mpf("0.9"), mp.mpf("1e-16"), 2)),
        "loghalo_triax": (0, [1.0, 0.9, 1e-16, 1.0 - 1.0 / 0.8**2],
                          lambda R, z, phi, t: lh_phi(R, z, phi, 1, mp.mpf("0.9"), mp.mpf("1e-16"), 1 - 1 / mp.mpf("0.8") ** 2)),
        "dehnenbar": (1, [0.01, float(-4 * tb), float(-2 * tb), 0.8, 1.85, 25.0 * numpy.pi / 180.0],
                      lambda R, z, phi, t: bar_phi(R, z, phi, t, mp.mpf("0.01"), -4 * tb, -2 * tb, mp.mpf("0.8"), omegab, mp.mpf(25.0 * numpy.pi / 180.0))),
    }
    # the spherical and Kuzmin components, packed as galpy packs them
    # (integrateFullOrbit.py:82-92,176-213,358-365); Phi from the closed form
    # with the packed arguments
    new = {
        "powersph": (7, ["1.3", "1.7"], powersph_phi),
        "powersph_a2": (7, ["0.7", "2"], powersph_phi),
        "kepler": (7, ["0.9", "3"], powersph_phi),
        "jaffe": (10, ["1.1", "0.8"], jaffe_phi),
        "isochrone": (14, ["1.2", "0.7"], isochrone_phi),
        "plummer": (17, ["0.9", "0.8"], plummer_phi),
        "pseudoiso": (18, ["1.4", "1.3"], pseudoiso_phi),
        "kuzmin": (19, ["1.0", "0.6"], kuzmin_phi),
        "burkert": (20, ["0.8", "2.0"], burkert_phi),
        "dehnencore": (33, ["1.5", "1.2"], lambda R, z, amp, a: dehnen_phi(R, z, amp, a, 0)),
        "dehnen": (34, ["1.3", "0.9", "1.5"], dehnen_phi),
        "dehnen_a05": (34, ["0.7", "1.4", "0.5"], dehnen_phi),
        "homogsphere": (35, ["1.0", "1.21", "1.331"], homog_phi),
        "flatpower": (12, ["1.1", "0.5", "0.81", "0.01"], flatpower_phi),
        "flatpower_a0": (12, ["0.9", "0", "0.64", "0.04"], flatpower_phi),
    }
    for name, (code, sargs, phi_fn) in new.items():
        margs = [mp.mpf(float(v)) for v in sargs]  # the packed doubles exactly
        cases[name] = (code, [float(v) for v in sargs],
                       (lambda f, m: (lambda R, z, phi, t: f(R, z, *m)))(phi_fn, margs))
    with_hessian = {"psc", "mn", "nfw", "hernquist", "loghalo_q09"} | set(new)
    rng = numpy.random.default_rng(2024)
    out = {}
    for name, (code, args, fn) in cases.items():
        pts = []
        for _ in range(24):
            R = float(rng.uniform(0.05, 3.0))
            z = float(rng.normal(0.0, 0.4))
            phi = float(rng.uniform(0.0, 2 * numpy.pi))
            t = float(rng.uniform(-30.0, 10.0)) if name == "dehnenbar" else 0.0
            FR, Fz, Fp = forces(fn, mp.mpf(R), mp.mpf(z), mp.mpf(phi), mp.mpf(t))
            row = {"R": R, "z": z, "phi": phi, "t": t, "pot": float(fn(mp.mpf(R), mp.mpf(z), mp.mpf(phi), mp.mpf(t))),
                   "FR": float(FR), "Fz": float(Fz), "Fphi": float(Fp)}
            if name in with_hessian:
                f2 = lambda a, b: fn(a, b, mp.mpf(phi), mp.mpf(t))
                R2 = mp.diff(f2, (mp.mpf(R), mp.mpf(z)), (2, 0))
                z2 = mp.diff(f2, (mp.mpf(R), mp.mpf(z)), (0, 2))
                row["R2"] = float(R2)
                row["z2"] = float(z2)
                row["Rz"] = float(mp.diff(f2, (mp.mpf(R), mp.mpf(z)), (1, 1)))
                if name in new:
                    # Poisson: Phi_RR + Phi_R / R + Phi_zz = 4 pi rho
                    lap = R2 - FR / R + z2
                    rho4 = four_pi_rho(name, mp.mpf(R), mp.mpf(z), [mp.mpf(v) for v in cases[name][1]])
                    assert abs(lap - rho4) <= mp.mpf("1e-25") * max(1, abs(rho4)), (name, R, z, lap, rho4)
            pts.append(row)
        out[name] = {"pot_type": [code], "pot_args": args, "points": pts}
    return out


def main():
    with open(os.path.join(HERE, "incgamma.json"), "w") as f:
        json.dump(incgamma(), f)
    with open(os.path.join(HERE, "potentials.json"), "w") as f:
        json.dump(potentials(), f)


if __name__ == "__main__":
    main()
