"""The Nystrom form of DOP853 used by the device (galpy_amd/csrc/dop853_rkn.h,
generated by gen_dop853_rkn.py) is the same method as leung_dop853.c's
first-order form: one step of both, in numpy, on a nonlinear force, agrees in
every stage position, the new state, both error sums and the dense-output
coefficients (a wrong expanded coefficient shows up at O(h) or O(h^2))."""

import importlib.util
import os
import subprocess
import sys

import numpy

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "galpy_amd", "csrc", "gen_dop853_rkn.py")


def _gen():
    spec = importlib.util.spec_from_file_location("gen_dop853_rkn", GEN)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _acc(q):
    # smooth, anisotropic, nonlinear (a softened point mass plus a tidal term)
    r2 = q @ q + 0.3
    return -q / r2**1.5 + numpy.array([0.05, -0.02, 0.01]) * q[::-1]


def test_generated_header_is_current():
    r = subprocess.run([sys.executable, GEN, "--check"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_nystrom_step_equals_first_order_step():
    g = _gen()
    T = g.tableau(g.parse_tableau())
    N = g.nystrom(T)
    a, b, bhh, er, d = T["a"], T["b"], T["bhh"], T["er"], T["d"]
    q0 = numpy.array([0.9, -0.35, 0.2])
    p0 = numpy.array([0.1, 0.95, -0.15])
    h = 0.37

    # leung_dop853.c form: k_j = f(y_j), y = (q, p), f = (p, a(q))
    def f(y):
        return numpy.concatenate([y[3:], _acc(y[:3])])

    y0 = numpy.concatenate([q0, p0])
    k = {1: f(y0)}
    Yref = {}
    for s in range(2, 13):
        Yref[s] = y0 + h * sum(a[s][j] * k[j] for j in sorted(a[s]))
        k[s] = f(Yref[s])
    kb = sum(b[j] * k[j] for j in sorted(b))
    yn = y0 + h * kb
    e3 = kb - sum(bhh[j] * k[j] for j in sorted(bhh))
    e5 = sum(er[j] * k[j] for j in sorted(er))
    k[13] = f(yn)
    for s in (14, 15, 16):
        Yref[s] = y0 + h * sum(a[s][j] * k[j] for j in sorted(a[s]))
        k[s] = f(Yref[s])
    rc = {i: h * sum(d[i][j] * k[j] for j in sorted(d[i])) for i in d}

    # Nystrom form with the generated (exact, rounded) coefficients
    F = lambda x: float(x)  # noqa: E731
    A = {1: _acc(q0)}
    for s in range(2, 13):
        Q = q0 + h * (F(N["QP"][s]) * p0 + h * sum(F(v) * A[l] for l, v in N["QA"][s].items()))
        assert numpy.abs(Q - Yref[s][:3]).max() < 1e-14, s
        A[s] = _acc(Q)
    qn = q0 + h * (F(N["BP"]) * p0 + h * sum(F(v) * A[l] for l, v in N["BA"].items()))
    pn = p0 + h * sum(b[j] * A[j] for j in sorted(b))
    assert numpy.abs(qn - yn[:3]).max() < 1e-14 and numpy.abs(pn - yn[3:]).max() < 1e-14
    e3q = F(N["E3P"]) * p0 + h * sum(F(v) * A[l] for l, v in N["E3A"].items())
    e5q = F(N["E5P"]) * p0 + h * sum(F(v) * A[l] for l, v in N["E5A"].items())
    assert numpy.abs(e3q - e3[:3]).max() < 1e-14 and numpy.abs(e5q - e5[:3]).max() < 1e-14
    A[13] = _acc(qn)
    for s in (14, 15, 16):
        Q = q0 + h * (F(N["QP"][s]) * p0 + F(N["QN"][s]) * pn
                      + h * sum(F(v) * A[l] for l, v in N["QA"][s].items()))
        assert numpy.abs(Q - Yref[s][:3]).max() < 1e-14, s
        A[s] = _acc(Q)
    for i in d:
        rq = h * (F(N["DP"][i]) * p0 + F(N["DN"][i]) * pn + h * sum(F(v) * A[l] for l, v in N["DA"][i].items()))
        assert numpy.abs(rq - rc[i][:3]).max() < 1e-13 * max(1.0, numpy.abs(rc[i][:3]).max()), i
    # the stage times of the dense stages: the position rows sum to c_s
    for s in (14, 15, 16):
        assert abs(float(N["QP"][s] + N["QN"][s]) - T["c"][s]) < 1e-15


def test_expanded_coefficients_are_consistent_with_the_tableau():
    """Row sums: QP_s = c_s (stages with c in the tableau), sum(b) = 1, the
    error weights sum to ~0; QA_s1 = sum_j a_sj a_j1."""
    g = _gen()
    T = g.tableau(g.parse_tableau())
    N = g.nystrom(T)
    for s in range(2, 12):
        assert abs(float(N["QP"][s]) - T["c"][s]) < 4e-15, s
    assert float(N["BP"]) == 1.0
    assert abs(float(N["E3P"])) < 1e-15 and abs(float(N["E5P"])) < 1e-15
    for s in range(3, 13):
        want = sum(T["a"][s][j] * T["a"][j].get(1, 0.0) for j in T["a"][s] if j in T["a"])
        assert abs(float(N["QA"][s].get(1, 0.0)) - want) < 1e-14 * max(1.0, abs(want)), s
