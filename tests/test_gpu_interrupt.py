"""The interrupt contract of the blocking host call (bovy_symplecticode.c:
38-45 and 300-311; integrateFullOrbit.py:737-738): a SIGINT delivered while
integrateFullOrbit runs makes every orbit not yet finished return err = -10
promptly, the Python caller turns that into KeyboardInterrupt, and the next
call runs normally (the flag is cleared).  Driven once by a real SIGINT to
the process and once by galpy_hip_set_interrupt from another thread, for a
fixed-step (symplec4, polled every 64 output steps) and the adaptive DOP853
kernel (polled inside its persistent loop)."""

import ctypes
import signal
import threading
import time

import numpy
import pytest

from helpers import ics

pytestmark = pytest.mark.gpu
_LOGTOL = float(numpy.log(1e-12))  # the C-ABI takes natural logs of rtol / atol
_NODT = -9999.99  # no fixed step: the integrator chooses


def _call(pot, yo, t, code, dt):
    from galpy_amd import _lib
    from galpy_amd import potential as P

    lib = _lib.load()
    npot, pt, pa, _ = P.parse_pot(pot)
    yo = numpy.array(yo, dtype=numpy.float64, order="C")  # the call overwrites its yo
    res = numpy.empty((yo.shape[0], t.size, 6))
    err = numpy.zeros(yo.shape[0], dtype=numpy.int32)
    lib.integrateFullOrbit(yo.shape[0], yo, t.size, numpy.ascontiguousarray(t), 0, npot, pt, pa, None,
                           dt, _LOGTOL, _LOGTOL, res, err, code, None)
    return res, err


def _interrupt_while(run, how, first=0.3):
    """Run `run()` while another thread keeps interrupting it from `first`
    seconds on (a SIGINT to the process, or the library's setter) until it
    returns; the process ignores SIGINT outside the library's own handler."""
    from galpy_amd import _lib

    lib = _lib.load()
    done = threading.Event()

    def poke():
        if done.wait(first):
            return
        while not done.is_set():
            if how == "sigint":
                signal.pthread_kill(threading.main_thread().ident, signal.SIGINT)
            else:
                lib.galpy_hip_set_interrupt(1)
            done.wait(0.1)

    old = signal.signal(signal.SIGINT, lambda *a: None)
    th = threading.Thread(target=poke)
    try:
        th.start()
        t0 = time.perf_counter()
        out = run()
        el = time.perf_counter() - t0
    finally:
        done.set()
        th.join()
        lib.galpy_hip_set_interrupt(0)
        signal.signal(signal.SIGINT, old)
    return out, el


@pytest.mark.parametrize("how", ["sigint", "setter"])
@pytest.mark.parametrize("method", ["symplec4", "dop853"])
def test_interrupt_returns_minus_ten_and_clears(method, how):
    from galpy_amd import potential as P

    mw = P.MWPotential2014_factory()
    # warm-up: code load and first-launch set-up before the timed call
    small = ics(64, seed=5)
    tw = numpy.linspace(0.0, 1.0, 11)
    code = 3 if method == "symplec4" else 6
    ref_small, e0 = _call(mw, small, tw, code, 0.01 if code == 3 else _NODT)
    assert numpy.all(e0 == 0)
    if method == "symplec4":
        # 1,024 orbits, 2e6 steps each: about 10 s if never interrupted
        yo, t, dt = ics(1024, seed=6), numpy.linspace(0.0, 20000.0, 4001), 0.01
    else:
        # 65,536 orbits to t = 50,000 with two outputs: several seconds
        yo, t, dt = ics(65536, seed=7), numpy.array([0.0, 50000.0]), _NODT
    (res, err), el = _interrupt_while(lambda: _call(mw, yo, t, code, dt), how)
    print(f"{method} {how}: returned after {el:.2f} s, {int((err == -10).sum())} of {err.size} interrupted",
          flush=True)
    assert numpy.all((err == 0) | (err == -10))
    assert numpy.any(err == -10)
    assert el < 3.0, f"the interrupted call took {el:.2f} s"
    # the flag is cleared: the same small call runs to completion, bit-identical
    again, e1 = _call(mw, small, tw, code, 0.01 if code == 3 else _NODT)
    assert numpy.all(e1 == 0)
    assert numpy.array_equal(again, ref_small)


def test_python_caller_raises_keyboardinterrupt():
    from galpy_amd import potential as P
    from galpy_amd.integrate import integrateFullOrbit_hip

    mw = P.MWPotential2014_factory()
    yo, t = ics(65536, seed=8), numpy.array([0.0, 50000.0])
    with pytest.raises(KeyboardInterrupt):
        _interrupt_while(lambda: integrateFullOrbit_hip(mw, yo.copy(), t, "dop853_hip"), "setter")
    res, err = integrateFullOrbit_hip(mw, ics(32, seed=9), numpy.linspace(0.0, 1.0, 5), "dop853_hip")
    assert numpy.all(err == 0)
