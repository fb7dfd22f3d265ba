"""The optimizer's Moré-Thuente line search against scipy's DCSRCH (the
MINPACK-2 routine behind line_search_wolfe1, which the reference calls at
math_utils.py:37-52): the same trial steps, the same count and the same
outcome on smooth functions of the rotation angle, on functions that need
every interval-update case, and on failures (maxiter, xtol, step bounds)."""
import math

import numpy as np
import pytest
from scipy.optimize._dcsrch import DCSRCH

from dftpy_amd import linesearch as LS


def _run_both(fun, dfun, stp0, c1=1e-4, c2=0.2, xtol=1e-12, amin=0.0, amax=np.pi, maxiter=100):
    trials_a, trials_b = [], []

    def fa(t):
        trials_a.append(t)
        return fun(t)

    def fb(t):
        trials_b.append(t)
        return fun(t)

    f0, g0 = fun(0.0), dfun(0.0)
    stp_a, task_a = LS.search(fa, dfun, f0, g0, stp0, c1=c1, c2=c2, xtol=xtol, stpmin=amin, stpmax=amax,
                              maxiter=maxiter)
    d = DCSRCH(fb, dfun, c1, c2, xtol, amin, amax)
    stp_b, _, _, task_b = d(stp0, phi0=f0, derphi0=g0, maxiter=maxiter)
    return (stp_a, task_a, trials_a), (stp_b, task_b.decode(), trials_b)


def _angle_family(rng):
    """E(t) along phi cos t + p sin t: a trigonometric polynomial in 2t with
    a descent direction at t = 0."""
    k = rng.integers(1, 4)
    a = rng.normal(size=k) * rng.uniform(0.1, 10.0)
    b = rng.normal(size=k)
    if sum((j + 1) * 2 * b[j] for j in range(k)) > 0:
        b = -b
    f = lambda t: float(sum(a[j] * math.cos(2 * (j + 1) * t) + b[j] * math.sin(2 * (j + 1) * t) for j in range(k)))
    df = lambda t: float(sum(-2 * (j + 1) * a[j] * math.sin(2 * (j + 1) * t) + 2 * (j + 1) * b[j]
                             * math.cos(2 * (j + 1) * t) for j in range(k)))
    return f, df


@pytest.mark.parametrize("seed", range(40))
def test_same_trials_as_dcsrch_angle_functions(seed):
    rng = np.random.default_rng(seed)
    f, df = _angle_family(rng)
    if df(0.0) >= 0:
        pytest.skip("not a descent direction")
    for stp0 in (1.0, float(rng.uniform(1e-3, 0.5))):
        a, b = _run_both(f, df, stp0)
        assert a[2] == b[2], (a, b)
        assert a[0] == b[0]
        assert (a[1] == b[1]) or (a[1].startswith("WARN") and b[1].startswith("WARN"))


@pytest.mark.parametrize("seed", range(30))
def test_same_trials_as_dcsrch_polynomials(seed):
    """Quartics with a far minimum (extrapolation, case 3/4) and steep walls
    (case 1), tight and loose curvature tolerances."""
    rng = np.random.default_rng(100 + seed)
    c = rng.normal(size=5)
    c[1] = -abs(c[1]) - 0.1
    c[4] = abs(c[4]) * rng.choice([1e-4, 1.0, 100.0]) + 1e-6
    # numpy scalars, as the optimizer's device sums arrive
    f = lambda t: np.float64(np.polyval(c[::-1], t))
    df = lambda t: np.float64(np.polyval((c[1:] * np.arange(1, 5))[::-1], t))
    for c2 in (0.9, 0.2, 0.01):
        for stp0 in (1.0, 1e-3, 20.0):
            a, b = _run_both(f, df, stp0, c2=c2, amax=50.0, xtol=1e-14)
            assert a[2] == b[2], (c2, stp0, a, b)
            assert a[0] == b[0]


def test_every_update_case_is_reached():
    """The four interval-update cases of the paper all occur on the test
    families above (guards against a branch that is never exercised)."""
    seen = set()
    orig = LS.step_update

    def spy(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
        sg = LS._sign(dp) * LS._sign(dx)
        seen.add(1 if fp > fx else 2 if sg < 0 else 3 if abs(dp) < abs(dx) else 4)
        return orig(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax)

    LS.step_update = spy
    try:
        for seed in range(30):
            rng = np.random.default_rng(100 + seed)
            c = rng.normal(size=5)
            c[1] = -abs(c[1]) - 0.1
            c[4] = abs(c[4]) * rng.choice([1e-4, 1.0, 100.0]) + 1e-6
            f = lambda t: float(np.polyval(c[::-1], t))
            df = lambda t: float(np.polyval((c[1:] * np.arange(1, 5))[::-1], t))
            for c2 in (0.9, 0.2, 0.01):
                for stp0 in (1.0, 1e-3, 20.0):
                    LS.search(f, df, f(0.0), df(0.0), stp0, c2=c2, stpmax=50.0)
    finally:
        LS.step_update = orig
    assert seen == {1, 2, 3, 4}


def test_failures_match_dcsrch():
    # maxiter exhausted: a function whose curvature condition is hard to meet
    f = lambda t: -t + 1e-3 * math.sin(50 * t)
    df = lambda t: -1 + 5e-2 * math.cos(50 * t)
    a, b = _run_both(f, df, 1.0, amax=np.pi, maxiter=4)
    assert a[0] is None and b[0] is None
    assert a[2] == b[2]
    # the step bound: a linear decrease reaches amax
    f = lambda t: -2.0 * t
    df = lambda t: -2.0
    a, b = _run_both(f, df, 1.0, amax=3.0)
    assert a[0] is None and b[0] is None
    assert a[1] == b[1] == "WARNING: STP = STPMAX"
    assert a[2] == b[2]
    # an ascent direction is an input error
    a, b = _run_both(lambda t: t, lambda t: 1.0, 1.0)
    assert a[0] is None and a[1] == b[1] == "ERROR: INITIAL G .GE. ZERO"


def test_non_finite_value_fails():
    """A NaN energy ends the search with a warning (no exception, no loop)."""
    f = lambda t: float("nan") if t > 0.5 else -t
    df = lambda t: float("nan") if t > 0.5 else -1.0
    stp, task = LS.search(f, df, 0.0, -1.0, 1.0, c2=0.2, stpmax=np.pi)
    assert stp is None and task.startswith("WARN")
