"""Moré-Thuente line search on the rotation angle theta.

The reference's density optimizer calls scipy's ``line_search_wolfe1``
(math_utils.py:37-52), i.e. MINPACK-2's DCSRCH: a search for a step that
satisfies the sufficient-decrease condition f(t) <= f(0) + c1 t f'(0) and
the curvature condition |f'(t)| <= c2 |f'(0)|, keeping an interval of
uncertainty [stx, sty] that is updated by safeguarded cubic / quadratic /
secant interpolation.  This module states that algorithm as published in
J. J. Moré and D. J. Thuente, "Line search algorithms with guaranteed
sufficient decrease", ACM Trans. Math. Software 20 (1994) 286-307
(sections 2-4: the modified function psi(t) = f(t) - f(0) - c1 t f'(0) while
no step has reached psi <= 0 with f' >= 0, the four interval-update cases,
the extrapolation bounds [1.1, 4] times the step, the bisection when the
interval does not shrink by 0.66), with MINPACK-2's termination tests and
messages.  The result has to be the reference's step for step: the trial
angles decide the optimizer's trajectory.

``search(f, fprime, f0, g0, stp, ...)`` returns (step or None, task) and
counts its calls through f.  A task starting with 'WARNING' or 'ERROR' means
failure (the reference's "Line-Search Failed", converged = 1).
"""
import math

__all__ = ["search", "step_update"]

_P5 = 0.5
_P66 = 0.66
_XTRAPL = 1.1
_XTRAPU = 4.0


def _sign(x):
    return 1 if x > 0 else (-1 if x < 0 else 0)


def _cubic_gamma(theta, da, db, clip):
    s = max(abs(theta), abs(da), abs(db))
    arg = (theta / s) ** 2 - (da / s) * (db / s)
    if clip:
        arg = max(0.0, arg)
    return s * math.sqrt(arg) if arg >= 0 else float("nan")


def step_update(stx, fx, dx, sty, fy, dy, stp, fp, dp, brackt, stpmin, stpmax):
    """One safeguarded step of the interval update (Moré-Thuente section 4,
    MINPACK-2 dcstep).  (stx, fx, dx): best step so far; (sty, fy, dy): the
    other end of the interval; (stp, fp, dp): the trial just evaluated.
    Returns the updated (stx, fx, dx, sty, fy, dy, new trial, brackt)."""
    sgnd = _sign(dp) * _sign(dx)
    if fp > fx:
        # case 1: higher value -- the minimum is bracketed; the cubic step if it
        # is closer to stx than the quadratic one, else their midpoint
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        gamma = _cubic_gamma(theta, dx, dp, False)
        if stp < stx:
            gamma = -gamma
        p = (gamma - dx) + theta
        q = ((gamma - dx) + gamma) + dp
        stpc = stx + (p / q) * (stp - stx)
        stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx)
        stpf = stpc if abs(stpc - stx) <= abs(stpq - stx) else stpc + (stpq - stpc) / 2.0
        brackt = True
    elif sgnd < 0:
        # case 2: lower value, derivatives of opposite sign -- bracketed; the
        # cubic or the secant step, whichever is farther from stp
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        gamma = _cubic_gamma(theta, dx, dp, False)
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = ((gamma - dp) + gamma) + dx
        stpc = stp + (p / q) * (stx - stp)
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
        brackt = True
    elif abs(dp) < abs(dx):
        # case 3: lower value, same-sign derivative decreasing in magnitude;
        # the cubic step only if it tends to infinity in the right direction
        theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp
        gamma = _cubic_gamma(theta, dx, dp, True)
        if stp > stx:
            gamma = -gamma
        p = (gamma - dp) + theta
        q = (gamma + (dx - dp)) + gamma
        r = p / q
        if r < 0 and gamma != 0:
            stpc = stp + r * (stx - stp)
        elif stp > stx:
            stpc = stpmax
        else:
            stpc = stpmin
        stpq = stp + (dp / (dp - dx)) * (stx - stp)
        if brackt:
            stpf = stpc if abs(stpc - stp) < abs(stpq - stp) else stpq
            if stp > stx:
                stpf = min(stp + _P66 * (sty - stp), stpf)
            else:
                stpf = max(stp + _P66 * (sty - stp), stpf)
        else:
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            stpf = max(stpmin, min(stpmax, stpf))
    else:
        # case 4: lower value, same-sign derivative not decreasing
        if brackt:
            theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp
            gamma = _cubic_gamma(theta, dy, dp, False)
            if stp > sty:
                gamma = -gamma
            p = (gamma - dp) + theta
            q = ((gamma - dp) + gamma) + dy
            stpf = stp + (p / q) * (sty - stp)
        elif stp > stx:
            stpf = stpmax
        else:
            stpf = stpmin
    if fp > fx:
        sty, fy, dy = stp, fp, dp
    else:
        if sgnd < 0:
            sty, fy, dy = stx, fx, dx
        stx, fx, dx = stp, fp, dp
    return stx, fx, dx, sty, fy, dy, stpf, brackt


def search(f, fprime, f0, g0, stp, c1=1e-4, c2=0.9, xtol=1e-14, stpmin=0.0, stpmax=50.0, maxiter=100):
    """Moré-Thuente search from f(0) = f0, f'(0) = g0 < 0 with first trial
    ``stp``.  ``maxiter`` counts the set-up pass and every trial, as DCSRCH's
    driver does (so at most maxiter - 1 evaluations of f).  Returns
    (step, task): step is None unless task starts with 'CONVERGENCE'."""
    f0, g0, stp = float(f0), float(g0), float(stp)
    task = None
    for bad, msg in ((stp < stpmin, "ERROR: STP .LT. STPMIN"), (stp > stpmax, "ERROR: STP .GT. STPMAX"),
                     (g0 >= 0, "ERROR: INITIAL G .GE. ZERO"), (c1 < 0, "ERROR: FTOL .LT. ZERO"),
                     (c2 < 0, "ERROR: GTOL .LT. ZERO"), (xtol < 0, "ERROR: XTOL .LT. ZERO"),
                     (stpmin < 0, "ERROR: STPMIN .LT. ZERO"), (stpmax < stpmin, "ERROR: STPMAX .LT. STPMIN")):
        if bad:
            task = msg
    if task is not None:
        return None, task
    brackt = False
    stage = 1
    gtest = c1 * g0
    width = stpmax - stpmin
    width1 = width / _P5
    stx, fx, gx = 0.0, f0, g0
    sty, fy, gy = 0.0, f0, g0
    stmin, stmax = 0.0, stp + _XTRAPU * stp
    for _ in range(maxiter - 1):
        fv = float(f(stp))
        gv = float(fprime(stp))
        ftest = f0 + stp * gtest
        if stage == 1 and fv <= ftest and gv >= 0:
            stage = 2
        task = None
        if brackt and (stp <= stmin or stp >= stmax):
            task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS"
        if brackt and stmax - stmin <= xtol * stmax:
            task = "WARNING: XTOL TEST SATISFIED"
        if stp == stpmax and fv <= ftest and gv <= gtest:
            task = "WARNING: STP = STPMAX"
        if stp == stpmin and (fv > ftest or gv >= gtest):
            task = "WARNING: STP = STPMIN"
        if fv <= ftest and abs(gv) <= c2 * -g0:
            task = "CONVERGENCE"
        if task is not None:
            return (stp if task == "CONVERGENCE" else None), task
        if stage == 1 and fv <= fx and fv > ftest:
            # the modified function psi with the auxiliary values
            stx, fxm, gxm, sty, fym, gym, stp, brackt = step_update(
                stx, fx - stx * gtest, gx - gtest, sty, fy - sty * gtest, gy - gtest,
                stp, fv - stp * gtest, gv - gtest, brackt, stmin, stmax)
            fx, fy = fxm + stx * gtest, fym + sty * gtest
            gx, gy = gxm + gtest, gym + gtest
        else:
            stx, fx, gx, sty, fy, gy, stp, brackt = step_update(
                stx, fx, gx, sty, fy, gy, stp, fv, gv, brackt, stmin, stmax)
        if brackt:
            if abs(sty - stx) >= _P66 * width1:
                stp = stx + _P5 * (sty - stx)
            width1 = width
            width = abs(sty - stx)
            stmin, stmax = min(stx, sty), max(stx, sty)
        else:
            stmin = stp + _XTRAPL * (stp - stx)
            stmax = stp + _XTRAPU * (stp - stx)
        stp = min(max(stp, stpmin), stpmax)
        if brackt and (stp <= stmin or stp >= stmax or stmax - stmin <= xtol * stmax):
            stp = stx
        if not math.isfinite(stp):
            return None, "WARNING: NON-FINITE STEP"
    return None, "WARNING: dcsrch did not converge within max iterations"
