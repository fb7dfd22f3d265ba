"""Host control of the forward-difference truncated-Newton loop
(optimization.py:92-148) on the CPU, with the device plan replaced by a
stand-in that does the same vector arithmetic on CPU tensors (the kernels
themselves are covered by the GPU tests): the pAp == 0 branch, the pAp < 0
branches and a non-finite pAp."""
import math

import numpy as np
import pytest
import torch

from dftpy_amd.optimization import Optimization, OptimizationError


class _Plan:
    dtype = torch.float64

    def lincomb(self, a, x, b, y, z):
        z.copy_(a * x + b * y)
        return z

    def dot3(self, a, b):
        return float((a * b).sum()), float((a * a).sum()), float((b * b).sum())

    def residual(self, V, phi, mu, res):
        res.copy_((V - mu) * phi)
        return float((res * res).sum())

    def cg_update(self, a, p, Ap, d, r):
        d.add_(a * p)
        r.sub_(a * Ap)
        return float((r * r).sum())


class _Ev:
    """V(phi1) from a callable of (phi1, p): lets a test choose Ap."""

    def __init__(self, fn):
        self.fn = fn

    def eval_phi(self, phi, p, c, s, V, aux):
        phi1 = c * phi + s * p
        V.copy_(self.fn(phi1, p))
        return np.zeros(9)


class _W:
    """phi = 2^20 everywhere and |eps p| below half its ulp, so that
    phi + eps p == phi exactly and a chosen V gives an exactly chosen Ap."""

    def __init__(self, n, ev, seed=0):
        g = torch.Generator().manual_seed(seed)
        self.P = _Plan()
        self.ev = ev
        self.phi = torch.full((n,), 2.0 ** 20, dtype=torch.float64)
        self.res = 0.05 * (torch.rand(n, generator=g, dtype=torch.float64) - 0.5)
        e = lambda: torch.zeros(n, dtype=torch.float64)
        self._t = dict(d=e(), r=e(), cgp=e(), Ap=e(), best=e(), tmp=e())

    def tn(self):
        return self._t


def _opt():
    o = object.__new__(Optimization)
    o.optimization_options = {"hvp": "fd", "maxfun": 50}
    o.nfeval = 0
    o.verbose = False
    return o


def test_pap_zero_first_step_gives_steepest_descent():
    mu = 0.0
    holder = {}
    # V(phi1) = res0 / phi1: (V - mu) phi1 = res0 exactly, so Ap = 0 and pAp = 0
    W = _W(64, _Ev(lambda phi1, p: holder["res0"] / phi1))
    holder["res0"] = W.res.clone()
    n = _opt()._tn_loop(W, mu)
    assert n == 1
    assert torch.equal(W.tn()["d"], -W.res)


def test_pap_zero_later_keeps_direction_built_so_far():
    mu = 0.0
    holder = {"calls": 0}
    diag = torch.linspace(1.0, 10.0, 64, dtype=torch.float64)  # r1 > 0.1 r0 after one step

    def V(phi1, p):
        holder["calls"] += 1
        if holder["calls"] == 1:  # a positive-definite first step: Ap = D p
            return (holder["res0"] + 1e-9 * diag * p) / phi1
        return holder["res0"] / phi1  # then no curvature

    W = _W(64, _Ev(V), seed=1)
    holder["res0"] = W.res.clone()
    n = _opt()._tn_loop(W, mu)
    assert n == 2
    # the first CG step only: alpha = r0 / pAp along p = -res
    p = -W.res
    alpha = float((p * p).sum() / (p * diag * p).sum())
    assert torch.allclose(W.tn()["d"], alpha * p, rtol=1e-6, atol=1e-12)


def test_negative_curvature_first_step():
    mu = 0.0
    holder = {}
    W = _W(32, _Ev(lambda phi1, p: (holder["res0"] - 1e-9 * 3.0 * p) / phi1), seed=2)
    holder["res0"] = W.res.clone()
    n = _opt()._tn_loop(W, mu)
    # pAp < 0 at the first step: direction = r0 / pAp p (optimization.py:116-120)
    assert n == 1
    assert torch.allclose(W.tn()["d"], (1.0 / -3.0) * -W.res, rtol=1e-6, atol=1e-12)


def test_non_finite_pap_raises():
    W = _W(16, _Ev(lambda phi1, p: torch.full_like(phi1, float("nan"))))
    with pytest.raises(OptimizationError):
        _opt()._tn_loop(W, 0.0)
