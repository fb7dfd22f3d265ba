"""Density optimization on phi = sqrt(rho), device resident.

Mirrors ``Optimization`` of optimization/optimization.py:31-547 (defaults
:64-78, TN :92-148, LBFGS :150-173 with math_utils.py:307-325, CG
math_utils.py:327-361, orthogonalisation :205-226, value/derivative
:228-280, main loop :283-525, convergence :527-540, chemical potential
:542-544).  Every field operation is a HIP kernel on device buffers; only
scalars (energies, dot products, line-search decisions) reach the host.

With a TotalFunctional built from this package's functionals, each line-
search trial is one fused evaluation at phi cos t + p sin t (the trial field
is never materialised), and the TN inner step uses the exact Hessian-vector
product (``hvp='exact'``, default) instead of the reference's forward
difference (``hvp='fd'`` reproduces that one).

Options beyond the reference's:
  hvp            'exact' | 'fd'
  precision      'fp64' | 'fp32' | 'mixed' (default: the dtype of guess_rho).
                 'mixed': FP64 energies, potentials, line search and phi;
                 the TN inner conjugate-gradient Hessian products in FP32
  mixed_phase1   mixed mode: run FP32 outer iterations first and switch to
                 the FP64 outer loop once |dE| < mixed_switch (Ha, default
                 1e-8 * |E|, about the FP32 energy noise) or the FP32
                 line search fails (cut after 8 trials)
  ls_first_step  'unit' (the reference code: first trial theta = 1) or
                 'theta' (first trial = the running theta)
  tn_graph       True (default): the device TN loop replays one captured
                 hipGraph per CG iteration instead of launching its kernels
                 (same results; launch mode is as fast at 192^3); False:
                 launches
  defer_sums     True (default; exact-HVP TN on device terms with the unit
                 first step): the orthonormalisation runs on the device with
                 the first line-search trial queued behind it, and the
                 residual norm is read with the next orthonormalisation sums
                 (each step's log line is completed then); same results as
                 False, which waits for each sum
  cg_spectra     'plan' (default): the exact-HVP inner CG keeps its spectra
                 in the work precision, in split rows (the Nyquist column
                 apart) where the lengths allow; 'bf16': FP32 inner CG loops
                 store the spectra between transform passes as bf16 (the
                 transforms still compute in FP32), on plans whose lengths
                 allow it (Plan.cg_spectra); the directions then carry bf16
                 rounding, the energies and the line search do not;
                 'padded': the padded rows of the evaluations
"""
import math
import time

import numpy as np
import torch

from . import _lib as L
from . import linesearch as LS
from .field import DirectField
from .functional import TotalFunctional, FunctionalOutput, AbstractFunctional

HARTREE_EV = 27.21138602436724


class OptimizationError(ValueError):
    """A non-finite energy, residual or curvature: the solve cannot go on."""


def _finite(x, what):
    if not math.isfinite(x):
        raise OptimizationError("non-finite %s (%r): the density or the functional produced NaN/Inf" % (what, x))
    return x


class HostTermsEvaluator:
    """Evaluation for an optimizer whose energy is not entirely made of this
    package's device terms: a TotalFunctional that holds plugin terms (any
    AbstractFunctional, total_functional.py:69-77), or any callable returning
    an object with ``.energy`` and ``.potential`` (the reference's
    EnergyEvaluator contract, optimization.py:83-86, :251, :275).

    The device terms of a TotalFunctional still run fused; each plugin term
    (or the callable) is called on DirectFields of the same grid, dtype and
    device with calcType {'E', 'V'} and the keyword arguments ``phi`` and
    ``lphi`` the reference passes, and its energy and potential are added.
    Every field operation around the calls (the trial field phi cos t +
    p sin t, its square, the sums for the chemical potential and the
    line-search derivative) is a device kernel of the plan.  Mirrors the
    DeviceEvaluator interface that the optimizer drives."""

    def __init__(self, evaluator, grid, dtype):
        from .plan import Plan
        self.grid = grid
        self.dtype = dtype
        if isinstance(evaluator, TotalFunctional):
            self.total = evaluator
            self.calls = evaluator.plugin_terms()
            self.dev = evaluator.evaluator(grid, dtype) if evaluator.has_device_terms() else None
        else:
            self.total = None
            self.calls = [("EVALUATOR", evaluator)]
            self.dev = None
        if self.dev is not None:
            self.P = self.dev.P
            self.terms = self.dev.terms
        else:
            self.P = Plan(tuple(int(x) for x in grid.nrR), grid.lattice,
                          precision="fp64" if dtype == torch.float64 else "fp32", full=grid.full,
                          device=grid.get_device())
            self.terms = dict(tf=False, vw=False, wt=False, lda=False, hartree=False, pseudo=False, gga=None,
                              lwt=None, rho0_fixed=None)
        self._tmp = None

    @property
    def configured_rho0(self):
        return self.dev.configured_rho0 if self.dev is not None else None

    def ensure_kernel(self, rho0):
        return self.dev.ensure_kernel(rho0) if self.dev is not None else False

    def energies(self, sums):
        out = self.dev.energies(sums[:-1]) if self.dev is not None else {}
        out["HOST"] = float(sums[-1])
        return out

    def eval_phi(self, phi, p, c, s, V, aux):
        P = self.P
        if self.dev is not None:
            sums = self.dev.eval_phi(phi, p, c, s, V, aux)
        else:
            sums = np.zeros(L.NSUM)
            V.zero_()
        if self._tmp is None:
            self._tmp = (P.empty(), P.empty())
        newphi, p2 = self._tmp
        if p is None:
            newphi.copy_(phi)
        else:
            P.lincomb(c, phi, s, p, newphi)
        with torch._C.DisableTorchFunctionSubclass():
            rho = newphi * newphi
        rf = DirectField(grid=self.grid, data=rho, dtype=self.dtype)
        pf = DirectField(grid=self.grid, data=newphi, dtype=self.dtype)
        e_host = 0.0
        dV = P.dV
        for key, f in self.calls:
            out = f(rf, calcType={"E", "V"}, phi=pf, lphi=False)
            e_host += _finite(float(out.energy), "energy of %s" % key)
            pot = out.potential
            pot = pot if isinstance(pot, torch.Tensor) else torch.as_tensor(np.asarray(pot))
            with torch._C.DisableTorchFunctionSubclass():
                pot = pot.to(device=V.device, dtype=V.dtype).reshape(V.shape).contiguous()
                V.add_(pot)
            sums[L.S_VRHO] += P.dot3(pot, rho)[0] * dV
            if p is not None:
                P.lincomb(c, p, -s, phi, p2)
                sums[L.S_DTHETA] += 2.0 * P.dot_vphip(pot, newphi, p2) * dV
        if self.dev is None:
            sums[L.S_RHO] = P.dot3(newphi, newphi)[1] * dV
        return np.concatenate([sums, [e_host]])


class _Work:
    """Device buffers of one precision."""

    def __init__(self, ev):
        self.ev = ev
        self.P = P = ev.P
        e = lambda: torch.empty(P.nr, dtype=P.dtype, device=P.device)
        self.e = e
        self.phi, self.V, self.Vt, self.res, self.res_prev, self.p = e(), e(), e(), e(), e(), e()
        self.aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
        self.auxt = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
        self.W0 = None  # HVP diagonal (V - mu + 2 rho a(rho)) at phi
        self._tn = None

    def tn(self):
        if self._tn is None:
            e = self.e
            work = torch.empty((4,) + tuple(self.P.nr), dtype=self.P.dtype, device=self.P.device)
            self._tn = dict(d=e(), work=work, r=work[0], cgp=work[1], Ap=work[2], best=work[3], tmp=e())
        return self._tn


class _State:
    """Loop state of one optimize_rho (the locals of optimization.py:283-525)."""


class Optimization:
    def __init__(self, optimization_method="CG-HS", optimization_options=None, EnergyEvaluator=None,
                 guess_rho=None):
        self.rho = guess_rho
        self.optimization_options = {} if optimization_options is None else optimization_options
        defaults = {"maxcor": 6, "ftol": 1.0e-7, "xtol": 1.0e-12, "maxfun": 50, "maxiter": 200, "maxls": 30,
                    "econv": 1.0e-5, "c1": 1e-4, "c2": 0.2, "algorithm": "EMM", "vector": "Orthogonalization",
                    "ncheck": 2, "h0": 1.0, "hvp": "exact", "precision": None, "mixed_switch": None,
                    "tn_graph": True, "cg_spectra": "plan",
                    "ls_first_step": "unit", "defer_sums": True}
        for k, v in defaults.items():
            self.optimization_options.setdefault(k, v)
        if self.optimization_options["cg_spectra"] not in ("plan", "bf16", "padded"):
            # layout and storage of the spectra in the exact-HVP inner CG: the
            # plan precision in split rows (default), bf16 on FP32 plans
            # (every transform still in FP32), or the padded rows
            raise ValueError("cg_spectra must be 'plan', 'bf16' or 'padded', not %r"
                             % (self.optimization_options["cg_spectra"],))
        if EnergyEvaluator is None:
            raise AttributeError("Must provide an energy evaluator")
        if not callable(EnergyEvaluator):
            raise TypeError("EnergyEvaluator must be callable: evaluator(rho, calcType=..., phi=..., lphi=...) "
                            "returning an object with .energy and .potential")
        if self.optimization_options["algorithm"] != "EMM" or \
                self.optimization_options["vector"] != "Orthogonalization":
            raise AttributeError("only algorithm='EMM' with vector='Orthogonalization' is on the device path")
        self.EnergyEvaluator = EnergyEvaluator
        # host_terms: energy terms evaluated through their own compute() (a
        # plugin AbstractFunctional, or an evaluator that is not a
        # TotalFunctional); their second derivative is unknown
        self.host_terms = not isinstance(EnergyEvaluator, TotalFunctional) or bool(EnergyEvaluator.plugin_terms())
        if isinstance(EnergyEvaluator, TotalFunctional):
            terms = EnergyEvaluator._merged_terms()
            no_exact = terms["gga"] is not None or terms["lwt"] is not None or self.host_terms
            if (terms["gga"] is not None or terms["lwt"] is not None) and \
                    self.optimization_options["precision"] in ("fp32", "mixed"):
                # their potentials carry grad rho / rho^(4/3)-type fields that
                # grow without bound where rho -> 0 (trial rotations reach
                # such points); in FP32 the transforms' rounding of those
                # fields swamps the potential
                raise ValueError("GGA / LWT kinetic functionals run in FP64 on the device path "
                                 "(precision=%r)" % self.optimization_options["precision"])
        else:
            no_exact = True
            prec = self.optimization_options["precision"]
            if prec not in (None, "fp64"):
                raise ValueError("an EnergyEvaluator that is not a TotalFunctional runs in FP64 (precision=%r)" % prec)
            self.optimization_options["precision"] = "fp64"
        if no_exact and self.optimization_options["hvp"] == "exact":
            # no closed-form second derivative of the GGA / LWT kinetic terms
            # or of plugin terms on the device: the reference's
            # forward-difference product
            self.optimization_options["hvp"] = "fd"
        self.optimization_method = optimization_method
        if optimization_method not in ("TN", "LBFGS", "GD") and not optimization_method.startswith("CG"):
            raise AttributeError("The %s direction method not implemented." % optimization_method)
        self.log = []
        self.timings = {}
        self.nfeval = 0
        self.nhvp = 0
        self._tn_chunk = 4
        self._tn_hist = []
        self._host_evs = {}
        self.verbose = True

    # ------------------------------------------------------------------
    def _print(self, *a):
        if self.verbose:
            print(*a, flush=True)

    def _eval(self, W, phi, p, t):
        c, s = (1.0, 0.0) if p is None else (math.cos(t), math.sin(t))
        self.nfeval += 1
        return W.ev.eval_phi(phi, p, c, s, W.Vt, W.auxt)

    def _ii(self):
        if not isinstance(self.EnergyEvaluator, TotalFunctional):
            return 0.0  # the evaluator's energy is the whole energy
        ew = self.EnergyEvaluator.ewald
        return ew.energy if ew else 0.0

    def _energy(self, W, sums):
        return float(sum(W.ev.energies(sums).values())) + self._ii_energy

    # --- directions ---------------------------------------------------
    def _direction_tn(self, W, mu, Wi=None):
        """get_direction_TN (optimization.py:92-148) with an exact or a
        forward-difference Hessian-vector product.  With ``Wi`` (mixed
        precision) the inner CG runs on FP32 copies of phi, W0 and the
        residual; the direction is returned to W.p in W's precision."""
        if Wi is not None:
            self._tn_inner_inputs(W, Wi)
            nd = self._tn_loop(Wi, mu)
            W.P.convert_from(Wi.tn()["d"], W.p)
            return nd
        if self.optimization_options["hvp"] != "exact" and W.P.dtype != torch.float64:
            # the forward difference (eps = 1e-9) is below FP32 resolution:
            # the product runs on FP64 copies of phi and the residual
            # (the residual and mu are re-evaluated in FP64 at that phi: the
            # difference quotient subtracts them from an FP64 potential)
            W64 = self._work("fp64")
            W64.P.bind_stream(True)
            W64.P.convert_from(W.phi, W64.phi)
            sums = self._eval(W64, W64.phi, None, 0.0)
            W64.V, W64.Vt = W64.Vt, W64.V
            mu64 = sums[L.S_VRHO] / sums[L.S_RHO]
            W64.P.residual(W64.V, W64.phi, mu64, W64.res)
            nd = self._tn_loop(W64, mu64)
            W.P.convert_from(W64.tn()["d"], W.p)
            return nd
        nd = self._tn_loop(W, mu)
        W.p.copy_(W.tn()["d"])
        return nd

    def _tn_inner_inputs(self, W, Wi):
        """FP32 copies of phi, the residual and the Hessian diagonal W0 for
        the mixed-precision inner CG, with the outer evaluator's WT kernel."""
        if W.ev.configured_rho0 is not None:
            Wi.ev.ensure_kernel(W.ev.configured_rho0)  # same WT kernel as the outer evaluator
        Wi.P.convert_from(W.phi, Wi.phi)
        Wi.P.convert_from(W.res, Wi.res)
        if Wi.W0 is None:
            Wi.W0 = Wi.e()
        Wi.P.convert_from(W.W0, Wi.W0)

    def _tn_loop(self, W, mu):
        P = W.P
        T = W.tn()
        if self.optimization_options["hvp"] == "exact":
            # the whole CG loop on the device; chunk = iterations queued per
            # host synchronisation, predicted from the previous call
            n, stat, pAp = P.tn_direction(W.phi, W.res, W.W0, T["d"], T["work"],
                                          self.optimization_options["maxfun"], self._tn_chunk)
            self.nhvp += n
            # queue about as many iterations as the shortest recent loop took;
            # the device queues 2 more per extra synchronisation after that
            self._tn_hist = (self._tn_hist + [n])[-3:]
            self._tn_chunk = max(2, min(self._tn_hist) - 1)
            _finite(pAp, "p.Ap in the truncated-Newton loop")
            if stat == L.TN_FAILED:
                self._print("!WARN : pAp small than zero :iter = ", n - 1, pAp)
            return n
        d, r, cgp, Ap, best = T["d"], T["r"], T["cgp"], T["Ap"], T["best"]
        d.zero_()
        P.lincomb(-1.0, W.res, 0.0, W.res, r)
        cgp.copy_(r)
        r0 = P.dot3(r, r)[1]
        rconv = r0 * 0.1
        rl = [1e10]
        eps = 1.0e-9
        it = 0
        for it in range(self.optimization_options["maxfun"]):
            # Ap = ((V(phi1) - mu) phi1 - res0) / eps, phi1 = phi + eps p
            tmp = T["tmp"]
            self.nfeval += 1
            W.ev.eval_phi(W.phi, cgp, 1.0, eps, Ap, None)
            P.lincomb(1.0, W.phi, eps, cgp, tmp)
            P.residual(Ap, tmp, mu, Ap)
            P.lincomb(1.0 / eps, Ap, -1.0 / eps, W.res, Ap)
            pAp = _finite(P.dot3(cgp, Ap)[0], "p.Ap in the truncated-Newton loop")
            if pAp > 0.0:
                r1 = _finite(P.cg_update(r0 / pAp, cgp, Ap, d, r), "CG residual norm")
            if pAp < 0.0:
                if it == 0:
                    P.lincomb(r0 / pAp, cgp, 0.0, d, d)
                else:
                    self._print("!WARN : pAp small than zero :iter = ", it, pAp)
                break
            if pAp == 0.0:
                # no curvature information along p: the reference's r0 / pAp
                # would be infinite.  First step: the steepest-descent
                # direction p = -res (the orthonormalisation rescales it);
                # later: keep the direction built so far.
                if it == 0:
                    d.copy_(cgp)
                self._print("!WARN : pAp equal to zero :iter = ", it)
                break
            if r1 < min(rl):
                best.copy_(d)
            rl.append(r1)
            if r1 < rconv:
                break
            elif r1 > 1000 * min(rl[:-1]):
                d.copy_(best)
                break
            elif it > 10 and abs(r0 - r1) < 0.1 * r0:
                d.copy_(best)
                break
            beta = r1 / r0
            r0 = r1
            P.lincomb(1.0, r, beta, cgp, cgp)
        return it + 1

    def _direction_lbfgs(self, W):
        """get_direction_LBFGS (optimization.py:150-173): two-loop recursion,
        first loop skipping index 0 as the reference does."""
        P = W.P
        lb = self._lbfgs
        q = W.p
        P.lincomb(-1.0, W.res, 0.0, W.res, q)
        n = len(lb["s"])
        al = [0.0] * n
        for i in range(n - 1, 0, -1):
            a = lb["rho"][i] * P.dot3(lb["s"][i], q)[0]
            al[i] = a
            P.lincomb(-a, lb["y"][i], 1.0, q, q)
        h0 = self.optimization_options["h0"]
        if not h0:
            gamma = 1.0 if n < 1 else P.dot3(lb["s"][-1], lb["y"][-1])[0] / P.dot3(lb["y"][-1], lb["y"][-1])[1]
            P.lincomb(gamma, q, 0.0, q, q)
        else:
            P.lincomb(h0, q, 0.0, q, q)
        for i in range(n):
            b = lb["rho"][i] * P.dot3(lb["y"][i], q)[0]
            P.lincomb(al[i] - b, lb["s"][i], 1.0, q, q)
        return 1

    def _direction_cg(self, W, first):
        """get_direction_CG (math_utils.py:327-361)."""
        P = W.P
        m = self.optimization_method
        if first or self._dir_prev is None:
            W.P.lincomb(-1.0, W.res, 0.0, W.res, W.p)
            return 1
        R, Rp, Dp = W.res, W.res_prev, self._dir_prev
        t = W.tn()["tmp"]
        P.lincomb(1.0, R, -1.0, Rp, t)  # R - Rp
        if m in ("CG-HS", "CG"):
            beta = P.dot3(R, t)[0] / P.dot3(Dp, t)[0]
        elif m == "CG-PR":
            beta = max(P.dot3(R, t)[0] / P.dot3(Rp, Rp)[1], 0.0)
        elif m == "CG-DY":
            beta = P.dot3(R, R)[1] / P.dot3(Dp, t)[0]
        elif m == "CG-CD":
            beta = -P.dot3(R, R)[1] / P.dot3(Dp, Rp)[0]
        elif m == "CG-LS":
            beta = P.dot3(R, t)[0] / P.dot3(Dp, Rp)[0]
        else:
            beta = P.dot3(R, R)[1] / P.dot3(Rp, Rp)[1]
        P.lincomb(-1.0, R, beta, Dp, W.p)
        return 1

    # --- orthogonalisation and line search -----------------------------
    @staticmethod
    def _orth_from_sums(sums):
        """OrthogonalNormalization (optimization.py:205-226) from one pass of
        sums: p' = s (p - f phi) with f = <p,phi>/<phi,phi>, |p - f phi|^2 =
        <p,p> - f <p,phi>, s = sqrt(Ne/|p - f phi|^2) (the volume element of
        Ne and of the norm cancels and is left out: ofdft_orth_normalize
        forms the same bits on the device).  Returns theta0, s, f and
        sum V phi p' (the derivative at theta = 0 over 2 dV, :370-380)."""
        pphi, phiphi, pp, vp, vphi = sums
        factor = pphi / phiphi
        pn = pp - factor * pphi
        theta = math.sqrt(pn / phiphi)
        s = math.sqrt(phiphi / pn)
        return theta, s, factor, s * (vp - factor * vphi)

    def _orth_norm(self, W):
        P = W.P
        theta, s, factor, vphip = self._orth_from_sums(P.orth_sums(W.p, W.phi, W.V))
        P.lincomb(s, W.p, -s * factor, W.phi, W.p)
        return theta, vphip

    def _deferred(self):
        """The exact-HVP TN loop on device terms with the unit first trial:
        the orthonormalisation, the first line-search trial and the residual
        norm are queued without waiting for their sums."""
        o = self.optimization_options
        return (bool(o["defer_sums"]) and self.optimization_method == "TN" and o["hvp"] == "exact"
                and not self.host_terms and o["ls_first_step"] != "theta")

    def _flush_res(self, st, raw=None):
        """Complete the log line of the step whose residual norm was left on
        the device (hvp_prepare_async): raw = its sum res^2 when the caller
        already has it from the same plan's deferred_sums()."""
        pend = st.res_pending
        if pend is None:
            return
        st.res_pending = None
        P, entry = pend
        if raw is None:
            raw = P.deferred_sums()[0]
        st.resN = _finite(raw * P.dV, "residual norm")
        entry["dP"] = st.resN
        self._print("{:<8d}{:<24.12E}{:<16.6E}{:<16.6E}{:<8d}{:<8d}{:<16.6E}".format(
            entry["step"], entry["energy"], entry["dE"], st.resN, entry["Nd"], entry["Nls"], entry["time"]))

    def _line_search(self, W, E0, g0, theta, maxiter=100, first=None):
        """first = (t, sums): the first trial, already evaluated."""
        o = self.optimization_options
        cache = {}
        count = [0]

        def phi_(t):
            count[0] += 1
            if first is not None and count[0] == 1 and t == first[0]:
                sums = first[1]
            else:
                sums = self._eval(W, W.phi, W.p, t)
            cache["t"], cache["sums"] = t, sums
            return self._energy(W, sums)

        def dphi_(t):
            if cache.get("t") != t:
                phi_(t)
            return float(cache["sums"][L.S_DTHETA])

        a1 = theta if o["ls_first_step"] == "theta" else 1.0
        if first is not None and first[0] != a1:
            first = None
        if not 0 < o["c1"] < o["c2"] < 1:
            # scipy's scalar_search_wolfe1 check (the reference's call path)
            raise ValueError("'c1' and 'c2' do not satisfy '0 < c1 < c2 < 1'.")
        stp, task = LS.search(phi_, dphi_, E0, g0, a1, c1=o["c1"], c2=o["c2"], xtol=o["xtol"], stpmin=0.0,
                              stpmax=np.pi, maxiter=maxiter)
        return stp, count[0], cache

    # --- precision switching ------------------------------------------
    def _work(self, prec):
        grid = self.rho.grid
        dt = torch.float64 if prec == "fp64" else torch.float32
        # the evaluator changes with the grid and with restarted ions
        if self.host_terms:
            hk = (prec, tuple(grid.nrR), grid.lattice.tobytes())
            if hk not in self._host_evs:
                self._host_evs[hk] = HostTermsEvaluator(self.EnergyEvaluator, grid, dt)
            ev = self._host_evs[hk]
        else:
            ev = self.EnergyEvaluator.evaluator(grid, dt)
        key = (prec, id(ev))
        if key not in self._works:
            self._works = {k: w for k, w in self._works.items() if k[0] != prec}
            self._works[key] = _Work(ev)
        ev.P.set_graph(bool(self.optimization_options.get("tn_graph")))
        ev.P.set_cg_spectra(self.optimization_options["cg_spectra"])
        return self._works[key]

    def _residual(self, W, mu, defer=False):
        """(V - mu) phi into W.res; with the exact TN also the HVP diagonal.
        defer: the norm is left on the device (None returned; the plan's
        deferred_sums() gives it)."""
        if defer:
            if W.W0 is None:
                W.W0 = W.e()
            W.P.hvp_prepare_async(W.phi, W.V, W.aux, mu, W.res, W.W0)
            return None
        if self.optimization_method == "TN" and self.optimization_options["hvp"] == "exact":
            if W.W0 is None:
                W.W0 = W.e()
            return W.P.hvp_prepare(W.phi, W.V, W.aux, mu, W.res, W.W0) * W.P.dV
        return W.P.residual(W.V, W.phi, mu, W.res) * W.P.dV

    def _start(self, W):
        """Evaluate at W.phi and set V, aux, mu, residual."""
        sums = self._eval(W, W.phi, None, 0.0)
        W.V, W.Vt = W.Vt, W.V
        W.aux, W.auxt = W.auxt, W.aux
        E = self._energy(W, sums)
        mu = sums[L.S_VRHO] / sums[L.S_RHO]
        resN = self._residual(W, mu)
        return E, mu, resN, sums

    # ------------------------------------------------------------------
    def optimize_rho(self, guess_rho=None, guess_phi=None, lphi=False):
        """optimize_rho (optimization.py:283-525): the outer loop on phi."""
        self._begin(guess_rho, guess_phi, lphi)
        st = self._st
        for it in range(1, self.optimization_options["maxiter"]):
            nd = self._direction(st)
            if self._advance(st, it, nd):
                break
        else:
            self._flush_res(st)
            self._print("!WARN: Not converged, but reached max steps")
        return self._end()

    # the outer loop in three parts, so that a batch driver can interleave
    # the iterations of many cells (batch.solve_batch_persistent)
    def _begin(self, guess_rho=None, guess_phi=None, lphi=False):
        if lphi:
            raise AttributeError("lphi=True is not on the device path")
        if guess_rho is None and self.rho is None and guess_phi is None:
            raise AttributeError("Must provide a guess density")
        if guess_rho is not None:
            self.rho = guess_rho
        if self.rho is None:
            self.rho = guess_phi * guess_phi
        o = self.optimization_options
        prec = o["precision"]
        if prec is None:
            prec = "fp64" if self.rho.dtype == torch.float64 else "fp32"
        if prec not in ("fp64", "fp32", "mixed"):
            raise ValueError("precision must be fp64, fp32 or mixed")
        self._works = {}
        self._ii_energy = self._ii()
        st = _State()
        st.phase1 = prec == "mixed" and bool(o.get("mixed_phase1"))
        st.cur = "fp32" if (prec == "fp32" or st.phase1) else "fp64"
        st.W = W = self._work(st.cur)
        # mixed precision: FP64 outer iterations with the TN inner CG in FP32
        st.Wi = None
        if prec == "mixed" and not st.phase1 and self.optimization_method == "TN" and o["hvp"] == "exact":
            st.Wi = self._work("fp32")
        for w in ([W, st.Wi] if st.Wi is not None else [W]):
            w.P.bind_stream(True)
        st.t0 = time.perf_counter()
        with torch._C.DisableTorchFunctionSubclass():
            # device scalars throughout: no host synchronisation before the
            # first evaluation
            if guess_phi is None:
                rho = self.rho.to(torch.float64)
                phi = torch.where(rho > 0, torch.sqrt(torch.where(rho > 0, rho, torch.ones_like(rho))),
                                  torch.full_like(rho, 1e-300))
                phi = phi * torch.sqrt(rho.sum() / (phi * phi).sum())
            else:
                phi = guess_phi.to(torch.float64)
            W.phi.copy_(phi.to(W.P.dtype))
            # electron count fixed in FP64 from the initial phi (optimization.py:305-315);
            # read on the host only where the FP64 constraint needs it
            self._nelec_sum = (phi * phi).sum()
            self._nelec = None
        st.E, st.mu, st.resN, st.sums = self._start(W)
        _finite(st.E, "energy")
        _finite(st.resN, "residual norm")
        st.hist = [st.E]
        self.log = [dict(step=0, energy=st.E, dE=st.E, dP=st.resN, Nd=1, Nls=1, time=time.perf_counter() - st.t0,
                         prec=st.cur)]
        self._print("{:8s}{:24s}{:16s}{:16s}{:8s}{:8s}{:16s}".format("Step", "Energy(a.u.)", "dE", "dP", "Nd", "Nls",
                                                                      "Time(s)"))
        self._print("{:<8d}{:<24.12E}{:<16.6E}{:<16.6E}{:<8d}{:<8d}{:<16.6E}".format(0, st.E, st.E, st.resN, 1, 1,
                                                                                   self.log[-1]["time"]))
        st.theta = 0.1
        st.converged = 2
        st.res_pending = None
        self._lbfgs = dict(s=[], y=[], rho=[])
        self._dir_prev = None
        st.first = True
        st.switch = o["mixed_switch"]
        if st.switch is None:
            st.switch = 1e-8 * abs(st.E)
        self._st = st

    def _direction(self, st):
        """The search direction into st.W.p; returns Nd."""
        m = self.optimization_method
        W = st.W
        if m == "TN":
            return self._direction_tn(W, st.mu, st.Wi if st.cur == "fp64" else None)
        if m == "LBFGS":
            return self._direction_lbfgs(W)
        if m == "GD":
            W.P.lincomb(-1.0, W.res, 0.0, W.res, W.p)
            return 1
        return self._direction_cg(W, st.first)

    def _advance(self, st, it, nd):
        """Orthonormalisation, line search and update of one outer iteration
        (optimization.py:363-516) after the direction; True when the loop
        stops (converged, or the line search failed)."""
        o = self.optimization_options
        m = self.optimization_method
        W = st.W
        defer = self._deferred()
        first = None
        if defer:
            # orthonormalise on the device and evaluate the first trial
            # (theta = 1) behind it; one synchronisation returns the trial's
            # sums, the orthonormalisation sums and the previous residual
            W.P.orth_normalize(W.p, W.phi, W.V)
            first = (1.0, self._eval(W, W.phi, W.p, 1.0))
            d6 = W.P.deferred_sums()
            if st.res_pending is not None and st.res_pending[0] is W.P:
                self._flush_res(st, d6[0])
            theta0, _, _, vphip = self._orth_from_sums(d6[1:])
        else:
            theta0, vphip = self._orth_norm(W)
        if vphip * 2.0 > 0:
            self._print("!WARN: Change to steepest decent")
            W.P.lincomb(-1.0, W.res, 0.0, W.res, W.p)
            theta0, vphip = self._orth_norm(W)
            first = None
        st.theta = min(theta0, st.theta)
        g0 = vphip * 2.0 * W.P.dV
        # an FP32 line search that cannot converge is cut short: the outer
        # loop then continues in FP64
        stp, nls, cache = self._line_search(W, st.E, g0, st.theta, 8 if (st.phase1 and st.cur == "fp32") else 100,
                                            first)
        failed = stp is None
        if failed and not (st.phase1 and st.cur == "fp32"):
            self._flush_res(st)
            st.converged = 1
            self._print("!!!ERROR : Line-Search Failed!!!")
            self._print("!!!ERROR : Density Optimization NOT Converged  !!!")
            return True
        if not failed:
            st.theta = stp
            t = cache["t"]
            sums = cache["sums"]
            old_phi = None
            if m == "LBFGS":
                old_phi = W.phi.clone()
            W.P.lincomb(math.cos(t), W.phi, math.sin(t), W.p, W.phi)
            W.V, W.Vt = W.Vt, W.V
            W.aux, W.auxt = W.auxt, W.aux
            Enew = _finite(self._energy(W, sums), "energy")
            st.mu = sums[L.S_VRHO] / sums[L.S_RHO]
            W.res_prev.copy_(W.res)
            self._flush_res(st)
            st.resN = self._residual(W, st.mu, defer=defer)
            if st.resN is not None:
                _finite(st.resN, "residual norm")
            if m == "LBFGS":
                self._lbfgs_update(W, old_phi)
            if m.startswith("CG"):
                if self._dir_prev is None:
                    self._dir_prev = W.e()
                self._dir_prev.copy_(W.p)
            st.first = False
            st.hist.append(Enew)
            st.E = Enew
            hist = st.hist
            self.log.append(dict(step=it, energy=st.E, dE=hist[-1] - hist[-2], dP=st.resN, Nd=nd, Nls=nls,
                                 time=time.perf_counter() - st.t0, prec=st.cur))
            if st.resN is None:
                # printed once the residual norm is read (_flush_res)
                st.res_pending = (W.P, self.log[-1])
            else:
                self._print("{:<8d}{:<24.12E}{:<16.6E}{:<16.6E}{:<8d}{:<8d}{:<16.6E}".format(
                    it, st.E, hist[-1] - hist[-2], st.resN, nd, nls, self.log[-1]["time"]))
        # mixed precision: leave FP32 near its noise floor
        if st.phase1 and st.cur == "fp32" and (failed or (len(st.hist) > 1 and
                                                          abs(st.hist[-1] - st.hist[-2]) < st.switch)):
            self._flush_res(st)
            W32 = W
            st.cur = "fp64"
            st.W = W = self._work("fp64")
            W.P.bind_stream(True)
            st.Wi = self._work("fp32") if (m == "TN" and o["hvp"] == "exact") else None
            W.P.convert_from(W32.phi, W.phi)
            # FP64 electron-count constraint: FP32 rotations drift the norm
            nphi = W.P.dot3(W.phi, W.phi)[1] * W.P.dV
            if self._nelec is None:
                self._nelec = float(self._nelec_sum) * W.P.dV
            W.P.lincomb(math.sqrt(self._nelec / nphi), W.phi, 0.0, W.phi, W.phi)
            E32 = st.hist[-1]
            st.E, st.mu, st.resN, st.sums = self._start(W)
            _finite(st.E, "energy")
            # the convergence test restarts on FP64 energies
            st.hist = [st.E]
            st.theta = 0.1
            st.first = True
            self._lbfgs = dict(s=[], y=[], rho=[])
            self._dir_prev = None
            self.log.append(dict(step=it, energy=st.E, dE=st.E - E32, dP=st.resN, Nd=0, Nls=0,
                                 time=time.perf_counter() - st.t0, prec="switch->fp64"))
            self._print("#### FP32 -> FP64 at step %d, E = %.12E" % (it, st.E))
            return False
        if self._check_converge(st.hist):
            self._flush_res(st)
            st.converged = 0
            self._print("#### Density Optimization Converged ####")
            return True
        return False

    def _end(self):
        st = self._st
        self._flush_res(st)
        for w in self._works.values():
            w.P.bind_stream(False)
        self.mu = st.mu
        self._print("Chemical potential (a.u.):", st.mu)
        self._print("Chemical potential (eV)  :", st.mu * HARTREE_EV)
        grid = self.rho.grid
        W = st.W
        with torch._C.DisableTorchFunctionSubclass():
            phi_out = W.phi.clone()
            rho_out = phi_out * phi_out
            Vout = W.V.clone()
        self.phi = DirectField(grid=grid, data=phi_out, dtype=phi_out.dtype)
        self.rho = DirectField(grid=grid, data=rho_out, dtype=rho_out.dtype)
        self.functional = FunctionalOutput(name="TOTAL", energy=st.E,
                                           potential=DirectField(grid=grid, data=Vout, dtype=Vout.dtype))
        self.energy = st.E
        self.converged = st.converged
        self.elapsed = time.perf_counter() - st.t0
        return self.rho

    def _lbfgs_update(self, W, old_phi):
        """LBFGS.update (math_utils.py:315-325): keep Bound+1 pairs."""
        lb = self._lbfgs
        P = W.P
        if len(lb["s"]) > self.optimization_options["maxcor"]:
            lb["s"].pop(0)
            lb["y"].pop(0)
            lb["rho"].pop(0)
        s = W.e()
        y = W.e()
        P.lincomb(1.0, W.phi, -1.0, old_phi, s)
        P.lincomb(1.0, W.res, -1.0, W.res_prev, y)
        lb["s"].append(s)
        lb["y"].append(y)
        lb["rho"].append(1.0 / P.dot3(y, s)[0])

    def _check_converge(self, hist):
        """check_converge (optimization.py:527-540)."""
        econv = self.optimization_options["econv"]
        nc = self.optimization_options["ncheck"]
        E = hist[-1]
        if econv is not None:
            if len(hist) - 1 < nc:
                return False
            for i in range(nc):
                if abs(hist[-2 - i] - E) > econv:
                    return False
        return True

    def get_chemical_potential(self, potential, rho, **kwargs):
        return (potential * rho).integral() / rho.integral()

    def __call__(self, guess_rho=None, calcType=("E", "V"), guess_phi=None, lphi=False):
        return self.optimize_rho(guess_rho=guess_rho, guess_phi=guess_phi, lphi=lphi)
