"""The reference's plugin and evaluator contracts on the device path:

* any AbstractFunctional joins a TotalFunctional (total_functional.py:69-77)
  and is summed with the fused device terms (here an external harmonic
  potential written by the test, as a user would write one);
* split=True returns every term's FunctionalOutput with its own potential
  (total_functional.py:115-131), keyed like the reference;
* calcType 'D' gives the kinetic energy densities (kedf/tf.py:66-75,
  vw.py:125-126, wt.py:123-129); entries the device path does not compute
  raise;
* Optimization accepts any evaluator returning .energy and .potential
  (optimization.py:83-86, :251, :275) and drives it with the reference's
  forward-difference Hessian product on the device vector kernels;
* failure paths: maxiter reached (converged = 2, optimization.py:514-516),
  line-search failure (converged = 1, :444-448), non-finite input (raises).

Bar: energies 1e-10 relative and potentials 1e-9 Ha max-abs against the
oracle (FP64); converged energies 1e-8 Ha against the oracle's TN with the
same term.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import workloads as W
from dftpy_amd.functional import AbstractFunctional, FunctionalOutput
from dftpy_amd.optimization import OptimizationError

pytestmark = pytest.mark.gpu

K_HARM = 0.02


class Harmonic(AbstractFunctional):
    """A user-written term: V(r) = k/2 |r - c|^2 (c the cell centre),
    E = int V rho.  Written with plain torch operations on the DirectField,
    the way a plugin author would."""

    def __init__(self, grid, k=K_HARM):
        self.type = "EXT"
        self.name = "EXT"
        self.vext = torch.tensor(_harmonic_numpy(grid.lattice, grid.nrR, k), device="cuda")

    def compute(self, rho, calcType=("E", "V"), **kwargs):
        v = self.vext.to(rho.dtype)
        out = FunctionalOutput(name="EXT")
        with torch._C.DisableTorchFunctionSubclass():
            out.energy = float((v * rho).sum()) * rho.grid.dV
        if "V" in calcType:
            out.potential = D.DirectField(grid=rho.grid, data=v.clone(), dtype=rho.dtype)
        return out


def _harmonic_numpy(lattice, nr, k):
    nr = tuple(int(n) for n in nr)
    f = np.stack(np.meshgrid(*[np.arange(n) / n for n in nr], indexing="ij"), axis=-1)
    r = (f - 0.5) @ np.asarray(lattice)
    return 0.5 * k * (r ** 2).sum(-1)


def _system(kedf="WT"):
    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(*W.EMPTY_CORE["Al"])})
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {"Al": O.empty_core_table(*W.EMPTY_CORE["Al"])}, PME=True)
    otot = O.TotalFunctional(og, kedf=kedf, pseudo=opp, ewald=O.Ewald(og, oi, PME=True))
    return wl, grid, pp, og, otot


def test_plugin_term_energy_potential_and_split():
    print("case plugin term al4_32 WT", flush=True)
    wl, grid, pp, og, otot = _system()
    ext = Harmonic(grid)
    tot = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp, EXT=ext)
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    vext = _harmonic_numpy(og.lattice, og.nr, K_HARM)
    E0, V0, parts0 = otot.compute(wl["rho_ini"], split=True)
    Eext = float((vext * wl["rho_ini"]).sum() * og.dV)
    out = tot.compute(rho)
    print("E", out.energy, E0 + Eext)
    assert abs(out.energy - (E0 + Eext)) <= 1e-10 * abs(E0 + Eext)
    dv = float((out.potential.as_subclass(torch.Tensor) - torch.tensor(V0 + vext, device="cuda")).abs().max())
    assert dv < 1e-9, dv
    # split: each term with its own potential
    sp = tot.get_energy_potential(rho, split=True)
    assert set(sp) == {"TOTAL", "KEDF-TF", "KEDF-VW", "KEDF-NL", "XC", "HARTREE", "PSEUDO", "EXT", "II"}
    r = wl["rho_ini"]
    cache = O.KernelCache()
    oV = {"KEDF-TF": O.tf_functional(r, og)[1], "KEDF-VW": O.vw_functional(r, og)[1],
          "KEDF-NL": O.wt_functional(r, og, cache=cache)[1], "XC": O.lda_functional(r, og)[1],
          "HARTREE": O.hartree_functional(r, og)[1], "PSEUDO": otot.pseudo.compute(r)[1], "EXT": vext}
    for k, v in oV.items():
        e = parts0.get(k, Eext)
        got = sp[k]
        dv = float((got.potential.as_subclass(torch.Tensor) - torch.tensor(v, device="cuda")).abs().max())
        print("%-8s E %.14e (oracle %.14e)  max|dV| %.2e" % (k, got.energy, e, dv))
        assert abs(got.energy - e) <= 1e-10 * max(abs(e), 1e-3), k
        assert dv < 1e-9, k
    assert abs(sp["II"].energy - parts0["II"]) < 1e-12 * abs(parts0["II"])
    assert abs(sp["TOTAL"].energy - out.energy) < 1e-10 * abs(out.energy)
    vs = sum(sp[k].potential for k in oV)
    assert float((vs - sp["TOTAL"].potential).abs().max()) == 0.0
    # KEDF split keys are the engine names (kedf/__init__.py:158-172)
    ke = D.KEDF("WT").compute(rho, split=True)
    assert set(ke) == {"TF", "VW", "WT-NL"}


def test_energy_density_and_unsupported_calctype():
    print("case calcType D al4_32 WT", flush=True)
    wl, grid, pp, og, otot = _system()
    tot = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    r = wl["rho_ini"]
    out = tot.get_energy_potential(rho, calcType={"E", "V", "D"}, split=True)
    ctf = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)
    cache = O.KernelCache()
    ref = {"KEDF-TF": ctf * np.cbrt(r ** 5), "KEDF-VW": O.vw_functional(r, og)[1] * r,
           "KEDF-NL": O.wt_functional(r, og, cache=cache)[1] * r / (2 * 5.0 / 6.0)}
    for k, v in ref.items():
        ed = out[k].energydensity.as_subclass(torch.Tensor).cpu().numpy()
        err = np.abs(ed - v).max()
        print(k, "max |ed - oracle| %.2e, int ed dV %.12f vs E %.12f" % (err, ed.sum() * og.dV, out[k].energy))
        assert err < 1e-9 * np.abs(v).max()
        assert abs(ed.sum() * og.dV - out[k].energy) < 1e-10 * abs(out[k].energy)
    assert not hasattr(out["XC"], "energydensity")  # the in-house LDA has none (semilocal_xc.py:283-341)
    tot_d = tot.compute(rho, calcType={"E", "D"})
    s = sum(ref.values())
    assert np.abs(tot_d.energydensity.as_subclass(torch.Tensor).cpu().numpy() - s).max() < 1e-9 * np.abs(s).max()
    with pytest.raises(AttributeError):
        tot.compute(rho, calcType={"E", "V2"})


def _oracle_tn_with_ext(otot, og, vext, rho_ini, econv):
    def ev(r, c, p):
        E, V = otot.compute(r, c, phi=p)
        if E is not None:
            E = E + float((vext * r).sum() * og.dV)
        return E, (None if V is None else V + vext)
    opt = O.Optimization(ev, og, "TN", {"econv": econv})
    opt.optimize_rho(rho_ini)
    return opt


def test_tn_with_plugin_term_matches_oracle():
    print("case TN with plugin term al4_32 TFvW", flush=True)
    wl, grid, pp, og, otot = _system(kedf="TFVW")
    ext = Harmonic(grid)
    tot = D.TotalFunctional(KEDF=D.KEDF("TFVW", y=1.0 / 9.0), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp,
                            EXT=ext)
    otot.opts["y"] = 1.0 / 9.0
    vext = _harmonic_numpy(og.lattice, og.nr, K_HARM)
    oopt = _oracle_tn_with_ext(otot, og, vext, wl["rho_ini"], 1e-9)
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options={"econv": 1e-9})
    assert opt.optimization_options["hvp"] == "fd"
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    print("GPU %.12f oracle %.12f (%d vs %d iterations)" % (opt.energy, oopt.energy, len(opt.log) - 1,
                                                            len(oopt.log) - 1))
    assert opt.converged == 0
    assert abs(opt.energy - oopt.energy) < 1e-8
    # the same forward-difference TN: the outer iteration count agrees (the
    # inner CG counts can differ by one here and there: eps = 1e-9 magnifies
    # the last-bit differences of the two FP64 evaluations)
    print("Nd GPU", [l["Nd"] for l in opt.log[1:]], "oracle", [l[4] for l in oopt.log[1:]])
    assert abs(len(opt.log) - len(oopt.log)) <= 1


def test_any_evaluator_callable():
    """An evaluator that is not a TotalFunctional (a plain function returning
    .energy/.potential) drives the same TN as the fused path."""
    print("case callable evaluator al4_32 TFvW", flush=True)
    wl, grid, pp, og, otot = _system(kedf="TFVW")
    tot = D.TotalFunctional(KEDF=D.KEDF("TFVW", y=1.0 / 9.0), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    calls = []

    def evaluator(rho, calcType=("E", "V"), **kwargs):
        calls.append(sorted(calcType))
        return tot.compute(rho, calcType=calcType)

    opt = D.Optimization(EnergyEvaluator=evaluator, optimization_method="TN", optimization_options={"econv": 1e-9})
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    ref = D.Optimization(EnergyEvaluator=tot, optimization_method="TN",
                         optimization_options={"econv": 1e-9, "hvp": "fd"})
    ref.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    print("callable %.12f fused %.12f, %d calls" % (opt.energy, ref.energy, len(calls)))
    assert opt.converged == 0 and len(calls) > 10
    # the callable's energy includes the ion-ion term (TotalFunctional.compute)
    assert abs(opt.energy - ref.energy) < 1e-9
    with pytest.raises(ValueError):
        D.Optimization(EnergyEvaluator=evaluator, optimization_options={"precision": "fp32"})
    with pytest.raises(TypeError):
        D.Optimization(EnergyEvaluator=42)


def _plain_total(grid, pp):
    return D.TotalFunctional(KEDF=D.KEDF("TFVW", y=1.0 / 9.0), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)


def test_failure_paths():
    print("case failure paths al4_32", flush=True)
    wl, grid, pp, og, otot = _system(kedf="TFVW")
    tot = _plain_total(grid, pp)
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    # maxiter reached: converged = 2 (optimization.py:514-516)
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN",
                         optimization_options={"econv": 1e-12, "maxiter": 3})
    opt.optimize_rho(guess_rho=rho)
    assert opt.converged == 2 and len(opt.log) == 3
    # line-search failure: xtol = 1 ends the search once the first trial
    # (theta = 1) brackets the minimum -- converged = 1 (:444-448)
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options={"xtol": 1.0})
    opt.optimize_rho(guess_rho=rho)
    assert opt.converged == 1
    # c1 >= c2 is refused as scipy's line_search_wolfe1 refuses it
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options={"c1": 0.5, "c2": 0.2})
    with pytest.raises(ValueError):
        opt.optimize_rho(guess_rho=rho)
    # a NaN in the density raises instead of iterating
    bad = wl["rho_ini"].copy()
    bad[3, 4, 5] = np.nan
    for method in ("TN", "LBFGS"):
        opt = D.Optimization(EnergyEvaluator=tot, optimization_method=method)
        with pytest.raises(OptimizationError):
            opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=bad))


def test_fp32_forward_difference_runs_in_fp64():
    """A plugin term forces the forward-difference product (eps = 1e-9,
    below FP32 resolution); in FP32 and in the FP32 phase of mixed precision
    it runs on FP64 copies of phi with an FP64 residual, so the FP32 solve
    reaches the FP64 minimum within the FP32 energy noise and the mixed solve
    reaches it to 1e-6 eV/atom."""
    print("case plugin TN fp32 / mixed al4_32 WT", flush=True)
    wl, grid, pp, og, otot = _system()
    tot = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp, EXT=Harmonic(grid))
    res = {}
    for prec, econv in (("fp64", 1e-9), ("fp32", 5e-5), ("mixed", 1e-9)):
        opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN",
                             optimization_options={"econv": econv, "precision": prec, "mixed_phase1": True})
        opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
        print(prec, opt.energy, opt.converged, len(opt.log) - 1)
        assert opt.converged == 0
        res[prec] = opt.energy
    assert abs(res["fp32"] - res["fp64"]) < 1e-4
    assert abs(res["mixed"] - res["fp64"]) * W.HARTREE / 4 < 1e-6
    # GGA / LWT kinetic terms are FP64-only on this path
    tg = D.TotalFunctional(KEDF=D.KEDF("GGA", k_str="LKT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    for prec in ("fp32", "mixed"):
        with pytest.raises(ValueError):
            D.Optimization(EnergyEvaluator=tg, optimization_options={"precision": prec})
