"""Parity at the benchmark configurations (FP64 bar: per-term energies to
1e-10 relative, potentials to 1e-9 Ha max-abs, against the oracle on the
same seeded inputs):

* fcc Al 3x3x3 (108 atoms), 96^3, WT + LDA + Hartree + PP: one evaluation,
  and the converged TN energy against the oracle's TN (econv 1e-9 Ha);
* fcc Al 6x6x6 (864 atoms), 192^3: one FP64 evaluation against the oracle,
  and the FP32 evaluation's error per term, for V and for the total (eV/atom).
"""
import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("KEDF-TF", "KEDF-VW", "KEDF-NL", "XC", "HARTREE", "PSEUDO", "II")


def _pair(name):
    wl = W.build(name)
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={s: D.empty_core(*W.EMPTY_CORE[s]) for s in wl["tables"]})
    tot = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {s: O.empty_core_table(*W.EMPTY_CORE[s]) for s in wl["tables"]}, PME=True)
    otot = O.TotalFunctional(og, kedf="WT", pseudo=opp, ewald=O.Ewald(og, oi, PME=True))
    return wl, grid, tot, og, otot


def _check_eval(wl, grid, tot, otot):
    E, V, parts = otot.compute(wl["rho_ini"], split=True)
    out = tot.get_energy_potential(D.DirectField(grid=grid, data=wl["rho_ini"]), split=True)
    for k in KEYS:
        print("%-8s GPU %.14e  oracle %.14e" % (k, out[k].energy, parts[k]))
        assert abs(out[k].energy - parts[k]) <= 1e-10 * max(abs(parts[k]), 1e-3), k
    full = tot.compute(D.DirectField(grid=grid, data=wl["rho_ini"]))
    dv = float((full.potential.as_subclass(torch.Tensor) - torch.tensor(V, device="cuda")).abs().max())
    print("max |V - V_oracle| = %.3e Ha" % dv)
    assert dv < 1e-9
    assert abs(full.energy - E) <= 1e-10 * abs(E)
    return parts


def test_al108_96_fp64_evaluation_matches_oracle():
    print("case al108_96 evaluation", flush=True)
    wl, grid, tot, og, otot = _pair("al108_96")
    _check_eval(wl, grid, tot, otot)


def test_al108_96_fp64_tn_solve_matches_oracle():
    print("case al108_96 TN solve", flush=True)
    wl, grid, tot, og, otot = _pair("al108_96")
    opts = {"econv": 1e-9, "maxiter": 100}
    oopt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "TN", dict(opts))
    oopt.optimize_rho(wl["rho_ini"])
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options=dict(opts))
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    nat = len(wl["symbols"])
    err = abs(opt.energy - oopt.energy) * W.HARTREE / nat
    print("GPU %.12f oracle %.12f Ha: %.3e eV/atom; %d vs %d iterations" % (opt.energy, oopt.energy, err,
                                                                          len(opt.log) - 1, len(oopt.log) - 1))
    assert opt.converged == 0
    assert abs(opt.energy - oopt.energy) < 5e-9


def test_al864_192_fp64_evaluation_and_fp32_error():
    print("case al864_192 evaluation FP64 and FP32", flush=True)
    wl, grid, tot, og, otot = _pair("al864_192")
    parts = _check_eval(wl, grid, tot, otot)
    nat = len(wl["symbols"])
    r32 = D.DirectField(grid=grid, data=wl["rho_ini"], dtype=torch.float32)
    o32 = tot.get_energy_potential(r32, split=True)
    tot32 = 0.0
    for k in KEYS:
        err = abs(o32[k].energy - parts[k]) * W.HARTREE / nat
        tot32 += o32[k].energy
        print("FP32 %-8s error %.3e eV/atom" % (k, err))
        assert err < 1e-5
    E = sum(parts[k] for k in KEYS)
    print("FP32 total error %.3e eV/atom" % (abs(tot32 - E) * W.HARTREE / nat))
    assert abs(tot32 - E) * W.HARTREE / nat < 1e-5


def _golden(name):
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_converged.json")) as fh:
        return json.load(fh)[name]


def test_al864_192_converged_energy_against_oracle_minimum():
    """The 864-atom 192^3 solve, mixed precision (FP32 outer phase, FP64
    refinement) and FP64, against the oracle's converged TN minimum
    (tests/golden/oracle_converged.json, made by make_converged.py in the
    build container): within 1e-6 eV/atom.  At the mixed solve's converged
    density, the FP64 per-term energies and V against the oracle evaluated
    there (1e-10 relative, 1e-9 Ha), and the FP32 evaluation's error there."""
    print("case al864_192 converged TN vs oracle minimum", flush=True)
    gold = _golden("al864_192")
    wl, grid, tot, og, otot = _pair("al864_192")
    nat = len(wl["symbols"])
    res = {}
    for prec, phase1 in (("mixed", True), ("fp64", False)):
        opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN",
                             optimization_options={"econv": 1e-8, "precision": prec, "mixed_phase1": phase1})
        opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
        err = abs(opt.energy - gold["energy_Ha"]) * W.HARTREE / nat
        print("%s: E %.10f Ha, oracle minimum %.10f Ha, error %.3e eV/atom" % (prec, opt.energy, gold["energy_Ha"], err))
        assert opt.converged == 0
        assert err < 1e-6
        res[prec] = opt
    rho = res["mixed"].rho
    r = rho.as_subclass(torch.Tensor).double().cpu().numpy()
    E, V, parts = otot.compute(r, split=True)
    out = tot.get_energy_potential(D.DirectField(grid=grid, data=r), split=True)
    o32 = tot.get_energy_potential(D.DirectField(grid=grid, data=r, dtype=torch.float32), split=True)
    for k in KEYS:
        e32 = abs(o32[k].energy - parts[k]) * W.HARTREE / nat
        print("%-8s FP64 %.14e oracle %.14e   FP32 error %.3e eV/atom" % (k, out[k].energy, parts[k], e32))
        assert abs(out[k].energy - parts[k]) <= 1e-10 * max(abs(parts[k]), 1e-3), k
    full = tot.compute(D.DirectField(grid=grid, data=r))
    dv = float((full.potential.as_subclass(torch.Tensor) - torch.tensor(V, device="cuda")).abs().max())
    f32 = tot.compute(D.DirectField(grid=grid, data=r, dtype=torch.float32))
    dv32 = float((f32.potential.as_subclass(torch.Tensor).double() - torch.tensor(V, device="cuda")).abs().max())
    print("at the converged density: max|V - V_oracle| FP64 %.3e Ha, FP32 %.3e Ha; total FP64 %.3e, FP32 %.3e eV/atom"
          % (dv, dv32, abs(full.energy - E) * W.HARTREE / nat, abs(f32.energy - E) * W.HARTREE / nat))
    assert dv < 1e-9
    assert abs(full.energy - E) <= 1e-10 * abs(E)
