"""GPU density optimization against the reference's saved notebook log
(examples/notebooks/density_optimization.ipynb:174-183) and the oracle's TN.

Tolerances: per-iteration energies 1e-9 Ha against the notebook's 12
printed digits; final energies 1e-9 Ha (FP64); mixed precision within
1e-6 eV/atom of the FP64 solve.
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import workloads as W
from tests.conftest import DATA

pytestmark = pytest.mark.gpu

NOTEBOOK_E = [-8.090977710718, -8.273226052119, -8.280858141000, -8.281101062862, -8.281135978745,
              -8.281138625688, -8.281138996216]
NOTEBOOK_ND = [2, 7, 5, 6, 5, 8]


def _notebook_system(kedf="TFvW"):
    a = 4.05 / W.BOHR
    cell = np.eye(3) * a
    ions = D.Ions(["Al"] * 4, W.FCC_BASIS * a, cell)
    nr = D.ecut2nr(ecut=35, lattice=cell)
    grid = D.DirectGrid(lattice=cell, nr=nr)
    pseudo = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": os.path.join(DATA, "Al_lda.oe01.recpot")})
    KE = D.Functional(type="KEDF", name=kedf)
    XC = D.Functional(type="XC", name="LDA")
    HARTREE = D.Functional(type="HARTREE")
    ev = D.TotalFunctional(KE=KE, XC=XC, HARTREE=HARTREE, PSEUDO=pseudo)
    rho = D.DirectField(grid=grid)
    rho[:] = ions.get_ncharges() / grid.volume
    return ions, grid, ev, rho


def test_notebook_trajectory_fd_hvp():
    """Reference-exact mode (forward-difference HVP, first trial = theta)
    reproduces every logged iteration of the notebook run."""
    ions, grid, ev, rho = _notebook_system()
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-6 * ions.nat, "hvp": "fd", "ls_first_step": "theta"})
    opt.optimize_rho(guess_rho=rho)
    E = [l["energy"] for l in opt.log]
    Nd = [l["Nd"] for l in opt.log[1:]]
    print(E, Nd)
    assert Nd == NOTEBOOK_ND
    assert np.allclose(E, NOTEBOOK_E, rtol=0, atol=2e-9)
    assert abs(opt.mu - 0.30115196532053357) < 1e-8
    assert abs(ev.Energy(opt.rho) - (-8.2811389962163)) < 1e-9


def test_notebook_exact_hvp_converges_to_same_minimum():
    ions, grid, ev, rho = _notebook_system()
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-10})
    opt.optimize_rho(guess_rho=rho)
    print(opt.log[-1])
    assert opt.converged == 0
    assert abs(opt.energy - (-8.2811389962163)) < 5e-9


def _al4_pair():
    wl = W.build("al4_32")
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {"Al": O.empty_core_table(3.0, 1.12)}, PME=True)
    oew = O.Ewald(og, oi, PME=True)
    otot = O.TotalFunctional(og, kedf="TFVW", kedf_options=dict(x=1.0, y=1.0 / 9.0), pseudo=opp, ewald=oew)
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(3.0, 1.12)})
    ev = D.TotalFunctional(KEDF=D.KEDF("TFVW", x=1.0, y=1.0 / 9.0), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    return wl, og, otot, grid, ev


def test_al4_32_tn_matches_oracle():
    wl, og, otot, grid, ev = _al4_pair()
    opts = {"econv": 1e-9, "maxiter": 200}
    oopt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "TN", dict(opts))
    oopt.optimize_rho(wl["rho_ini"])
    for hvp in ("fd", "exact"):
        rho = D.DirectField(grid=grid, data=wl["rho_ini"])
        opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                             optimization_options=dict(opts, hvp=hvp))
        opt.optimize_rho(guess_rho=rho)
        print(hvp, opt.energy, oopt.energy, len(opt.log), len(oopt.log))
        assert abs(opt.energy - oopt.energy) < 1e-9
        if hvp == "fd":
            n = min(4, len(oopt.log))
            for a, b in zip(opt.log[:n], oopt.log[:n]):
                assert abs(a["energy"] - b[1]) < 1e-8


def test_al4_32_lbfgs_and_cg_reach_tn_minimum():
    wl, og, otot, grid, ev = _al4_pair()
    ref = None
    for m in ("TN", "LBFGS", "CG-HS"):
        rho = D.DirectField(grid=grid, data=wl["rho_ini"])
        opt = D.Optimization(EnergyEvaluator=ev, optimization_method=m,
                             optimization_options={"econv": 1e-10, "maxiter": 400})
        opt.optimize_rho(guess_rho=rho)
        print(m, opt.energy, len(opt.log), opt.converged)
        if ref is None:
            ref = opt.energy
        assert abs(opt.energy - ref) < 1e-8


def test_al4_32_lbfgs_trajectory_matches_oracle():
    """LBFGS (optimization.py:150-173, math_utils.py:307-325) iteration by
    iteration against the oracle's restatement, both FP64: the first loop
    of the two-loop recursion skips index 0 and the memory keeps maxcor + 1
    pairs in both, so the energies, line-search counts and iteration
    counts agree."""
    print("case LBFGS trajectory al4_32", flush=True)
    wl, og, otot, grid, ev = _al4_pair()
    opts = {"econv": 1e-10, "maxiter": 400}
    oopt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "LBFGS", dict(opts))
    oopt.optimize_rho(wl["rho_ini"])
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="LBFGS", optimization_options=dict(opts))
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    print("GPU %d iterations, oracle %d; final %.12f vs %.12f" % (len(opt.log) - 1, len(oopt.log) - 1, opt.energy,
                                                                 oopt.energy))
    assert opt.converged == 0 and oopt.converged == 0
    n = min(len(opt.log), len(oopt.log))
    for a, b in zip(opt.log[:n], oopt.log[:n]):
        print("%3d %.12f %.12f %d %d" % (a["step"], a["energy"], b[1], a["Nls"], b[5]))
    for a, b in zip(opt.log[:12], oopt.log[:12]):
        assert abs(a["energy"] - b[1]) < 1e-9
        assert a["Nls"] == b[5]
    assert abs(len(opt.log) - len(oopt.log)) <= 2
    assert abs(opt.energy - oopt.energy) < 1e-9
    # memory of maxcor + 1 pairs (the reference pops only above Bound)
    assert len(opt._lbfgs["s"]) == min(opts.get("maxcor", 6) + 1, len(opt.log) - 1)
    assert len(oopt.lb_s) == min(7, len(oopt.log) - 1)


@pytest.mark.parametrize("phase1", [False, True])
def test_al4_32_mixed_precision_within_1e6_ev_per_atom(phase1):
    wl, og, otot, grid, ev = _al4_pair()
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    o64 = D.Optimization(EnergyEvaluator=ev, optimization_method="TN", optimization_options={"econv": 1e-9})
    o64.optimize_rho(guess_rho=rho)
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    omx = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-9, "precision": "mixed", "mixed_phase1": phase1})
    omx.optimize_rho(guess_rho=rho)
    err = abs(omx.energy - o64.energy) * W.HARTREE / 4
    print("mixed", phase1, omx.energy, o64.energy, err, [l["prec"] for l in omx.log])
    assert omx.converged == 0
    assert err < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tn_graph_replays_give_the_same_solve(prec):
    """The device TN loop with each CG iteration replayed from one captured
    hipGraph (many replays per solve, with the optimizer's own launches and
    host reads between them) reproduces the launch-mode solve exactly."""
    print("case TN graph mode al108_96", prec, flush=True)
    wl = W.build("al108_96")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(3.0, 1.12)})
    ev = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    res = {}
    for graph in (False, True):
        opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                             optimization_options={"econv": 1e-9, "precision": prec, "tn_graph": graph,
                                                   "mixed_phase1": True})
        opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
        res[graph] = opt
        print(graph, opt.energy, [l["Nd"] for l in opt.log])
        assert opt.converged == 0
    assert res[True].energy == res[False].energy
    assert [l["Nd"] for l in res[True].log] == [l["Nd"] for l in res[False].log]
    assert [l["energy"] for l in res[True].log] == [l["energy"] for l in res[False].log]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_deferred_sums_give_the_same_solve(prec, capsys):
    """The orthonormalisation on the device with the first line-search trial
    queued behind it, and the residual norm read one step later
    (defer_sums=True), reproduce the solve that waits for each sum: the same
    energies, residual norms, inner and line-search counts at every step, and
    the same printed log lines."""
    print("case deferred sums al108_96", prec, flush=True)
    wl = W.build("al108_96")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(3.0, 1.12)})
    ev = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    res, out = {}, {}
    for defer in (False, True):
        opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                             optimization_options={"econv": 1e-9, "precision": prec, "defer_sums": defer,
                                                   "mixed_phase1": True})
        capsys.readouterr()
        opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
        # the time column aside, the printed lines must agree
        out[defer] = [ln.split()[:6] for ln in capsys.readouterr().out.splitlines() if ln[:1].isdigit()]
        res[defer] = opt
        assert opt.converged == 0
    a, b = res[False], res[True]
    assert b.energy == a.energy
    for key in ("energy", "dE", "dP", "Nd", "Nls", "prec"):
        assert [l[key] for l in b.log] == [l[key] for l in a.log], key
    assert out[True] == out[False] and len(out[True]) == len(a.log) - (1 if prec == "mixed" else 0)
    assert b.nfeval == a.nfeval
