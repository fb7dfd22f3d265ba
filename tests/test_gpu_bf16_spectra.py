"""Spectrum layouts of the device TN loop's Hessian products
(ofdft_plan_set_cg_spectra, Optimization option cg_spectra): split rows in
the plan precision (the default) against the padded rows, and bf16 split
rows (cg_spectra='bf16').

The spectra between the transform passes of the exact Hessian product are
stored as bf16 pairs (n2/2 columns per row, the Nyquist column in a plane of
its own) while every transform, multiplier and pointwise step computes in
FP32.  What the tests hold it to:

* one Hessian product of the bf16 loop against the FP32 product of the same
  plan (the host-visible Ap = (r0 - r1) / alpha of one CG step): relative L2
  error below 2e-2 (bf16 keeps 8 significant bits, 2^-9 = 2e-3 relative per
  stored value; a product passes four stored spectra), and p.Ap, a sum over
  the spectrum, to 1e-3 relative -- on cubic and non-cubic grids and every
  tile width, so that the Nyquist tiles of the column passes are exercised;
* the bf16 loop replayed from a captured hipGraph gives bitwise the launch-
  mode solve, and switching a plan back to FP32 spectra gives bitwise the
  FP32 result again;
* the converged energy of mixed-precision solves with bf16 spectra is
  within 1e-6 eV/atom of the oracle's converged minimum
  (tests/golden/oracle_converged.json) or of the FP64 GPU solve.
"""
import json
import os

import numpy as np
import pytest
import torch

import dftpy_amd as D
from dftpy_amd import workloads as W
from dftpy_amd import batch as B
from dftpy_amd.plan import Plan

pytestmark = pytest.mark.gpu

AP_TOL = 2e-2
PAP_TOL = 1e-3


def _golden(name):
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_converged.json")) as fh:
        return json.load(fh)[name]


def _inputs(nr, cell, seed=0, prec="fp32"):
    """A plan with TF + vW + WT + LDA + Hartree (no PP), a smooth positive
    phi, and its residual and Hessian diagonal at mu = 0.3."""
    P = Plan(nr, cell, precision=prec)
    rng = np.random.default_rng(seed)
    x = [np.arange(n) / n for n in nr]
    X = np.meshgrid(*x, indexing="ij")
    rho = 0.02 * (1.0 + 0.3 * np.cos(2 * np.pi * X[0]) * np.sin(2 * np.pi * (X[1] + X[2]))
                  + 0.05 * rng.random(tuple(nr)))
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=False, rho0=float(rho.mean()))
    t = lambda a: torch.tensor(a, dtype=P.dtype, device=P.device)
    phi = t(np.sqrt(rho))
    V = P.empty()
    aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
    P.eval_phi(phi, None, 1.0, 0.0, V, aux)
    res, W0 = P.empty(), P.empty()
    P.hvp_prepare(phi, V, aux, 0.3, res, W0)
    return P, phi, res, W0


def _one_step(P, phi, res, W0):
    """One CG step of the device loop: (Ap of the step, p.Ap, direction)."""
    d = P.zeros()
    work = torch.zeros((4,) + tuple(P.nr), dtype=P.dtype, device=P.device)
    n, st, pAp = P.tn_direction(phi, res, W0, d, work, 1, 1)
    assert n == 1
    r0v = (-res).double()
    r1v = work[0].double()
    r0 = float((r0v * r0v).sum())
    alpha = r0 / pAp
    ap = (r0v - r1v) / alpha
    return ap, pAp, d.clone()


CASES = [((32, 32, 32), (0, 0)), ((32, 32, 32), (8, 16)), ((64, 64, 64), (32, 8)), ((96, 96, 96), (16, 16)),
         ((48, 64, 96), (16, 8)), ((64, 96, 128), (32, 32)), ((96, 48, 64), (8, 32)), ((128, 32, 48), (8, 8))]


@pytest.mark.parametrize("nr,tiles", CASES)
def test_bf16_hessian_product_within_bf16_rounding(nr, tiles):
    print("case bf16 Hessian product", nr, "tiles", tiles, flush=True)
    cell = np.diag([0.35 * n for n in nr])
    P, phi, res, W0 = _inputs(nr, cell)
    assert P.cg_spectra_tiles == (0, 0)
    ap32, pap32, d32 = _one_step(P, phi, res, W0)
    # the FP32 loop's Ap against the plan's standalone FP32 product
    ref = P.empty()
    P.hvp_fast(phi, -res, W0, ref)
    ref = ref.double()
    e32 = float((ap32 - ref).norm() / ref.norm())
    P.set_cg_spectra("bf16", tiles)
    widest = max(w for w in (32, 16, 8) if (nr[2] // 2) % w == 0)
    assert P.cg_spectra_tiles == tuple(t or widest for t in tiles)
    ap16, pap16, d16 = _one_step(P, phi, res, W0)
    e16 = float((ap16 - ref).norm() / ref.norm())
    epap = abs(pap16 - pap32) / abs(pap32)
    print("Ap rel. L2 error: FP32 loop %.2e, bf16 loop %.2e; p.Ap rel. error %.2e" % (e32, e16, epap))
    assert e32 < 1e-5
    assert e16 < AP_TOL
    assert e16 > 10 * e32  # the bf16 path did run
    assert epap < PAP_TOL
    # back to FP32 spectra: the FP32 loop again, bit for bit
    P.set_cg_spectra("plan")
    assert P.cg_spectra_tiles == (0, 0)
    ap32b, pap32b, d32b = _one_step(P, phi, res, W0)
    assert pap32b == pap32
    assert torch.equal(d32b, d32)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("nr", [(32, 32, 32), (48, 64, 96), (96, 48, 64), (192, 96, 128)])
def test_split_rows_give_the_padded_rows_result(prec, nr):
    """The default layout of the loop's spectra (split rows: n2/2 columns,
    the Nyquist column in a plane of its own) against the padded rows of the
    other passes: the same transforms of the same values, so the same
    product; only the grouping of p.Ap's FP64 partial sums differs."""
    print("case split vs padded rows", prec, nr, flush=True)
    cell = np.diag([0.35 * n for n in nr])
    P, phi, res, W0 = _inputs(nr, cell, prec=prec)
    assert P.cg_spectra()[0] == "plan"
    ap1, pap1, d1 = _one_step(P, phi, res, W0)
    P.set_cg_spectra("padded")
    assert P.cg_spectra() == ("padded", (0, 0))
    ap0, pap0, d0 = _one_step(P, phi, res, W0)
    epap = abs(pap1 - pap0) / abs(pap0)
    ed = float((d1.double() - d0.double()).norm() / d0.double().norm())
    print("p.Ap rel. difference %.2e, direction rel. difference %.2e" % (epap, ed))
    assert epap < 1e-12
    assert ed < (1e-6 if prec == "fp32" else 1e-14)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("nr", [(32, 32, 32), (48, 64, 96), (192, 96, 128)])
def test_split_rows_evaluation_is_bitwise_the_padded_one(prec, nr):
    """Evaluations (energies, potential, the line-search derivative) take
    the split rows too: the same arithmetic at other addresses, so the same
    bits as the padded rows."""
    print("case split vs padded rows, evaluation", prec, nr, flush=True)
    cell = np.diag([0.35 * n for n in nr])
    P, phi, res, W0 = _inputs(nr, cell, prec=prec)
    p = torch.sin(phi * 7.0)
    out = {}
    for mode in ("plan", "padded"):
        P.set_cg_spectra(mode)
        V = P.empty()
        aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
        sums = P.eval_phi(phi, p, 0.8, 0.6, V, aux)
        out[mode] = (np.array(sums), V.clone(), aux.clone())
    assert P.cg_spectra()[0] == "padded"
    assert np.array_equal(out["plan"][0], out["padded"][0])
    assert torch.equal(out["plan"][1], out["padded"][1])
    assert torch.equal(out["plan"][2], out["padded"][2])


def test_bf16_tile_selection_and_fallback():
    # n2 = 90: n2/2 = 45 has no tile width; FP64 plans keep their spectra
    cell = np.eye(3) * 20.0
    P = Plan((32, 32, 90), cell, precision="fp32")
    P.set_cg_spectra("bf16")
    assert P.cg_spectra() == ("padded", (0, 0))  # no split rows for an odd n2/2
    P = Plan((32, 32, 48), cell, precision="fp32")
    P.set_cg_spectra("bf16", (0, 0))
    assert P.cg_spectra_tiles == (8, 8)
    P.set_cg_spectra("bf16", (8, 16))  # 24 columns are no whole number of 16-wide tiles
    assert P.cg_spectra_tiles == (0, 0) and P.cg_spectra()[0] != "bf16"
    with pytest.raises(Exception):
        P.set_cg_spectra("bf16", (12, 8))
    P = Plan((32, 32, 32), cell, precision="fp64")
    P.set_cg_spectra("bf16")
    assert P.cg_spectra()[0] == "plan"  # FP64 plans keep their precision, in split rows
    with pytest.raises(ValueError):
        P.set_cg_spectra("fp16")


def _al108():
    wl = W.build("al108_96")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(3.0, 1.12)})
    ev = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    return wl, grid, ev


def test_bf16_graph_replays_give_the_same_solve():
    print("case bf16 spectra, TN graph mode al108_96", flush=True)
    wl, grid, ev = _al108()
    res = {}
    for graph in (False, True):
        opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                             optimization_options={"econv": 1e-9, "precision": "mixed", "tn_graph": graph,
                                                   "mixed_phase1": True, "cg_spectra": "bf16"})
        opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
        res[graph] = opt
        print(graph, opt.energy, [l["Nd"] for l in opt.log])
        assert opt.converged == 0
    assert res[True].energy == res[False].energy
    assert [l["Nd"] for l in res[True].log] == [l["Nd"] for l in res[False].log]
    # the same evaluator with FP32 spectra afterwards: the FP32 trajectory
    o32 = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-9, "precision": "mixed", "mixed_phase1": True})
    o32.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    o64 = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-9, "precision": "fp64"})
    o64.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    nat = len(wl["symbols"])
    err = abs(res[True].energy - o64.energy) * W.HARTREE / nat
    print("bf16 mixed %.12f, FP32 mixed %.12f, FP64 %.12f: bf16 error %.2e eV/atom"
          % (res[True].energy, o32.energy, o64.energy, err))
    assert o32.converged == 0 and o64.converged == 0
    assert err < 1e-6
    assert abs(o32.energy - o64.energy) * W.HARTREE / nat < 1e-6


@pytest.mark.parametrize("name", ["batch00", "batch05"])
def test_bf16_mixed_solve_against_oracle_minimum(name):
    cell = next(c for c in W.batch_cells(64) if c["name"] == name)
    print("case bf16 spectra mixed solve", name, cell["nr"], flush=True)
    r = B.solve_cell(cell, "mixed", econv=1e-8, options={"cg_spectra": "bf16", "mixed_phase1": True})
    g = _golden(name)["energy_Ha"]
    err = abs(r["energy"] - g) * W.HARTREE / r["nat"]
    print("%s: E %.10f Ha, oracle minimum %.10f Ha, error %.2e eV/atom, %d iterations"
          % (name, r["energy"], g, err, r["iterations"]))
    assert r["converged"] == 0
    assert err < 1e-6


def test_bf16_al864_192_against_oracle_minimum():
    print("case bf16 spectra al864_192 mixed solve", flush=True)
    wl = W.build("al864_192")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={s: D.empty_core(*W.EMPTY_CORE[s]) for s in wl["tables"]},
                       PME=True)
    ev = D.TotalFunctional(KEDF=D.KEDF(wl["kedf"], **wl["kedf_options"]), XC=D.XC("LDA"), HARTREE=D.Hartree(),
                           PSEUDO=pp)
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": 1e-8, "precision": "mixed", "mixed_phase1": True,
                                               "cg_spectra": "bf16"})
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    g = _golden("al864_192")["energy_Ha"]
    err = abs(opt.energy - g) * W.HARTREE / len(wl["symbols"])
    print("al864_192 bf16: E %.10f Ha, oracle minimum %.10f Ha, error %.2e eV/atom, Nd %s"
          % (opt.energy, g, err, [l["Nd"] for l in opt.log]))
    assert opt.converged == 0
    assert err < 1e-6
