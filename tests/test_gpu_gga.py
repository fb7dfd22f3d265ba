"""GGA kinetic functionals on the GPU (GGA, kedf/gga.py:1212-1326) against
the oracle: energy (1e-10 relative) and potential (1e-9 of max |V|) of
every enhancement factor of the reference's list at 32^3, on the half,
full and sheared grids for LKT, FP32 against FP64, the reference's LC94
known answer on Al_fde_rho.pp, the stress, and a TN solve with the
forward-difference product against the oracle's TN."""
import os

import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import formats as F
from dftpy_amd import workloads as W
from tests.conftest import DATA

pytestmark = pytest.mark.gpu

NAMES = sorted(O.GGA_KEDFS)


def _smooth(nr, cell, ne):
    n = nr[0]
    i = np.arange(n) / n
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    rng = np.random.default_rng(4)
    f = np.ones((n, n, n))
    for _ in range(6):
        k = rng.integers(-3, 4, size=3)
        f += 0.05 * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
    vol = abs(np.linalg.det(cell))
    return f * ne / vol


def _pair(name, full=False, shear=None, prec=torch.float64):
    wl = W.build("al4_32")
    cell = wl["cell"] if shear is None else wl["cell"] @ (np.eye(3) + shear)
    grid = D.DirectGrid(cell, nr=wl["nr"], full=full)
    og = O.Grid(cell, wl["nr"], full=full)
    rho = _smooth(wl["nr"], cell, 12.0)
    tot = D.TotalFunctional(KEDF=D.KEDF("GGA", k_str=name))
    out = tot.get_energy_potential(D.DirectField(grid=grid, data=rho, dtype=prec), split=True)
    V = tot.compute(D.DirectField(grid=grid, data=rho, dtype=prec)).potential
    E, Vo = O.gga_functional(rho, og, name)
    return out["KEDF-GGA"].energy, V.detach().double().cpu().numpy(), E, Vo


@pytest.mark.parametrize("name", NAMES)
def test_gga_energy_and_potential_match_oracle(name):
    print("case gga", name, flush=True)
    e, v, eo, vo = _pair(name)
    err = np.abs(v - vo).max() / np.abs(vo).max()
    print("%-8s E GPU %.14e oracle %.14e  V rel %.2e" % (name, e, eo, err))
    assert abs(e - eo) <= 1e-10 * abs(eo)
    assert err <= 1e-9


@pytest.mark.parametrize("full,shear", [(True, None), (False, np.array([[0, .03, -.02], [.01, .02, .04], [-.03, 0, -.01]]))])
def test_gga_full_and_sheared_grids(full, shear):
    print("case gga LKT full=%s sheared=%s" % (full, shear is not None), flush=True)
    e, v, eo, vo = _pair("LKT", full=full, shear=shear)
    assert abs(e - eo) <= 1e-10 * abs(eo)
    assert np.abs(v - vo).max() <= 1e-9 * np.abs(vo).max()


def test_gga_fp32_against_fp64():
    print("case gga LKT fp32", flush=True)
    e64, v64, _, _ = _pair("LKT")
    e32, v32, _, _ = _pair("LKT", prec=torch.float32)
    assert abs(e32 - e64) <= 1e-6 * abs(e64)
    assert np.abs(v32 - v64).max() <= 1e-5 * np.abs(v64).max()


def test_gga_lc94_known_answer():
    """examples/test/test_functional.py:10-18 on the GPU."""
    print("case gga LC94 known answer", flush=True)
    rho = F.read_density(os.path.join(DATA, "Al_fde_rho.pp"))
    out = D.KEDF("GGA", k_str="lc94").compute(rho)
    print("LC94 energy %.16f" % out.energy)
    assert abs(out.energy - 1.6821337114254904) <= 1e-10 * 1.6821337114254904


def test_gga_total_stress_and_tn_solve():
    """KEDF GGA (LKT) + LDA + Hartree + PME local PP on the 4-atom cell:
    get_stress against the oracle; the reference's forward-difference
    Hessian product (optimization.py:106-114, eps = 1e-9) on the GPU against
    the oracle's at the same phi and p; and a TN solve with that product
    against the oracle's TN minimum.

    The solve's bar follows from how the forward difference behaves on this
    functional: dividing by eps = 1e-9 turns a rounding-level change of the
    potential into a 1e-6-level change of the product, and the TN loop and
    line search amplify it.  Run on the oracle with the potential multiplied
    by (1 + 1e-15 N(0,1)) (three seeds), the reference algorithm itself ends
    with 'Line-Search Failed' (converged = 1) 2-5e-8 Ha above the clean
    minimum; with 1e-14 it fails 1.1e-5 Ha above it.  So the GPU solve must
    end converged (0) or at that line-search failure (1), within 2e-5 Ha
    of the oracle's minimum; the product itself is the deterministic check."""
    print("case gga LKT total", flush=True)
    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(*W.EMPTY_CORE["Al"])})
    tot = D.TotalFunctional(KEDF=D.KEDF("GGA", k_str="LKT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {"Al": O.empty_core_table(*W.EMPTY_CORE["Al"])}, PME=True)
    otot = O.TotalFunctional(og, kedf="GGA", kedf_options={"functional": "LKT"}, pseudo=opp,
                             ewald=O.Ewald(og, oi, PME=True))
    rho = _smooth(wl["nr"], wl["cell"], 12.0)
    got = tot.get_stress(D.DirectField(grid=grid, data=rho), split=True)
    ref = otot.get_stress(rho, split=True)
    for k in ("KEDF-GGA", "TOTAL"):
        assert np.abs(got[k] - ref[k]).max() <= 1e-9 * np.abs(ref[k]).max(), k

    # one forward-difference product Ap = ((V(phi + eps p) - mu) phi1 - res0) / eps
    eps = 1e-9
    phi = np.sqrt(wl["rho_ini"])
    rng = np.random.default_rng(7)
    p = _smooth(wl["nr"], wl["cell"], 1.0) * (1.0 + 0.1 * rng.standard_normal(phi.shape))
    p -= p.mean()

    def gpu_v(r):
        out = tot.compute(D.DirectField(grid=grid, data=r))
        return out.potential.detach().double().cpu().numpy()

    def ora_v(r, ph):
        return otot.compute(r, ("E", "V"), phi=ph)[1]

    phi1 = phi + eps * p
    prods = []
    for vf in (lambda r, ph: gpu_v(r), ora_v):
        V0 = vf(phi * phi, phi)
        mu = np.sum(V0 * phi * phi) / np.sum(phi * phi)
        res0 = (V0 - mu) * phi
        prods.append(((vf(phi1 * phi1, phi1) - mu) * phi1 - res0) / eps)
    ap, apo = prods
    rel = np.abs(ap - apo).max() / np.abs(apo).max()
    pap, papo = np.sum(p * ap), np.sum(p * apo)
    print("forward-difference product: max rel %.2e, p.Ap %.12e oracle %.12e" % (rel, pap, papo))
    assert rel < 1e-3
    assert abs(pap - papo) < 1e-5 * abs(papo)

    opts = {"econv": 1e-9, "maxiter": 80}
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options=dict(opts))
    assert opt.optimization_options["hvp"] == "fd"
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    oopt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "TN", dict(opts))
    oopt.optimize_rho(wl["rho_ini"])
    print("GPU %.12f (converged %d)  oracle %.12f (converged %d)" % (opt.energy, opt.converged, oopt.energy,
                                                                     oopt.converged))
    assert oopt.converged == 0
    # the first outer step takes the same number of CG iterations
    assert opt.log[1]["Nd"] == oopt.log[1][4]
    assert opt.converged in (0, 1)
    assert abs(opt.energy - oopt.energy) < 2e-5
