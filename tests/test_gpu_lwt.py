"""LWT / LMGP nonlocal kinetic functionals on the GPU (LWTPotentialEnergy,
kedf/lwt.py:83-341) against the oracle: energy (1e-10 relative) and
potential (1e-9 of max |V|) of KEDF 'LWT' / 'LMGP' / 'LMGPA' / 'LMGPG' at
32^3 with linear and Hermite interpolation, kdd 1 / 2 / 3, plain tables,
the MGP kinetic-electron term, given kf bounds (points below the first node
and the kf clamp), low and negative densities, full and sheared grids;
FP32 against FP64; the reference's LMGP known answer on Al_fde_rho.pp; and
a TN solve with the forward-difference product against the oracle's TN."""
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

SHEAR = np.array([[0, .03, -.02], [.01, .02, .04], [-.03, 0, -.01]])


def _smooth(n, cell, ne, seed=4, amp=0.05):
    i = np.arange(n) / n
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    rng = np.random.default_rng(seed)
    f = np.ones((n, n, n))
    for _ in range(6):
        k = rng.integers(-3, 4, size=3)
        f += amp * np.cos(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
    return f * ne / abs(np.linalg.det(cell))


def _pair(name, opts, full=False, shear=None, prec=torch.float64, rho=None):
    wl = W.build("al4_32")
    cell = wl["cell"] if shear is None else wl["cell"] @ (np.eye(3) + shear)
    grid = D.DirectGrid(cell, nr=wl["nr"], full=full)
    og = O.Grid(cell, wl["nr"], full=full)
    if rho is None:
        rho = _smooth(32, cell, 12.0)
    tot = D.TotalFunctional(KEDF=D.KEDF(name, **opts))
    out = tot.get_energy_potential(D.DirectField(grid=grid, data=rho, dtype=prec), split=True)
    otot = O.TotalFunctional(og, kedf=name, kedf_options=dict(opts), xc=False, hartree=False)
    E, Vo, parts = otot.compute(rho, split=True)
    return (out["KEDF-NL"].energy, out["TOTAL"].energy, out["TOTAL"].potential.detach().double().cpu().numpy(),
            parts["KEDF-NL"], E, Vo)


def _check(res, etol=1e-10, vtol=1e-9):
    enl, e, v, enlo, eo, vo = res
    err = np.abs(v - vo).max() / np.abs(vo).max()
    print("E_NL GPU %.14e oracle %.14e  E %.14e / %.14e  V rel %.2e" % (enl, enlo, e, eo, err), flush=True)
    assert abs(enl - enlo) <= etol * abs(enlo)
    assert abs(e - eo) <= etol * abs(eo)
    assert err <= vtol


CASES = {
    "lwt_linear": ("LWT", dict()),
    "lwt_hermite": ("LWT", dict(interp="hermite")),
    "lwt_kdd1": ("LWT", dict(kdd=1)),
    "lwt_kdd2": ("LWT", dict(kdd=2)),
    "lwt_hermite_kdd2": ("LWT", dict(interp="hermite", kdd=2)),
    "lwt_order1": ("LWT", dict(order=1)),
    "lwt_ratio_nsp": ("LWT", dict(nsp=9)),
    "lwt_delta_x": ("LWT", dict(delta=0.2, x=0.8)),
    "lwt_lump": ("LWT", dict(lumpfactor=0.3)),
    "lwt_lump3": ("LWT", dict(lumpfactor=[0.1, 0.05, 2.0], interp="hermite")),
    "lmgp": ("LMGP", dict(maxpoints=100)),
    "lmgpa": ("LMGPA", dict(maxpoints=100)),
    "lmgpg": ("LMGPG", dict(maxpoints=100, interp="hermite")),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_lwt_energy_and_potential_match_oracle(case):
    name, opts = CASES[case]
    print("case lwt", case, flush=True)
    _check(_pair(name, opts))


@pytest.mark.parametrize("interp", ["linear", "hermite"])
def test_lwt_given_bounds_low_and_negative_density(interp):
    """kfmin above the smallest kf (points below the first node, with
    pot1 = p_0 (kf - kf_0) / kf_0), kfmax below the largest (the clamp),
    and a region of tiny and negative densities (the low-density weight
    and its zero for rho < 0)."""
    print("case lwt bounds / low density", interp, flush=True)
    wl = W.build("al4_32")
    rho = _smooth(32, wl["cell"], 12.0, seed=7, amp=0.12)
    rho[:3, :4, :5] = 3e-7
    rho[0, 0, :3] = -2e-9
    kf = 2.0 * np.cbrt(3.0 * np.pi ** 2 * np.abs(rho[rho > 1e-3]))
    opts = dict(interp=interp, kfmin=1.05 * kf.min(), kfmax=0.95 * kf.max())
    _check(_pair("LWT", opts, rho=rho))


@pytest.mark.parametrize("full,shear", [(True, None), (False, SHEAR), (True, SHEAR)])
def test_lwt_full_and_sheared_grids(full, shear):
    print("case lwt full=%s sheared=%s" % (full, shear is not None), flush=True)
    _check(_pair("LWT", dict(interp="hermite", lumpfactor=0.2), full=full, shear=shear))


def test_lwt_fp32_against_fp64():
    print("case lwt fp32", flush=True)
    r64 = _pair("LWT", dict())
    r32 = _pair("LWT", dict(), prec=torch.float32)
    err = np.abs(r32[2] - r64[2]).max() / np.abs(r64[2]).max()
    print("FP32 E_NL rel %.2e  V rel %.2e" % (abs(r32[0] - r64[0]) / abs(r64[0]), err))
    assert abs(r32[0] - r64[0]) <= 1e-5 * abs(r64[0])
    # two FP32 transforms per kf node, each node's p and m summed into pot1,
    # pot3 and the pot2 spectrum in FP32: a few 1e-5 of max |V|
    assert err <= 1e-4


def test_lwt_energy_only_equals_energy_of_full_call():
    """calcType E runs the reference's kdd = 1 path; the energy is the same."""
    print("case lwt energy only", flush=True)
    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    rho = D.DirectField(grid=grid, data=_smooth(32, wl["cell"], 12.0))
    k = D.KEDF("LWT", interp="hermite")
    e1 = k.compute(rho, calcType=("E",)).energy
    e2 = k.compute(rho).energy
    assert abs(e1 - e2) <= 1e-12 * abs(e2)


def test_lmgp_known_answer():
    """examples/test/test_functional.py:28-33 on the GPU (the reference
    asserts 1e-3; the oracle matches all digits)."""
    print("case lmgp known answer", flush=True)
    rho = F.read_density(os.path.join(DATA, "Al_fde_rho.pp"))
    out = D.Functional(type="KEDF", name="LMGP").compute(rho)
    print("LMGP energy %.16f" % out.energy)
    assert abs(out.energy - 2.9145872498710492) <= 1e-10 * 2.9145872498710492


def test_lwt_stress_is_not_implemented():
    """The reference has no LWT stress (name2functions raises)."""
    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    tot = D.TotalFunctional(KEDF=D.KEDF("LWT"))
    with pytest.raises(AttributeError):
        tot.get_stress(D.DirectField(grid=grid, data=_smooth(32, wl["cell"], 12.0)))


KF_FIXED = dict(kfmin=0.3, kfmax=5.0)


def _lwt_pair(name, opts):
    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(*W.EMPTY_CORE["Al"])})
    tot = D.TotalFunctional(KEDF=D.KEDF(name, **opts), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {"Al": O.empty_core_table(*W.EMPTY_CORE["Al"])}, PME=True)
    otot = O.TotalFunctional(og, kedf=name, kedf_options=dict(opts), pseudo=opp, ewald=O.Ewald(og, oi, PME=True))
    o = {"econv": 1e-9, "maxiter": 80}
    opt = D.Optimization(EnergyEvaluator=tot, optimization_method="TN", optimization_options=dict(o))
    assert opt.optimization_options["hvp"] == "fd"
    opt.optimize_rho(guess_rho=D.DirectField(grid=grid, data=wl["rho_ini"]))
    oopt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "TN", dict(o))
    oopt.optimize_rho(wl["rho_ini"])
    print("GPU %.12f  oracle %.12f  (%d / %d iterations, status %d / %d)"
          % (opt.energy, oopt.energy, len(opt.log), len(oopt.log), opt.converged, oopt.converged))
    return opt, oopt


@pytest.mark.parametrize("name,opts", [("LWT", dict(KF_FIXED)), ("LWT", dict(KF_FIXED, interp="hermite")),
                                       ("LMGP", dict(KF_FIXED, maxpoints=100, interp="hermite"))])
def test_lwt_tn_solve(name, opts):
    """KEDF LWT / LMGP + LDA + Hartree + PME local PP on the 4-atom cell: a
    TN solve (forward-difference product, as the reference) against the
    oracle's TN, with the kf node range given (kfmin, kfmax: lwt.py:36-40),
    so that the functional is the same function at every call.  With
    linear interpolation the potential is not the exact gradient of the
    energy and both line searches fail (converged = 1) at the same point;
    with Hermite interpolation both converge.  These outcomes are stable
    under rounding: the oracle gives the same status with its potential
    multiplied by (1 + 1e-15 N(0,1)) (linear: 6e-9 Ha apart) or 1e-14 N(0,1)
    (Hermite: 4e-13 Ha, LMGP 8e-12 Ha apart)."""
    print("case lwt TN", name, opts, flush=True)
    opt, oopt = _lwt_pair(name, opts)
    assert opt.converged == oopt.converged
    if opts.get("interp") == "hermite":
        assert opt.converged == 0
        # both stop on two energy changes below econv; with the
        # forward-difference product the stopping points differ by a few econv
        assert abs(opt.energy - oopt.energy) < 3e-8
    else:
        assert opt.converged == 1
        assert abs(opt.energy - oopt.energy) < 1e-7


def test_lwt_tn_solve_growing_kf_range():
    """The same solve with the kf range that grows with the densities seen
    (guess_kf_bound, lwt.py:36-80).  The range widens by whole 0.5 steps of
    kfmax (and factors of kfmin) as soon as one evaluated density reaches
    past it, including the forward-difference points phi + 1e-9 p, so which
    evaluation widens it turns on last-digit rounding of max(rho), and from
    there the trajectories of any two implementations part (here after the
    second step).  The bar: the first two outer steps equal the oracle's
    (1e-8 Ha, same CG and line-search counts), and the solve ends converged
    or at a failed line search, the reference's two outcomes."""
    print("case lwt TN growing kf range", flush=True)
    opt, oopt = _lwt_pair("LWT", {})
    for k in (1, 2):
        g, o = opt.log[k], oopt.log[k]
        print(k, g["energy"], o[1], g["Nd"], o[4], g["Nls"], o[5])
        assert abs(g["energy"] - o[1]) < 1e-8
        assert (g["Nd"], g["Nls"]) == (o[4], o[5])
    assert opt.converged in (0, 1) and oopt.converged in (0, 1)
