"""GPU parity of the drop-in API (LocalPseudo, Ewald, TotalFunctional,
DirectField.fft/ifft) with the oracle and with the reference's known
answers (test_read_pp.py:12, test_pme.py:32-36, test_ewald.py:33-37,
test_functional.py:27)."""
import os

import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import workloads as W
from tests.conftest import DATA

pytestmark = pytest.mark.gpu


def _fixture(full=True):
    oi, rho, og = O.read_qepp(os.path.join(DATA, "Al_fde_rho.pp"), full=full)
    grid = D.DirectGrid(og.lattice, nr=og.nr, full=full)
    ions = D.Ions(oi.symbols, oi.positions, og.lattice)
    ions.set_charges({"Al": 3.0})
    return oi, rho, og, grid, ions


@pytest.mark.parametrize("pme", [True, False])
def test_local_pseudo_fixture_known_answer(pme):
    print("case pp fixture pme=%s" % pme, flush=True)
    oi, rho, og, grid, ions = _fixture()
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": os.path.join(DATA, "al.lda.recpot")}, PME=pme)
    r = D.DirectField(grid=grid, data=rho)
    e = pp(r).energy
    opp = O.LocalPseudo(og, oi, {"Al": O.read_recpot(os.path.join(DATA, "al.lda.recpot"))}, PME=pme)
    eo = opp.compute(rho)[0]
    print("PP energy", e, eo)
    assert abs(e - 1.3497046) < 1e-4          # test_read_pp.py:12
    assert abs(e - eo) < 1e-10 * abs(eo)
    dv = np.abs(pp.vreal.asnumpy() - opp.vreal).max()
    assert dv < 1e-9, dv


def test_local_forces_pme_vs_exact_and_oracle():
    oi, rho, og, grid, ions = _fixture()
    r = D.DirectField(grid=grid, data=rho)
    F = {}
    for pme in (True, False):
        pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": os.path.join(DATA, "Al_lda.oe01.recpot")}, PME=pme)
        F[pme] = pp.forces(r)
        opp = O.LocalPseudo(og, oi, {"Al": O.read_recpot(os.path.join(DATA, "Al_lda.oe01.recpot"))}, PME=pme)
        Fo = opp.forces(rho)
        print("forces", pme, F[pme], Fo)
        assert np.abs(F[pme] - Fo).max() < 1e-10
    assert np.allclose(F[True], F[False], atol=1e-4)  # test_pme.py:34


def test_local_forces_match_energy_derivative():
    """Forces on a displaced 4-atom cell equal -dE_PP/dR (central difference).
    An all-odd 27^3 grid has no Nyquist planes, on which the reference's
    half-space force sum and the irfftn energy treat G differently."""
    wl = W.build("al4_32")
    pos = wl["positions"].copy()
    pos[1] += np.array([0.05, -0.03, 0.02])
    grid = D.DirectGrid(wl["cell"], nr=(27, 27, 27), full=False)
    wl["rho_ini"] = W.initial_density((27, 27, 27), 12.0, abs(np.linalg.det(wl["cell"])))
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    def energy(p, pme):
        ions = D.Ions(wl["symbols"], p, wl["cell"])
        pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={"Al": D.empty_core(3.0, 1.12)}, PME=pme)
        return pp(rho).energy, pp
    for pme in (False, True):
        _, pp = energy(pos, pme)
        F = pp.forces(rho)
        h = 1e-4
        for d in range(3):
            pp_ = pos.copy(); pp_[1, d] += h
            pm_ = pos.copy(); pm_[1, d] -= h
            fd = -(energy(pp_, pme)[0] - energy(pm_, pme)[0]) / (2 * h)
            print(pme, d, F[1, d], fd)
            assert abs(F[1, d] - fd) < 1e-6


def test_ewald_fixture_known_answers():
    oi, rho, og, grid, ions = _fixture()
    for pme in (False, True):
        ew = D.Ewald(ions=ions, grid=grid, PME=pme)
        er, ek = ew._sums()
        print("ewald", pme, ew.Energy_corr(), er, ek)
        assert abs(ew.Energy_corr() - (-13.003839613564997)) < 1e-5   # test_ewald.py:33
        assert abs(er - 0.0) < 1e-5
        assert abs(ek - 7.606943756474895) < 1e-5
        oe = O.Ewald(og, oi, PME=pme)
        assert abs(ew.energy - oe.energy) < 1e-10


def test_total_functional_api_wt_fixture():
    """KEDF('WT') through the drop-in API on the fixture (test_functional.py:27)."""
    oi, rho, og, grid, ions = _fixture()
    ke = D.Functional(type="KEDF", name="WT")
    out = ke(D.DirectField(grid=grid, data=rho))
    assert abs(out.energy - 2.916818700014412) < 1e-10
    assert out.potential.shape == tuple(og.nr)


def test_field_fft_roundtrip_and_conventions():
    """fft()[0,0,0] == integral; ifft(fft(f)) == f (test_field.py:124-131);
    Gaussian transform pair on a 51^3-class grid with 2,3,5 sizes."""
    for full in (True, False):
        grid = D.DirectGrid(np.eye(3) * 10.0, nr=(50, 48, 45), full=full)
        rng = np.random.default_rng(0)
        f = D.DirectField(grid=grid, data=rng.random(tuple(grid.nrR)))
        F = f.fft()
        assert abs(complex(F[0, 0, 0].cpu()) - f.integral()) < 1e-10
        g = F.ifft()
        assert float((g - f).abs().max()) < 1e-12
        ref = np.fft.fftn(f.asnumpy()) * grid.dV if full else np.fft.rfftn(f.asnumpy()) * grid.dV
        assert np.abs(F.asnumpy() - ref).max() < 1e-12 * np.abs(ref).max()


def _displaced_cell():
    cell = np.eye(3) * 7.65339
    frac = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
    pos = frac @ cell + np.random.default_rng(2).normal(0.0, 0.15, size=(4, 3))
    return ["Al", "Mg", "Al", "Al"], pos, cell, np.array([3.0, 2.0, 3.0, 3.0])


@pytest.mark.parametrize("pme", [False, True])
def test_ewald_forces_match_oracle(pme):
    """Ion-ion forces (ewald.py:509-519) against the oracle on the fixture
    (zero by symmetry, test_ewald.py:37-40) and on a displaced Al/Mg cell."""
    print("case ewald forces pme=%s" % pme, flush=True)
    oi, rho, og, grid, ions = _fixture()
    F = D.Ewald(ions=ions, grid=grid, PME=pme).forces
    assert np.abs(F).max() < 1e-5
    syms, pos, cell, q = _displaced_cell()
    for nr in ((24, 24, 24), (27, 30, 32)):
        grid = D.DirectGrid(cell, nr=nr, full=False)
        ions = D.Ions(syms, pos, cell, charges=q)
        F = D.Ewald(ions=ions, grid=grid, PME=pme).forces
        og = O.Grid(cell, nr, full=False)
        Fo = O.Ewald(og, O.Ions(syms, pos, cell, charges=q), PME=pme).forces
        print(nr, np.abs(F - Fo).max(), np.abs(Fo).max())
        assert np.abs(F - Fo).max() < 1e-10 * max(1.0, np.abs(Fo).max())


def test_total_forces_split_and_ewald_exact_vs_pme():
    """TotalFunctional.get_forces (total_functional.py:148-174): TOTAL =
    PSEUDO + II; the ion-ion forces of the 256-atom Al/Mg cell at 128^3 agree
    between the exact and PME reciprocal sums."""
    wl = W.build("almg256_128")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={s: D.empty_core(*W.EMPTY_CORE[s]) for s in wl["tables"]})
    tot = D.TotalFunctional(KEDF=D.KEDF("WT"), XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    rho = D.DirectField(grid=grid, data=wl["rho_ini"])
    f = tot.get_forces(rho, split=True)
    assert set(f) == {"TOTAL", "PSEUDO", "II"}
    assert np.abs(f["TOTAL"] - f["PSEUDO"] - f["II"]).max() < 1e-12
    assert np.abs(tot.get_forces(rho) - f["TOTAL"]).max() == 0.0
    Fx = D.Ewald(ions=ions, grid=grid, PME=False).forces
    print("Ewald exact vs PME at 128^3: %.3e (max |F| %.3e)" % (np.abs(Fx - f["II"]).max(), np.abs(Fx).max()))
    assert np.abs(Fx - f["II"]).max() < 1e-4
    # momentum conservation: exact for the exact sums, to the PME accuracy for PME
    assert np.abs(Fx.sum(axis=0)).max() < 1e-9 * np.abs(Fx).max() * len(Fx)
    assert np.abs(f["II"].sum(axis=0)).max() < 1e-5
