"""The INI job driver end to end on the GPU (``python -m dftpy_amd
config.ini``): fcc Al, 4 atoms, POSCAR + recpot files, 32^3, TFvW + LDA +
Hartree + PME local PP, TN to econv 1e-9 Ha/atom, energy / forces / stress
and the density written as XSF; checked against the oracle at the GPU's
final density (energy 1e-10 relative, forces and stress 1e-9 of their
largest component) and against the oracle's own TN minimum (1e-7 Ha)."""
import os

import numpy as np
import pytest
import torch

import oracle as O
from dftpy_amd import formats as F
from dftpy_amd import interface as I
from dftpy_amd import workloads as W
from dftpy_amd.ions import Ions
from tests.conftest import DATA

pytestmark = pytest.mark.gpu


def test_ini_job_matches_oracle(tmp_path):
    print("case INI job al4_32", flush=True)
    wl = W.build("al4_32")
    ions = Ions(symbols=wl["symbols"], positions=wl["positions"], cell=wl["cell"])
    F.write_vasp(str(tmp_path / "POSCAR"), ions=ions)
    ini = tmp_path / "config.ini"
    ini.write_text("""
[JOB]
task = Optdensity
calctype = Energy Force Stress
[PATH]
pppath = {data}
cellpath = {tmp}
[PP]
Al = al.lda.recpot
[CELL]
cellfile = POSCAR
[GRID]
nr = 32 32 32
[KEDF]
kedf = TFVW
y = 1/9
[OPT]
method = TN
econv = 1e-9
[DENSITY]
densityoutput = {tmp}/rho.xsf
""".format(data=DATA, tmp=tmp_path))
    (res,) = I.main([str(ini)])
    rho = res["density"]
    r = rho.detach().double().cpu().numpy()
    E = res["energypotential"]["TOTAL"].energy
    # the oracle on the same cell, tables and density
    og = O.Grid(ions.cell, (32, 32, 32), full=False)
    oi = O.Ions(wl["symbols"], ions.positions, ions.cell)
    opp = O.LocalPseudo(og, oi, {"Al": O.read_recpot(os.path.join(DATA, "al.lda.recpot"))}, PME=True)
    oew = O.Ewald(og, oi, PME=True)
    otot = O.TotalFunctional(og, kedf="TFVW", kedf_options=dict(x=1.0, y=1.0 / 9.0), pseudo=opp, ewald=oew)
    Eo = otot.compute(r, calc=("E",))[0]
    print("GPU %.12f  oracle at GPU density %.12f" % (E, Eo))
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    fo = opp.forces(r) + oew.forces
    assert np.abs(res["forces"]["TOTAL"] - fo).max() <= 1e-9 * np.abs(fo).max() + 1e-12
    so = otot.get_stress(r)
    assert np.abs(res["stress"]["TOTAL"] - so).max() <= 1e-9 * np.abs(so).max()
    # the oracle's own TN from the same uniform start reaches the same minimum
    rho0 = np.full((32, 32, 32), oi.get_ncharges() / og.volume)
    oopt = O.Optimization(lambda x, c, p: otot.compute(x, c, phi=p), og, "TN", {"econv": 4e-9})
    oopt.optimize_rho(rho0)
    print("GPU minimum %.12f  oracle minimum %.12f" % (E, oopt.energy))
    assert abs(E - oopt.energy) < 1e-7
    # the density file
    _, back, _ = F.read_all(str(tmp_path / "rho.xsf"))
    assert np.abs(back.detach().cpu().numpy() - r).max() <= 1e-12 * np.abs(r).max()


def test_ini_lwt_job_matches_api(tmp_path):
    """[KEDF] kedf = LMGPA with the schema's LWT options (INI default interp
    hermite) runs the same solve as the Python API with those options."""
    print("case INI LMGPA job al4_32", flush=True)
    wl = W.build("al4_32")
    ions = Ions(symbols=wl["symbols"], positions=wl["positions"], cell=wl["cell"])
    F.write_vasp(str(tmp_path / "POSCAR"), ions=ions)
    ini = tmp_path / "config.ini"
    ini.write_text("""
[JOB]
task = Optdensity
calctype = Energy
[PATH]
pppath = {data}
cellpath = {tmp}
[PP]
Al = al.lda.recpot
[CELL]
cellfile = POSCAR
[GRID]
nr = 32 32 32
[KEDF]
kedf = LMGPA
maxpoints = 100
[OPT]
method = TN
econv = 1e-8
""".format(data=DATA, tmp=tmp_path))
    (res,) = I.main([str(ini)])
    E = res["energypotential"]["TOTAL"].energy
    import dftpy_amd as D
    # the ions as the job read them from its POSCAR (the file's digits)
    ions2 = F.read(str(tmp_path / "POSCAR"), format="vasp")
    grid = D.DirectGrid(ions2.cell, nr=(32, 32, 32), full=False)
    pp = D.LocalPseudo(grid=grid, ions=ions2, PP_list={"Al": os.path.join(DATA, "al.lda.recpot")})
    ev = D.TotalFunctional(KEDF=D.KEDF("LMGPA", maxpoints=100, interp="hermite"), XC=D.XC("LDA"),
                           HARTREE=D.Hartree(), PSEUDO=pp)
    # the job's starting density, built the way ConfigParser builds it: the
    # reference's forward-difference HVP (used with LWT) turns last-digit
    # differences of the start into 1e-7 differences of the first products
    rho = D.DirectField(grid=grid)
    with torch._C.DisableTorchFunctionSubclass():
        rho.fill_(12.0 / grid.volume)
        rho.mul_(12.0 / (float(rho.sum()) * grid.dV))
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN", optimization_options={"econv": 4e-8})
    opt.optimize_rho(guess_rho=rho)
    # the job reports the energy of a fresh evaluation at the final density
    # (interface.py:245-317).  LMGPA's kf node range only grows with the
    # densities evaluated (lwt.py:36-80), so after a failed line search
    # (whose trial rotations reach far) that evaluation can differ from the
    # optimizer's last energy; the API side takes the same evaluation.
    E_api = ev.compute(opt.rho).energy
    print("INI %.12f  API %.12f (optimizer %.12f, converged %d)" % (E, E_api, opt.energy, opt.converged))
    assert abs(E - E_api) < 1e-9
    assert opt.converged in (0, 1)
