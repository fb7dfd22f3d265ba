"""ASE calculator over the GPU path (api/api4ase.py:11-119).

``DFTpyCalculator(config=...)`` runs the INI job (ConfigParser and
OptimizeDensityConf) for the calculator's atoms; ``DFTpyCalculator(
evaluator=..., rho=...)`` runs the density optimization with a given
TotalFunctional from the last density. Energies are returned in eV, forces
in eV/Angstrom and stress in eV/Angstrom^3 as Voigt (xx, yy, zz, yz, xz,
xy), as ase expects.

ase is optional: when it is importable the calculator derives from its
``Calculator``; otherwise a small base class with the same calling
convention (``get_potential_energy(atoms)``, ``get_forces``, ``get_stress``,
``calculate(atoms, properties, system_changes)``, ``results``) stands in,
and any object with ``get_chemical_symbols()``, ``get_positions()``,
``get_cell()`` and ``copy()`` serves as the atoms."""
import numpy as np
import torch

from .field import DirectField
from .grid import DirectGrid
from .interface import ConfigParser, OptimizeDensityConf, evaluator2results, HARTREE_EV
from .ions import Ions, BOHR
from .optimization import Optimization

try:  # pragma: no cover - ase is not a dependency
    from ase.calculators.calculator import Calculator, all_changes
except ImportError:
    all_changes = ["positions", "numbers", "cell", "pbc", "initial_charges", "initial_magmoms"]

    class Calculator:
        """The part of ase's Calculator that DFTpyCalculator relies on:
        the atoms of the last calculation, ``results`` and the change
        check that decides whether to run again."""

        implemented_properties = []

        def __init__(self, atoms=None, **kwargs):
            self.atoms = None
            self.results = {}
            self.parameters = dict(kwargs)
            if atoms is not None:
                atoms.calc = self
                self.atoms = atoms.copy()

        @staticmethod
        def _state(atoms):
            num = atoms.get_atomic_numbers() if hasattr(atoms, "get_atomic_numbers") else None
            syms = list(atoms.get_chemical_symbols())
            pbc = np.asarray(atoms.get_pbc()) if hasattr(atoms, "get_pbc") else np.ones(3, bool)
            return dict(positions=np.asarray(atoms.get_positions(), dtype=float),
                        numbers=np.asarray(num) if num is not None else np.asarray(syms),
                        cell=np.asarray(atoms.get_cell(), dtype=float), pbc=pbc)

        def check_state(self, atoms, tol=1e-15):
            if self.atoms is None:
                return list(all_changes)
            a, b = self._state(self.atoms), self._state(atoms)
            changes = []
            for k in ("positions", "numbers", "cell", "pbc"):
                x, y = a[k], b[k]
                if x.shape != y.shape:
                    changes.append(k)
                elif x.dtype.kind == "f":
                    if np.abs(x - y).max(initial=0.0) > tol:
                        changes.append(k)
                elif not np.array_equal(x, y):
                    changes.append(k)
            return changes

        def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms.copy()

        def get_property(self, name, atoms=None, allow_calculation=True):
            if name not in self.implemented_properties:
                raise NotImplementedError("{} property not implemented".format(name))
            if atoms is None:
                atoms = self.atoms
                system_changes = []
            else:
                system_changes = self.check_state(atoms)
                if system_changes:
                    self.results = {}
            if name not in self.results:
                if not allow_calculation:
                    return None
                self.calculate(atoms, [name], system_changes)
            return np.copy(self.results[name]) if isinstance(self.results[name], np.ndarray) \
                else self.results[name]

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("energy", atoms)

        def get_forces(self, atoms=None):
            return self.get_property("forces", atoms)

        def get_stress(self, atoms=None):
            return self.get_property("stress", atoms)

FORCE_EV_A = HARTREE_EV / BOHR  # Ha/bohr -> eV/Angstrom (constants.py:72)
STRESS_EV_A3 = HARTREE_EV / BOHR ** 3  # Ha/bohr^3 -> eV/Angstrom^3 (constants.py:75)


class DFTpyCalculator(Calculator):
    """DFTpy calculator for ase (api/api4ase.py:11-119)."""

    implemented_properties = ["energy", "forces", "stress"]

    def __init__(self, config=None, mp=None, optimizer=None, evaluator=None, rho=None, zero_stress=False,
                 **kwargs):
        Calculator.__init__(self, **kwargs)
        self.config = config
        self.mp = mp
        self.optimizer = optimizer
        self.evaluator = evaluator
        self.zero_stress = zero_stress
        self.dftpy_results = {"density": rho}

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        """api4ase.py:25-50."""
        Calculator.calculate(self, atoms, properties, system_changes)
        if len(system_changes) > 0:
            self.run(system_changes, properties=properties)
        self.energy = self.dftpy_results["energypotential"]["TOTAL"].energy * HARTREE_EV
        self.results["energy"] = self.energy
        if "forces" in properties:
            if "TOTAL" not in self.dftpy_results["forces"]:
                self.calc_forces()
            self.forces = self.dftpy_results["forces"]["TOTAL"] * FORCE_EV_A
            self.results["forces"] = self.forces
        if "stress" in properties:
            voigt = np.zeros(6)
            if not self.zero_stress:
                if "TOTAL" not in self.dftpy_results["stress"]:
                    self.calc_stress()
                s = self.dftpy_results["stress"]["TOTAL"]
                voigt[:3] = np.diag(s)
                voigt[3], voigt[4], voigt[5] = s[1, 2], s[0, 2], s[0, 1]
            self.stress = voigt * STRESS_EV_A3
            self.results["stress"] = self.stress

    def run(self, system_changes=all_changes, properties=("energy",)):
        """api4ase.py:52-60."""
        if self.config is not None:
            self.run_config(system_changes=system_changes, properties=properties)
        elif self.evaluator is not None:
            self.run_normal(system_changes=system_changes, properties=properties)
        else:
            raise ValueError("!!!Error: Please give a 'config' or 'evaluator' for the calculator")

    def run_config(self, system_changes=all_changes, properties=("energy",), **kwargs):
        """The INI job for the current atoms (api4ase.py:62-74): the
        density and pseudopotential of the last run are reused when
        MATH.reuse is set and the cell is unchanged."""
        pseudo = self.dftpy_results.get("pseudo", None)
        rho = self.dftpy_results.get("density", None)
        ions = Ions.from_ase(self.atoms)
        grid = None
        if "cell" not in system_changes and rho is not None:
            grid = rho.grid
        if not self.config["MATH"]["reuse"]:
            rho = None
        if pseudo is not None:
            if "cell" in system_changes or grid is None:
                pseudo = None
            else:
                pseudo.restart(ions=ions, grid=grid)
        config, others = ConfigParser(self.config, ions=ions, rhoini=rho, pseudo=pseudo, grid=grid)
        self.dftpy_results = OptimizeDensityConf(config, **others)

    def run_normal(self, system_changes=all_changes, properties=("energy",), **kwargs):
        """Optimize from the last density with the given evaluator
        (api4ase.py:76-107). A new cell keeps the grid shape; the density
        values are carried over and renormalised to the ions' charge."""
        rho = self.dftpy_results.get("density", None)
        ions = Ions.from_ase(self.atoms)
        if rho is None:
            raise ValueError("!!!Error: Please give a initial density for the calculator")
        grid = rho.grid
        if "cell" in system_changes:
            grid = DirectGrid(lattice=ions.cell, nr=grid.nrR, full=grid.full, device=grid.get_device())
            with torch._C.DisableTorchFunctionSubclass():
                rho = DirectField(grid=grid, data=rho.detach().clone())
        for key, func in self.evaluator.funcDict.items():
            if func.type == "PSEUDO":
                func.restart(ions=ions, grid=grid)
        if "cell" in system_changes:
            ntot = ions.get_ncharges()
            with torch._C.DisableTorchFunctionSubclass():
                rho.mul_(ntot / (float(rho.sum()) * grid.dV))
        if self.optimizer is None:
            self.optimizer = Optimization(EnergyEvaluator=self.evaluator)
        rho = self.optimizer.optimize_rho(guess_rho=rho)
        calc = {"E"}
        if "forces" in properties:
            calc.add("F")
        if "potential" in properties:
            calc.add("V")
        if "stress" in properties and not self.zero_stress:
            calc.add("S")
        self.dftpy_results = evaluator2results(self.evaluator, rho=rho, calcType=calc, ions=ions, **kwargs)

    def calc_forces(self):
        """api4ase.py:109-113."""
        evaluator = self.dftpy_results.get("evaluator", None)
        rho = self.dftpy_results.get("density", None)
        results = evaluator2results(evaluator, rho=rho, calcType={"F"})
        self.dftpy_results["forces"] = results.get("forces", None)

    def calc_stress(self):
        """api4ase.py:115-119."""
        evaluator = self.dftpy_results.get("evaluator", None)
        rho = self.dftpy_results.get("density", None)
        results = evaluator2results(evaluator, rho=rho, calcType={"S"})
        self.dftpy_results["stress"] = results.get("stress", None)
