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
    """An ase calculator over the GPU density optimization.  Its interface
    is the reference's (api/api4ase.py:11-119): the constructor arguments,
    the ase properties energy / forces / stress and their units.

    ``dftpy_results`` holds the last job's results (density, evaluator,
    energies; forces and stress once asked for).  A change of the atoms
    reruns the density optimization: with ``config`` as an INI job, which
    reuses the last density and the pseudopotential tables while the cell
    is unchanged and MATH.reuse is set; with ``evaluator`` from the last
    density, carried over to a new cell at the same grid shape and
    renormalised to the ions' charge."""

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

    # -- ase ---------------------------------------------------------------
    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        Calculator.calculate(self, atoms, properties, system_changes)
        if system_changes:
            self.run(system_changes, properties=properties)
        self.energy = self.dftpy_results["energypotential"]["TOTAL"].energy * HARTREE_EV
        self.results["energy"] = self.energy
        if "forces" in properties:
            self.forces = self._lazy("forces") * FORCE_EV_A
            self.results["forces"] = self.forces
        if "stress" in properties:
            self.stress = (np.zeros(6) if self.zero_stress else self._voigt(self._lazy("stress"))) * STRESS_EV_A3
            self.results["stress"] = self.stress

    @staticmethod
    def _voigt(s):
        """ase's Voigt order xx, yy, zz, yz, xz, xy of a symmetric 3x3 stress."""
        return np.array([s[0, 0], s[1, 1], s[2, 2], s[1, 2], s[0, 2], s[0, 1]])

    def _lazy(self, kind):
        """The TOTAL forces or stress (Ha/bohr, Ha/bohr^3) of the last
        density, evaluated on first request after a job."""
        have = self.dftpy_results.get(kind) or {}
        if "TOTAL" not in have:
            res = evaluator2results(self.dftpy_results.get("evaluator"), rho=self.dftpy_results.get("density"),
                                    calcType={"F" if kind == "forces" else "S"})
            self.dftpy_results[kind] = res.get(kind)
        return self.dftpy_results[kind]["TOTAL"]

    # -- density optimization ----------------------------------------------
    def run(self, system_changes=all_changes, properties=("energy",)):
        """The density optimization for the current atoms."""
        if self.config is not None:
            self._job_config(system_changes)
        elif self.evaluator is not None:
            self._job_evaluator(system_changes, properties)
        else:
            raise ValueError("DFTpyCalculator has nothing to compute with: give it a config (INI options) "
                             "or an evaluator (TotalFunctional)")

    def _job_config(self, system_changes):
        ions = Ions.from_ase(self.atoms)
        rho = self.dftpy_results.get("density")
        # the last grid (and with it the pseudopotential tables) only while
        # the cell is unchanged; the last density as the start with MATH.reuse
        grid = rho.grid if (rho is not None and "cell" not in system_changes) else None
        pseudo = self.dftpy_results.get("pseudo") if grid is not None else None
        if pseudo is not None:
            pseudo.restart(ions=ions, grid=grid)
        if not self.config["MATH"]["reuse"]:
            rho = None
        config, others = ConfigParser(self.config, ions=ions, rhoini=rho, pseudo=pseudo, grid=grid)
        self.dftpy_results = OptimizeDensityConf(config, **others)

    def _job_evaluator(self, system_changes, properties):
        ions = Ions.from_ase(self.atoms)
        rho = self.dftpy_results.get("density")
        if rho is None:
            raise ValueError("DFTpyCalculator(evaluator=...) needs the initial density: pass rho=DirectField")
        new_cell = "cell" in system_changes
        grid = rho.grid
        if new_cell:
            grid = DirectGrid(lattice=ions.cell, nr=grid.nrR, full=grid.full, device=grid.get_device())
        # the local pseudopotential follows the ions (and sets their charges)
        for func in self.evaluator.funcDict.values():
            if func.type == "PSEUDO":
                func.restart(ions=ions, grid=grid)
        if new_cell:
            with torch._C.DisableTorchFunctionSubclass():
                data = rho.detach().clone()
                data.mul_(ions.get_ncharges() / (float(data.sum()) * grid.dV))
            rho = DirectField(grid=grid, data=data)
        if self.optimizer is None:
            self.optimizer = Optimization(EnergyEvaluator=self.evaluator)
        rho = self.optimizer.optimize_rho(guess_rho=rho)
        wanted = {"E"} | {c for c, p in (("F", "forces"), ("V", "potential")) if p in properties}
        if "stress" in properties and not self.zero_stress:
            wanted.add("S")
        self.dftpy_results = evaluator2results(self.evaluator, rho=rho, calcType=wanted, ions=ions)
