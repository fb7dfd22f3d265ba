"""The ASE calculator's bookkeeping on the CPU (api/api4ase.py:11-119):
when it runs again, the units of energy / forces / stress and the Voigt
order of the stress, with the density optimization replaced by fixed
results."""
import numpy as np

from dftpy_amd import api4ase as A
from dftpy_amd.ions import Ions
from tests.fake_atoms import Atoms

# CODATA 2014 as ase.units computes them: Ha in eV and Bohr in Angstrom
HA_EV = 27.21138602436724
BOHR_A = 0.5291772105638411


class _Out:
    def __init__(self, e):
        self.energy = e


class _Stub(A.DFTpyCalculator):
    """run() replaced: energy = sum of x coordinates (bohr), forces and
    stress fixed, and every run counted."""

    def __init__(self, **kw):
        super().__init__(evaluator=object(), rho=object(), **kw)
        self.runs = []

    def run(self, system_changes=A.all_changes, properties=("energy",)):
        self.runs.append((sorted(system_changes), sorted(properties)))
        ions = Ions.from_ase(self.atoms)
        f = np.arange(ions.nat * 3, dtype=float).reshape(-1, 3)
        s = np.array([[1.0, 6.0, 5.0], [6.0, 2.0, 4.0], [5.0, 4.0, 3.0]])
        self.dftpy_results = {"energypotential": {"TOTAL": _Out(float(ions.positions[:, 0].sum()))},
                              "forces": {"TOTAL": f} if "forces" in properties else {},
                              "stress": {"TOTAL": s}}

    def _lazy(self, kind):
        if kind == "forces" and "TOTAL" not in self.dftpy_results["forces"]:
            n = len(self.atoms.get_chemical_symbols())
            self.dftpy_results["forces"] = {"TOTAL": np.arange(n * 3, dtype=float).reshape(-1, 3)}
        return super()._lazy(kind)


def _atoms():
    a = 4.05
    return Atoms(["Al"] * 4, np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]]) * a, np.eye(3) * a)


def test_from_ase_converts_angstrom_to_bohr():
    at = _atoms()
    ions = Ions.from_ase(at)
    assert ions.symbols == ["Al"] * 4
    assert np.allclose(ions.cell, np.eye(3) * 4.05 / BOHR_A, rtol=0, atol=1e-12)
    assert np.allclose(ions.positions, at.get_positions() / BOHR_A, rtol=0, atol=1e-12)


def test_calculator_units_voigt_and_rerun_on_change():
    at = _atoms()
    calc = _Stub()
    at.calc = calc
    e = at.get_potential_energy()
    x = at.get_positions()[:, 0].sum() / BOHR_A
    assert abs(e - x * HA_EV) < 1e-9 * abs(e)
    assert len(calc.runs) == 1
    # same atoms: no new run for energy; forces come from the lazy evaluation
    assert at.get_potential_energy() == e
    f = at.get_forces()
    assert len(calc.runs) == 1
    assert np.allclose(f, np.arange(12.0).reshape(4, 3) * HA_EV / BOHR_A, rtol=1e-14, atol=0)
    s = at.get_stress()
    # Voigt: xx yy zz yz xz xy
    assert np.allclose(s, np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0]) * HA_EV / BOHR_A ** 3, rtol=1e-14, atol=0)
    # moved atom: one new run, with 'positions' among the changes
    p = at.get_positions()
    p[0, 0] += 0.1
    at.set_positions(p)
    e2 = at.get_potential_energy()
    assert len(calc.runs) == 2 and "positions" in calc.runs[-1][0] and "cell" not in calc.runs[-1][0]
    assert abs(e2 - e - 0.1 / BOHR_A * HA_EV) < 1e-9 * abs(e)
    # new cell
    at.set_cell(at.get_cell() * 1.01, scale_atoms=True)
    at.get_potential_energy()
    assert "cell" in calc.runs[-1][0]


def test_zero_stress_and_missing_inputs():
    at = _atoms()
    calc = _Stub(zero_stress=True)
    at.calc = calc
    assert np.all(at.get_stress() == 0.0)
    bare = A.DFTpyCalculator()
    bare.atoms = at.copy()
    try:
        bare.run()
    except ValueError as e:
        assert "config" in str(e) and "evaluator" in str(e)
    else:
        raise AssertionError("a calculator without config or evaluator must raise")
