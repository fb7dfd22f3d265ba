"""A stand-in for ase.Atoms (ase is not installed): symbols, positions and
cell in Angstrom, periodic in all directions, with the accessors the
calculator reads."""
import numpy as np

from dftpy_amd.ions import ATOMIC_NUMBERS


class Atoms:
    def __init__(self, symbols, positions, cell):
        self.symbols = list(symbols)
        self.positions = np.array(positions, dtype=float).reshape(-1, 3)
        self.cell = np.array(cell, dtype=float).reshape(3, 3)
        self.calc = None

    def get_chemical_symbols(self):
        return list(self.symbols)

    def get_atomic_numbers(self):
        return np.array([ATOMIC_NUMBERS[s] for s in self.symbols])

    def get_positions(self):
        return self.positions.copy()

    def set_positions(self, p):
        self.positions = np.array(p, dtype=float).reshape(-1, 3)

    def get_cell(self):
        return self.cell.copy()

    def set_cell(self, cell, scale_atoms=False):
        cell = np.array(cell, dtype=float).reshape(3, 3)
        if scale_atoms:
            frac = np.linalg.solve(self.cell.T, self.positions.T).T
            self.positions = frac @ cell
        self.cell = cell

    def get_pbc(self):
        return np.ones(3, dtype=bool)

    def copy(self):
        return Atoms(self.symbols, self.positions, self.cell)

    def get_potential_energy(self):
        return self.calc.get_potential_energy(self)

    def get_forces(self):
        return self.calc.get_forces(self)

    def get_stress(self):
        return self.calc.get_stress(self)
