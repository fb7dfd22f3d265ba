"""Ions: symbols, Cartesian positions (bohr), cell rows (bohr), charges
(the parts of ions.py:5-135 the hot path uses; the reference builds on
ase.Atoms, which is not a dependency here)."""
import numpy as np


class Ions:
    def __init__(self, symbols, positions=None, cell=None, scaled_positions=None, charges=None):
        self.symbols = list(symbols)
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        if positions is None:
            positions = np.asarray(scaled_positions, dtype=np.float64) @ self.cell
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3).copy()
        self._charges = None if charges is None else np.asarray(charges, dtype=np.float64)

    @property
    def nat(self):
        return len(self.symbols)

    def __len__(self):
        return self.nat

    @property
    def charges(self):
        if self._charges is None:
            raise AttributeError("Please call 'set_charges' before use 'charges'.")
        return self._charges

    def set_charges(self, charges=None):
        """ions.py:94-105."""
        if isinstance(charges, dict):
            vals = []
            for s in self.symbols:
                if s not in charges:
                    raise AttributeError(f"{s} not in the charges")
                vals.append(charges[s])
            charges = vals
        elif isinstance(charges, (float, int)):
            charges = np.ones(self.nat) * charges
        self._charges = np.asarray(charges, dtype=np.float64)

    def get_ncharges(self):
        return float(self.charges.sum())

    def get_scaled_positions(self):
        f = np.linalg.solve(self.cell.T, self.positions.T).T
        f %= 1.0
        f %= 1.0
        return f

    @property
    def symbols_uniq(self):
        return np.sort(np.unique(self.symbols))
