"""Ions: symbols, Cartesian positions (bohr), cell rows (bohr), charges
(the parts of ions.py:5-135 the hot path and the file formats use; the
reference builds on ase.Atoms, which is not a dependency here).

The first three positional arguments are (symbols, positions, cell); the
reference's keyword arguments ``numbers``, ``scaled_positions``,
``charges`` and ``units`` ('au', or 'ase' for Angstrom input converted to
bohr as Ions.convert_units does, ions.py:62-78) are accepted by name."""
import numpy as np

BOHR = 0.5291772105638411  # Angstrom per bohr (ase.units.Bohr, CODATA 2014)

# chemical symbols by atomic number (0 = unknown, 'X')
CHEMICAL_SYMBOLS = (
    "X H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
ATOMIC_NUMBERS = {s: z for z, s in enumerate(CHEMICAL_SYMBOLS)}


def symbol_of(label):
    """A chemical symbol from a symbol or an atomic number (int or digits)."""
    if isinstance(label, (int, np.integer)):
        return CHEMICAL_SYMBOLS[int(label)]
    label = str(label)
    return CHEMICAL_SYMBOLS[int(label)] if label.isdigit() else label


class Ions:
    def __init__(self, symbols=None, positions=None, cell=None, scaled_positions=None, charges=None,
                 numbers=None, units="au"):
        if symbols is None:
            if numbers is None:
                raise AttributeError("Ions need symbols or numbers")
            symbols = [symbol_of(int(z)) for z in np.asarray(numbers).ravel()]
        self.symbols = [symbol_of(s) for s in symbols]
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3).copy()
        if units not in ("au", "ase"):
            raise AttributeError("units must be 'au' or 'ase'")
        if units == "ase":
            self.cell /= BOHR
        if positions is None:
            positions = np.asarray(scaled_positions, dtype=np.float64).reshape(-1, 3) @ self.cell
        else:
            positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
            if units == "ase":
                positions = positions / BOHR
        self.positions = np.asarray(positions, dtype=np.float64).reshape(-1, 3).copy()
        self._charges = None if charges is None else np.asarray(charges, dtype=np.float64)

    @property
    def numbers(self):
        return np.array([ATOMIC_NUMBERS.get(s, 0) for s in self.symbols], dtype=np.int64)

    def get_charges(self):
        """Atomic charges, zeros when none are set (ase's initial charges)."""
        return np.zeros(self.nat) if self._charges is None else self._charges.copy()

    @property
    def zval(self):
        """{symbol: charge} of the set charges (ions.py:107-115)."""
        if self._charges is None:
            return {}
        return {s: float(c) for s, c in zip(self.symbols, self._charges)}

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

    @staticmethod
    def from_ase(atoms):
        """Ions from an ase.Atoms-like object in Angstrom (ions.py:67-69):
        anything with get_chemical_symbols(), get_positions() and
        get_cell(); its initial charges are not carried over."""
        return Ions(symbols=list(atoms.get_chemical_symbols()), positions=np.asarray(atoms.get_positions()),
                    cell=np.asarray(atoms.get_cell()), units="ase")
