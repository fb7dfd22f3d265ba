"""Local pseudopotential and Ewald ion-ion energy on the GPU.

LocalPseudo mirrors functional/pseudo/__init__.py:74-571: PP_list maps a
symbol to a CASTEP .recpot file (pseudo/recpot.py:11-41) or to a
``RadialPseudo`` table (the BasePseudo injection path,
abstract_pseudo.py:49-111).  The radial table is fitted once on the host
with the reference's own call (scipy ``splrep(gp[1:], vp[1:], k=3)``,
pseudo/__init__.py:696); V_loc(r), forces and the Ewald sums run in HIP
kernels (``ofdft_vloc_build``, ``ofdft_local_forces``, ``ofdft_ewald``).
"""
import ctypes
import math
import os

import numpy as np
import torch
from scipy.interpolate import splrep
from scipy import special

from . import _lib as L
from .field import DirectField
from .functional import AbstractFunctional, FunctionalOutput
from .workloads import BOHR, HARTREE


class RadialPseudo:
    """Reciprocal-space radial local PP: q grid ``r`` (bohr^-1), values
    ``v`` (Ha bohr^3), valence ``zval`` (BasePseudo attributes)."""

    def __init__(self, r, v, zval, info=None):
        self.r = np.asarray(r, dtype=np.float64)
        self.v = np.asarray(v, dtype=np.float64)
        self._zval = zval
        self.info = info or {}

    @property
    def zval(self):
        return self._zval

    radial_grid = property(lambda self: self.r)
    local_potential = property(lambda self: self.v)


def read_recpot(fname):
    """CASTEP recpot table v(q) (the format pseudo/recpot.py:11-41 reads):
    after the "END COMMENT" line come a version line, q_max in 1/Angstrom,
    then v(q) in eV Angstrom^3 on an even grid from 0 to q_max, closed by a
    terminator line ("1000", or any one-character line).  Returned in
    atomic units, with the valence charge read off the Coulomb tail
    v ~ -4 pi Z / q^2 between the first two points (as the reference
    rounds it)."""
    with open(fname) as fh:
        lines = [l.strip() for l in fh]
    try:
        start = next(i for i, l in enumerate(lines) if "END COMMENT" in l) + 2
        stop = next(i for i in range(start + 1, len(lines)) if lines[i] == "1000" or len(lines[i]) == 1)
    except StopIteration:
        raise AttributeError("recpot file %s: no END COMMENT header or no terminator line after the "
                             "v(q) table" % fname) from None
    qmax = float(lines[start]) * BOHR
    v = np.array(" ".join(lines[start + 1:stop]).split(), dtype=np.float64) / HARTREE / BOHR ** 3
    q = np.linspace(0.0, qmax, num=len(v))
    zval = round((v[0] - v[1]) * q[1] ** 2 / (4.0 * np.pi))
    return RadialPseudo(q, v, zval)


def sym6(v):
    """3x3 symmetric matrix from the six components xx xy xz yy yz zz."""
    a = np.empty((3, 3))
    a[0, 0], a[0, 1], a[0, 2], a[1, 1], a[1, 2], a[2, 2] = v
    a[1, 0], a[2, 0], a[2, 1] = a[0, 1], a[0, 2], a[1, 2]
    return a

def empty_core(Z, rc, qmax=30.0, npts=2001):
    """Ashcroft empty-core table as a RadialPseudo (zval recovered as
    recpot.py:41 does)."""
    from .workloads import empty_core_table
    q, v = empty_core_table(Z, rc, qmax, npts)
    zval = round((v[0] - v[1]) * (q[1] ** 2) / (4.0 * np.pi))
    return RadialPseudo(q, v, zval)


def _load_pp(item):
    if isinstance(item, RadialPseudo):
        return item
    if hasattr(item, "radial_grid") and hasattr(item, "local_potential"):
        return RadialPseudo(item.radial_grid, item.local_potential, item.zval)
    fname = os.fspath(item)
    if not os.path.isfile(fname):
        raise FileNotFoundError("'{}' PP file not found".format(fname))
    if fname.lower().endswith(".recpot"):
        return read_recpot(fname)
    raise AttributeError("Pseudopotential '{}' is not supported on the device path".format(fname))


class LocalPseudo(AbstractFunctional):
    """Local pseudopotential term (pseudo/__init__.py:74-271)."""

    def __init__(self, grid=None, ions=None, PP_list=None, PME=True, BsplineOrder=10, **kwargs):
        self.type = "PSEUDO"
        self.name = "PSEUDO"
        if PP_list is None:
            raise AttributeError("Must specify PP_list for Pseudopotentials")
        if grid is None:
            raise AttributeError("Must specify grid for Pseudopotentials")
        if ions is None:
            raise AttributeError("Must specify ions for Pseudopotentials")
        self.PP_list = dict(PP_list)
        self.pps = {k: _load_pp(v) for k, v in self.PP_list.items()}
        for s in ions.symbols_uniq:
            if s not in self.pps:
                raise ValueError("There is no pseudopotential for {:4s} atom".format(s))
        self.grid = grid
        self.ions = ions
        self.ions.set_charges({k: p.zval for k, p in self.pps.items()})
        self.PME = PME
        self.BsplineOrder = BsplineOrder
        self._vreal = None
        self._vreal32 = None
        self._ewald = None
        self._plan = None
        self._version = 0

    def restart(self, grid=None, ions=None, full=False, duplicate=False):
        """Drop V_loc, the Ewald object and the plan so that they are built
        again for new ions and/or a new grid (pseudo/__init__.py:152-176);
        ``duplicate=True`` returns a new object instead."""
        if duplicate:
            return LocalPseudo(grid=grid if grid is not None else self.grid,
                               ions=ions if ions is not None else self.ions,
                               PP_list=self.PP_list, PME=self.PME, BsplineOrder=self.BsplineOrder)
        if ions is not None:
            for s in ions.symbols_uniq:
                if s not in self.pps:
                    raise ValueError("There is no pseudopotential for {:4s} atom".format(s))
            self.ions = ions
            self.ions.set_charges({k: p.zval for k, p in self.pps.items()})
        if grid is not None:
            if self._plan is not None and (tuple(grid.nrR) != tuple(self.grid.nrR)
                                           or not np.array_equal(grid.lattice, self.grid.lattice)):
                self._plan = None
            self.grid = grid
        self._vreal = None
        self._vreal32 = None
        self._ewald = None
        self._tot = None
        # device evaluators built over the old V_loc see a new version
        self._version += 1

    @property
    def zval(self):
        return {k: p.zval for k, p in self.pps.items()}

    def _fp64_plan(self):
        if self._plan is None:
            from .plan import Plan
            g = self.grid
            self._plan = Plan(tuple(int(x) for x in g.nrR), g.lattice, precision="fp64", full=g.full,
                              device=g.get_device())
        return self._plan

    def _tables(self):
        keys = sorted(self.pps)
        ctabs, keep = [], []
        for k in keys:
            p = self.pps[k]
            t, c, _ = splrep(p.r[1:], p.v[1:], k=3)
            t = np.ascontiguousarray(t, dtype=np.float64)
            c = np.ascontiguousarray(c, dtype=np.float64)
            keep += [t, c]
            ctabs.append(L.VLine(t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(t), p.r[1],
                                 p.r[1] - p.r[0], p.r[-1], p.v[0], p.v[1], p.v[2]))
        arr = (L.VLine * len(ctabs))(*ctabs)
        return keys, arr, keep

    def local_PP(self):
        """V_loc(r) on the device (pseudo/__init__.py:237-246)."""
        if self._vreal is not None:
            return
        P = self._fp64_plan()
        keys, arr, keep = self._tables()
        idx = {k: i for i, k in enumerate(keys)}
        species = np.array([idx[s] for s in self.ions.symbols], dtype=np.int32)
        pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
        out = DirectField(grid=self.grid, dtype=torch.float64)
        P._stream()
        with torch._C.DisableTorchFunctionSubclass():
            L.check(P.lib.ofdft_vloc_build(P._h, len(keys), arr, self.ions.nat,
                                           pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           species.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           1 if self.PME else 0, int(self.BsplineOrder), L.ptr(out)))
        torch.cuda.current_stream(P.device).synchronize()
        del keep
        self._vreal = out

    @property
    def vreal(self):
        self.local_PP()
        return self._vreal

    def vreal_device(self, dtype):
        """V_loc as a plain device tensor in the requested precision."""
        v = self.vreal
        with torch._C.DisableTorchFunctionSubclass():
            if dtype == torch.float64:
                return v.contiguous()
            if self._vreal32 is None:
                self._vreal32 = v.to(torch.float32).contiguous()
            return self._vreal32

    def _device_terms(self):
        return dict(pseudo=True, obj=self, owner=self)

    def compute(self, density, calcType=("E", "V"), **kwargs):
        """E = int V_loc rho (pseudo/__init__.py:219-234)."""
        from .functional import TotalFunctional
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(PSEUDO=self)
            self._tot._ewald = False
        ev = self._tot.evaluator(density.grid, density.dtype)
        with torch._C.DisableTorchFunctionSubclass():
            sums = ev.eval_rho(density.contiguous(), None)
        out = FunctionalOutput(name="eN", energy=float(sums[L.E_PP]))
        if "V" in calcType:
            out.potential = self.vreal if density.dtype == torch.float64 else \
                DirectField(grid=self.grid, data=self.vreal_device(density.dtype))
        return out

    def forces(self, rho):
        """Local-PP ionic forces (pseudo/__init__.py:260-270), Ha/bohr."""
        if rho is None:
            raise AttributeError("Must specify rho")
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be DirectField")
        self.local_PP()
        P = self._fp64_plan()
        out = np.zeros((self.ions.nat, 3))
        with torch._C.DisableTorchFunctionSubclass():
            r = rho.to(torch.float64).contiguous()
            P._stream()
            L.check(P.lib.ofdft_local_forces(P._h, L.ptr(r), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def stress(self, rho, energy=None):
        """Local-PP stress (pseudo/__init__.py:248-258): _StressPME or _Stress,
        Ha/bohr^3, 3x3."""
        if rho is None:
            raise AttributeError("Must specify rho")
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be DirectField")
        self.local_PP()
        if energy is None:
            energy = self.compute(rho, calcType=("E",)).energy
        P = self._fp64_plan()
        s6 = np.zeros(6)
        with torch._C.DisableTorchFunctionSubclass():
            r = rho.to(torch.float64).contiguous()
            P._stream()
            L.check(P.lib.ofdft_local_stress(P._h, L.ptr(r), s6.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        g = self.grid
        dV, vol = g.dV, g.volume
        if self.PME:
            # -sum / Omega^2 with fft() = dV rfftn() on rho and Q, times 2 nnr
            st = -sym6(s6) * dV * dV / vol ** 2 * 2.0 * float(np.prod(g.nrR))
        else:
            st = -sym6(s6) * dV / vol * 2.0
        st = st - np.eye(3) * energy
        return st / vol

    def get_ewald(self, PME=None):
        if self._ewald is None:
            self._ewald = Ewald(ions=self.ions, grid=self.grid, PME=self.PME if PME is None else PME,
                                BsplineOrder=self.BsplineOrder, plan=self._fp64_plan())
        return self._ewald


class Ewald:
    """Ion-ion energy (ewald.py:197-507): eta from Get_Best_eta (:256-278),
    Energy_corr (:468-482) on the host; the real-space and reciprocal sums
    in HIP kernels."""

    def __init__(self, precision=1.0e-8, ions=None, grid=None, PME=False, BsplineOrder=10, plan=None, **kwargs):
        if ions is None:
            raise AttributeError("Must pass ions to Ewald")
        if grid is None:
            raise AttributeError("Must pass grid or rho to Ewald")
        self.ions = ions
        self.grid = grid
        self.precision = precision
        self.PME = PME
        self.order = BsplineOrder
        self._plan = plan
        gmax = math.sqrt(grid.get_reciprocal().g2max)
        self.eta = self.Get_Best_eta(precision, gmax, ions)
        self._energy = None
        self._forces = None

    def Get_Best_eta(self, precision, gmax, ions):
        q2 = float(np.sum(ions.charges * ions.charges))
        eta = 1.6
        while True:
            ub = 4.0 * np.pi * ions.nat * q2 * np.sqrt(eta / np.pi) * special.erfc(gmax / 2.0 * np.sqrt(1.0 / eta))
            if ub < precision:
                return eta
            eta = eta - 0.01

    def Energy_corr(self):
        dc = -np.sqrt(self.eta / np.pi) * float(np.sum(self.ions.charges ** 2))
        gz = -4.0 * np.pi * (1.0 / (4.0 * self.eta * self.grid.volume) / 2.0) * self.ions.get_ncharges() ** 2
        return dc + gz

    def _fp64_plan(self):
        if self._plan is None:
            from .plan import Plan
            g = self.grid
            self._plan = Plan(tuple(int(x) for x in g.nrR), g.lattice, "fp64", full=g.full, device=g.get_device())
        return self._plan

    def _sums(self):
        P = self._fp64_plan()
        rmax = special.erfcinv(self.precision / 3.0) / np.sqrt(self.eta)
        pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
        q = np.ascontiguousarray(self.ions.charges, dtype=np.float64)
        er, ek = ctypes.c_double(), ctypes.c_double()
        P._stream()
        L.check(P.lib.ofdft_ewald(P._h, self.ions.nat, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(self.eta), float(rmax),
                                  1 if self.PME else 0, int(self.order), ctypes.byref(er), ctypes.byref(ek)))
        return er.value, ek.value

    @property
    def energy(self):
        if self._energy is None:
            er, ek = self._sums()
            self._energy = self.Energy_corr() + er + ek
        return self._energy

    @property
    def stress(self):
        """Ion-ion stress (ewald.py:523-538): Stress_real plus Stress_rec_PME
        or Stress_rec, Ha/bohr^3, 3x3."""
        if getattr(self, "_stress", None) is None:
            P = self._fp64_plan()
            rmax = special.erfcinv(self.precision / 3.0) / np.sqrt(self.eta)
            pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
            q = np.ascontiguousarray(self.ions.charges, dtype=np.float64)
            real6, rec6 = np.zeros(6), np.zeros(6)
            P._stream()
            dp = ctypes.POINTER(ctypes.c_double)
            L.check(P.lib.ofdft_ewald_stress(P._h, self.ions.nat, pos.ctypes.data_as(dp), q.ctypes.data_as(dp),
                                             float(self.eta), float(rmax), 1 if self.PME else 0, int(self.order),
                                             real6.ctypes.data_as(dp), rec6.ctypes.data_as(dp)))
            vol = self.grid.volume
            # Stress_rec_PME doubles the half-space sum and scales by 2 pi/Omega^2
            # (|fft(Q)|^2 / dV^2 = |rfftn(Q)|^2); Stress_rec scales by 4 pi/Omega^2
            rec = sym6(rec6) * (4.0 * np.pi / vol ** 2)
            s_g0 = self.ions.get_ncharges() ** 2 * 4.0 * np.pi * (1.0 / (4.0 * self.eta * vol ** 2) / 2.0)
            self._stress = sym6(real6) * (-0.5 / vol) + rec + np.eye(3) * s_g0
        return self._stress

    @property
    def forces(self):
        """Ion-ion forces (ewald.py:509-519): Forces_real plus Forces_rec_PME
        or Forces_rec, Ha/bohr, shape (nat, 3)."""
        if self._forces is None:
            P = self._fp64_plan()
            rmax = special.erfcinv(self.precision / 3.0) / np.sqrt(self.eta)
            pos = np.ascontiguousarray(self.ions.positions, dtype=np.float64)
            q = np.ascontiguousarray(self.ions.charges, dtype=np.float64)
            out = np.zeros((self.ions.nat, 3))
            P._stream()
            dp = ctypes.POINTER(ctypes.c_double)
            L.check(P.lib.ofdft_ewald_forces(P._h, self.ions.nat, pos.ctypes.data_as(dp), q.ctypes.data_as(dp),
                                             float(self.eta), float(rmax), 1 if self.PME else 0, int(self.order),
                                             out.ctypes.data_as(dp)))
            self._forces = out
        return self._forces
