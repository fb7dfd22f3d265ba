"""Functional plugin layer with the reference's names and contracts.

``AbstractFunctional.compute(rho, calcType, **kwargs) -> FunctionalOutput``
(functional/abstract_functional.py:12-44), ``FunctionalOutput``
(functional/functional_output.py:8-111), the ``Functional(type=...)`` factory
(functional/__init__.py:21-51), ``KEDF`` (kedf/__init__.py:95-191), ``XC``
(xc/__init__.py:9-69), ``Hartree`` (hartree.py:11-45) and ``TotalFunctional``
(total_functional.py:5-174).

Every term is evaluated on the GPU.  A TotalFunctional made of these
classes is fused: one call runs one device evaluation (five transform
launches for up to three convolutions plus the pointwise terms) instead of
one pass per term.
"""
import copy
import ctypes
from abc import ABC, abstractmethod

import numpy as np
import torch

from . import _lib as L
from .field import DirectField
from .plancache import plan_for

__all__ = ["FunctionalOutput", "AbstractFunctional", "KEDF", "XC", "Hartree", "TotalFunctional", "Functional",
           "ZeroFunctional"]


class FunctionalOutput:
    """Energy / potential record with + - * / (functional_output.py:8-111)."""

    attr_list = ["energy", "potential", "energydensity", "v2rho2", "v3rho3", "v4rho4", "force", "stress"]

    def __init__(self, name=None, energy=None, potential=None, energydensity=None, v2rho2=None, v3rho3=None,
                 v4rho4=None, force=None, stress=None):
        if name is None:
            raise AttributeError("Functional name must be specified")
        self.name = name
        for k, v in (("energy", energy), ("potential", potential), ("energydensity", energydensity),
                     ("v2rho2", v2rho2), ("v3rho3", v3rho3), ("v4rho4", v4rho4), ("force", force),
                     ("stress", stress)):
            if v is not None:
                setattr(self, k, v)

    def __iter__(self):
        for key in self.attr_list:
            if hasattr(self, key):
                yield key, getattr(self, key)

    def copy(self):
        out = FunctionalOutput(name=self.name)
        for k, v in self:
            setattr(out, k, v.clone() if isinstance(v, torch.Tensor) else copy.deepcopy(v))
        return out

    def sum(self, other=None):
        if other is None:
            return self.copy()
        name = self.name if self.name == other.name else self.name + other.name
        result = FunctionalOutput(name=name)
        for key, value in self:
            setattr(result, key, value + getattr(other, key) if hasattr(other, key) else value)
        for key, value in other:
            if not hasattr(result, key):
                setattr(result, key, value)
        return result

    def sub(self, other):
        result = self.copy()
        for key, value in self:
            setattr(result, key, value - getattr(other, key) if hasattr(other, key) else value)
        for key, value in other:
            if not hasattr(result, key):
                setattr(result, key, -value)
        return result

    def mul(self, x):
        result = self.copy()
        for key, value in result:
            setattr(result, key, value * x)
        return result

    def div(self, x):
        if x == 0:
            raise ValueError("Dividing zero")
        result = self.copy()
        for key, value in result:
            setattr(result, key, value / x)
        return result

    __add__ = sum
    __sub__ = sub
    __mul__ = mul
    __truediv__ = div


class AbstractFunctional(ABC):
    def __repr__(self):
        return getattr(self, "name", self.__class__.__name__) + ", " + repr(self.__dict__)

    def __call__(self, rho, *args, **kwargs):
        return self.compute(rho, *args, **kwargs)

    @abstractmethod
    def compute(self, rho, *args, **kwargs):
        pass

    @property
    def name(self):
        if getattr(self, "_name", None) is None:
            self._name = self.type
        return self._name

    @name.setter
    def name(self, value):
        self._name = value

    @property
    def type(self):
        if getattr(self, "_type", None) is None:
            self._type = self.__class__.__name__
        return self._type

    @type.setter
    def type(self, value):
        self._type = value

    # fusion hook: the device terms this functional contributes
    def _device_terms(self):
        return None


class ZeroFunctional(AbstractFunctional):
    def __init__(self, *args, **kwargs):
        self.type = "ZERO"
        self.name = "ZERO"

    def compute(self, rho, *args, calcType=("E", "V"), **kwargs):
        out = FunctionalOutput(name="ZERO", energy=0.0)
        if "V" in calcType:
            out.potential = DirectField(grid=rho.grid, dtype=rho.dtype)
        return out


# KEDF name expansion (kedf/__init__.py:27-69, restricted to the hot path)
_KEDF_NAMES = {
    "TF": ("TF",), "VW": ("VW",), "TFVW": ("TF", "VW"), "X_TF_Y_VW": ("TF", "VW"), "XTFYVW": ("TF", "VW"),
    "WT": ("TF", "VW", "WT-NL"), "WT-NL": ("WT-NL",), "NONE": (),
}


class _KernelCache(dict):
    """ke_kernel_saved (kedf/__init__.py:101-111): WT kernel density and shape."""

    def __init__(self):
        super().__init__(rho0=0.0, shape=None)


class KEDF(AbstractFunctional):
    """Kinetic functional 'TF', 'VW', 'TFVW' (x_TF_y_vW) or 'WT'.

    Options follow the reference: x (TF fraction and WT kernel lambda),
    y (vW fraction), alpha = beta (WT exponents, 5/6), rho0 (fixed WT
    kernel density; default: mean density, kernel rebuilt when it moves by
    more than 1e-6 as in wt.py:64-72)."""

    def __init__(self, name="WT", kedf=None, **kwargs):
        self.type = "KEDF"
        self.name = (kedf or name)
        key = self.name.upper()
        if key not in _KEDF_NAMES:
            raise AttributeError("%s KEDF to be implemented" % self.name)
        self.options = dict(x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0=None)
        self.options.update({k: v for k, v in kwargs.items() if k in ("x", "y", "alpha", "beta", "rho0", "sigma")})
        if self.options.get("sigma") is not None:
            raise AttributeError("vW smoothing (sigma) is not on the device path")
        self.options["kedf"] = self.name
        self.engines = _KEDF_NAMES[key]
        self.ke_kernel_saved = _KernelCache()
        self.energies = {}

    def _device_terms(self):
        o = self.options
        return dict(tf="TF" in self.engines, vw="VW" in self.engines, wt="WT-NL" in self.engines,
                    x=o["x"], y=o["y"], alpha=o["alpha"], beta=o["beta"], rho0=o["rho0"], owner=self)

    def _total(self):
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(KEDF=self)
        return self._tot

    def compute(self, density, calcType=("E", "V"), split=False, **kwargs):
        tot = self._total()
        if split:
            out = tot.get_energy_potential(density, calcType=calcType, split=True)
            return {k.split("-", 1)[1] if k.startswith("KEDF-") else k: v for k, v in out.items()
                    if k.startswith("KEDF-")}
        return tot.compute(density, calcType=calcType)


    def stress(self, density, name=None, kedf=None, split=False, **kwargs):
        """KEDF.stress (kedf/__init__.py:193-235): per engine ("TF", "VW",
        "WT-NL") when split, else their sum; energies at ``density``."""
        out, _ = _density_stress(self._total(), density)
        if split:
            return out
        tot = np.zeros((3, 3))
        for v in out.values():
            tot = tot + v
        return tot


class XC(AbstractFunctional):
    """LDA exchange-correlation (Slater + Perdew-Zunger, the in-house path of
    xc/semilocal_xc.py:283-341; LibXC is not on the device path)."""

    def __init__(self, xc=None, core_density=None, libxc=None, name=None, **kwargs):
        xc = xc or name
        self.type = "XC"
        self.name = xc or "XC"
        if not xc or str(xc).upper() != "LDA":
            raise AttributeError("only the in-house LDA is on the device path (got %r)" % (xc,))
        if core_density is not None:
            raise AttributeError("nonlinear core correction is not on the device path")
        self.options = dict(kwargs)
        self.energy = None

    def _device_terms(self):
        return dict(lda=True, owner=self)

    def compute(self, density, calcType=("E", "V"), **kwargs):
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(XC=self)
        out = self._tot.compute(density, calcType=calcType)
        if "E" in calcType:
            self.energy = out.energy
        return out

    def stress(self, density, pseudo=None, energy=None, **kwargs):
        """XC.stress (xc/__init__.py:72-87) with _LDAStress
        (xc/semilocal_xc.py:343-346): (E - sum V_xc rho dV) / Omega on the
        diagonal; E is ``energy`` when given, else the LDA energy at density."""
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(XC=self)
        out, _ = _density_stress(self._tot, density, energy_xc=energy)
        return out["XC"]


class Hartree(AbstractFunctional):
    """Hartree: V = IFFT[4 pi rho(G)/G^2], E = 1/2 int V rho (hartree.py:24-45)."""

    _energy = None

    def __init__(self, **kwargs):
        self.type = "HARTREE"
        self.name = "HARTREE"
        self.options = kwargs

    def __repr__(self):
        return "HARTREE"

    def _device_terms(self):
        return dict(hartree=True, owner=self)

    def compute(self, density, calcType=("E", "V"), **kwargs):
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(HARTREE=self)
        out = self._tot.compute(density, calcType=calcType)
        if "E" in calcType:
            Hartree._energy = out.energy
        return out

    @property
    def energy(self):
        return self._energy

    def stress(self, density, **kwargs):
        """HartreeFunctionalStress (hartree.py:51-90) with the Hartree energy
        at ``density``."""
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(HARTREE=self)
        out, _ = _density_stress(self._tot, density)
        return out["HARTREE"]


_SPLIT_KEYS = [("KEDF-TF", L.E_TF, "tf"), ("KEDF-VW", L.E_VW, "vw"), ("KEDF-NL", L.E_NL, "wt"),
               ("XC", L.E_XC, "lda"), ("HARTREE", L.E_H, "hartree"), ("PSEUDO", L.E_PP, "pseudo")]


def sym6(v):
    """3x3 symmetric matrix from the six components xx xy xz yy yz zz."""
    a = np.empty((3, 3))
    a[0, 0], a[0, 1], a[0, 2], a[1, 1], a[1, 2], a[2, 2] = v
    a[1, 0], a[2, 0], a[2, 1] = a[0, 1], a[0, 2], a[1, 2]
    return a


def _density_stress(total, rho, energy_xc=None):
    """Stress of the density terms of ``total`` at rho, keyed like the
    reference's split output (KEDF parts by their engine names "TF", "VW",
    "WT-NL"; "XC"; "HARTREE"), with each term's energy at rho.  One FP64
    energy evaluation (which also builds the WT kernel) and one
    ofdft_functional_stress call; the reference's factors:
    TF kedf/tf.py:52-62, vW kedf/vw.py:97-112, WT kedf/wt.py:75-106,
    LDA xc/semilocal_xc.py:343-346 (xc/__init__.py:72-87), Hartree
    hartree.py:68-90."""
    if not isinstance(rho, DirectField):
        raise TypeError("rho must be DirectField")
    grid = rho.grid
    ev = total.evaluator(grid, torch.float64)
    t = ev.terms
    with torch._C.DisableTorchFunctionSubclass():
        r = rho.to(torch.float64).contiguous()
        sums = ev.eval_rho(r, None)
        rho0 = float(r.mean())
    parts = ev.energies(sums)
    tkf = 2.0 * (3.0 * rho0 * np.pi ** 2) ** (1.0 / 3.0)
    raw = np.zeros(L.NSTRESS)
    ev.P._stream()
    with torch._C.DisableTorchFunctionSubclass():
        L.check(ev.P.lib.ofdft_functional_stress(ev.P._h, L.ptr(r), float(tkf),
                                                 raw.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    vol, dV = grid.volume, grid.dV
    eye = np.eye(3)
    out = {}
    if t["tf"]:
        out["TF"] = eye * (-2.0 / 3.0 * parts["KEDF-TF"] / vol)
    if t["vw"]:
        out["VW"] = sym6(raw[6:12]) * (-0.25 / vol * dV * t["y"])
    if t["wt"]:
        a, b = t["alpha"], t["beta"]
        s = sym6(raw[12:18]) * (dV * dV / vol ** 2)
        s = s * (np.pi ** 2 / (a * b * rho0 ** (a + b - 2) * tkf / 2.0))
        out["WT-NL"] = s - eye * (2.0 / 3.0 * parts["KEDF-NL"] / vol)
    if t["lda"]:
        e = parts["XC"] if energy_xc is None else energy_xc
        out["XC"] = eye * ((e - raw[18] * dV) / vol)
    if t["hartree"]:
        out["HARTREE"] = sym6(raw[0:6]) * (dV * dV * 8.0 * np.pi / vol ** 2) - eye * (parts["HARTREE"] / vol)
    return out, parts


class DeviceEvaluator:
    """The fused evaluation of a TotalFunctional on one grid and precision:
    holds the device plan with the merged functional switches and V_loc."""

    def __init__(self, total, grid, dtype):
        from .plan import Plan
        self.total = total
        self.grid = grid
        # an evaluator owns its plan: the functional switches live in the plan
        self.P = Plan(tuple(int(x) for x in grid.nrR), grid.lattice,
                      precision="fp64" if dtype == torch.float64 else "fp32", full=grid.full,
                      device=grid.get_device())
        self.dtype = dtype
        self.terms = total._merged_terms()
        self.configured_rho0 = None
        if self.terms["pseudo"]:
            self.P.set_vloc(self.terms["pseudo_obj"].vreal_device(dtype))
        self._configure(self.terms.get("rho0_fixed"))

    def _configure(self, rho0):
        t = self.terms
        wt_on = bool(t["wt"]) and rho0 is not None
        self.P.set_functional(tf=t["tf"], vw=t["vw"], wt=wt_on, lda=t["lda"], hartree=t["hartree"],
                              pseudo=t["pseudo"], x=t["x"], y=t["y"], alpha=t["alpha"], beta=t["beta"],
                              rho0=float(rho0) if wt_on else 0.0)
        self.configured_rho0 = rho0

    def ensure_kernel(self, rho0):
        """WT kernel cache semantics (wt.py:64-72): rebuild when the mean
        density moved by more than 1e-6.  Returns True when rebuilt."""
        if not self.terms["wt"] or self.terms.get("rho0_fixed") is not None:
            return False
        if self.configured_rho0 is None or abs(self.configured_rho0 - rho0) > 1e-6:
            self._configure(rho0)
            return True
        return False

    def energies(self, sums):
        return {k: float(sums[i]) for k, i, flag in _SPLIT_KEYS if self.terms[flag]}

    def eval_rho(self, rho, V=None):
        if self.terms["wt"] and self.terms.get("rho0_fixed") is None:
            with torch._C.DisableTorchFunctionSubclass():
                self.ensure_kernel(float(rho.to(torch.float64).mean()))
        return self.P.eval(rho, V)

    def eval_phi(self, phi, p, c, s, V, aux):
        sums = self.P.eval_phi(phi, p, c, s, V, aux)
        if self.terms["wt"] and self.terms.get("rho0_fixed") is None:
            if self.ensure_kernel(sums[L.S_RHO] / self.grid.volume):
                sums = self.P.eval_phi(phi, p, c, s, V, aux)
        return sums


class TotalFunctional(AbstractFunctional):
    """Sum of functionals plus the ion-ion constant (total_functional.py:5-174)."""

    def __init__(self, ewald=None, **kwargs):
        self.funcDict = {k: v for k, v in kwargs.items() if v is not None}
        self.UpdateNameType()
        self._ewald = ewald
        self._evaluators = {}

    def __repr__(self):
        return self.funcDict.__repr__()

    def __getitem__(self, key):
        return self.funcDict[key]

    def __setitem__(self, key, value):
        self.funcDict.update({key: value})
        self.UpdateNameType()

    def UpdateNameType(self):
        self.name = ""
        self.type = ""
        for key, f in self.funcDict.items():
            if not isinstance(f, AbstractFunctional):
                raise TypeError("{} must be AbstractFunctional".format(key))
            setattr(self, key, f)
            self.name += getattr(f, "name") + " "
            self.type += getattr(f, "type") + " "
        self._evaluators = {}

    def UpdateFunctional(self, keysToRemove=(), newFuncDict=None):
        for key in keysToRemove:
            self.funcDict.pop(key, None)
        self.funcDict.update(newFuncDict or {})
        self.UpdateNameType()

    def Subset(self, keys):
        return TotalFunctional(**{k: self.funcDict[k] for k in keys})

    @property
    def ewald(self):
        """The PSEUDO term's Ewald object unless one was given (:91-105);
        ``ewald=False`` switches the ion-ion term off."""
        if self._ewald is False:
            return None
        if self._ewald is None:
            for f in self.funcDict.values():
                if f.type == "PSEUDO":
                    return f.get_ewald()
            return None
        return self._ewald

    def _merged_terms(self):
        t = dict(tf=False, vw=False, wt=False, lda=False, hartree=False, pseudo=False, x=1.0, y=1.0,
                 alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0_fixed=None, pseudo_obj=None)
        seen = set()
        for key, f in self.funcDict.items():
            d = f._device_terms()
            if d is None:
                raise TypeError("%s (%s) has no device implementation" % (key, type(f).__name__))
            for flag in ("tf", "vw", "wt", "lda", "hartree", "pseudo"):
                if d.get(flag):
                    if flag in seen:
                        raise ValueError("term %s appears twice in the TotalFunctional" % flag)
                    seen.add(flag)
                    t[flag] = True
            if f.type == "KEDF":
                t.update(x=d["x"], y=d["y"], alpha=d["alpha"], beta=d["beta"], rho0_fixed=d["rho0"])
                if d["wt"] and abs(d["alpha"] - d["beta"]) > 1e-9:
                    raise AttributeError("the device WT path requires alpha == beta")
            if d.get("pseudo"):
                t["pseudo_obj"] = d["obj"]
        return t

    def evaluator(self, grid, dtype=torch.float64):
        key = (tuple(int(x) for x in grid.nrR), grid.lattice.tobytes(), bool(grid.full), dtype)
        ev = self._evaluators.get(key)
        if ev is None:
            ev = DeviceEvaluator(self, grid, dtype)
            self._evaluators[key] = ev
        return ev

    def compute(self, rho, calcType=("E", "V"), **kwargs):
        return self.get_energy_potential(rho, calcType=calcType, **kwargs)

    def Energy(self, rho, **kwargs):
        return self.compute(rho, calcType=("E",)).energy

    def get_energy_potential(self, rho, calcType=("E", "V"), split=False, **kwargs):
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be a DirectField")
        if rho.dim() != 3:
            raise NotImplementedError("spin-polarised densities are not on the device path")
        ev = self.evaluator(rho.grid, rho.dtype)
        V = None
        if "V" in calcType or split:
            V = DirectField(grid=rho.grid, dtype=rho.dtype)
        with torch._C.DisableTorchFunctionSubclass():
            sums = ev.eval_rho(rho.contiguous(), V)
        parts = ev.energies(sums)
        ew = self.ewald
        ii = ew.energy if ew else None
        total = FunctionalOutput(name="TOTAL", energy=float(sum(parts.values())) + (ii or 0.0))
        if "V" in calcType:
            total.potential = V
        if not split:
            return total
        out = {"TOTAL": total}
        for k, e in parts.items():
            out[k] = FunctionalOutput(name=k, energy=e)
        if ii is not None:
            out["II"] = FunctionalOutput(name="II", energy=ii)
        return out

    def get_forces(self, rho, ions=None, split=False, **kwargs):
        """Forces of every term that has them plus the ion-ion (Ewald)
        forces (total_functional.py:148-174); Ha/bohr, shape (nat, 3).  On
        this path the local PP is the only density term with forces."""
        ew = self.ewald
        if ions is None:
            if ew:
                ions = ew.ions
            else:
                raise ValueError("Please provide ions for forces")
        forces = {"TOTAL": np.zeros((ions.nat, 3))}
        for key, f in self.funcDict.items():
            if hasattr(f, "forces"):
                fr = f.forces(rho)
                if fr is not None:
                    if split:
                        forces[key] = fr
                    forces["TOTAL"] = forces["TOTAL"] + fr
        if ew:
            fr = ew.forces
            if split:
                forces["II"] = fr
            forces["TOTAL"] = forces["TOTAL"] + fr
        return forces if split else forces["TOTAL"]

    def get_stress(self, rho, split=False, **kwargs):
        """Stress tensor (total_functional.py:176-209), Ha/bohr^3: the KEDF
        parts keyed KEDF-TF, KEDF-VW, KEDF-NL, then XC, HARTREE, PSEUDO and
        the ion-ion II; every term with its energy at rho."""
        dens, parts = _density_stress(self, rho)
        stress = {"TOTAL": np.zeros((3, 3))}
        for k, v in dens.items():
            key = "KEDF-" + k.split("-")[-1] if k in ("TF", "VW", "WT-NL") else k
            stress[key] = v
            stress["TOTAL"] = stress["TOTAL"] + v
        for f in self.funcDict.values():
            if f.type == "PSEUDO":
                s = f.stress(rho, energy=parts.get("PSEUDO"))
                stress["PSEUDO"] = s
                stress["TOTAL"] = stress["TOTAL"] + s
        ew = self.ewald
        if ew:
            stress["II"] = ew.stress
            stress["TOTAL"] = stress["TOTAL"] + stress["II"]
        for i in range(1, 3):
            for j in range(i - 1, -1, -1):
                stress["TOTAL"][i, j] = stress["TOTAL"][j, i]
        return stress if split else stress["TOTAL"]


def Functional(type=None, optional_kwargs=None, **kwargs):
    """Factory by type string (functional/__init__.py:21-51)."""
    from .pseudo import LocalPseudo
    table = {"KEDF": KEDF, "XC": XC, "HARTREE": Hartree, "PSEUDO": LocalPseudo}
    type = type.upper()
    if optional_kwargs is not None:
        kwargs.update(optional_kwargs)
    try:
        cls = table[type]
    except KeyError:
        raise AttributeError("Unknown functional type: {}.".format(type))
    return cls(**kwargs)
