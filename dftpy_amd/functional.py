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
from . import lwt as lwt_mod
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
    "LWT": ("TF", "VW", "LWT-NL"), "LMGP": ("TF", "VW", "LWT-NL"), "LMGPA": ("TF", "VW", "LWT-NL"),
    "LMGPG": ("TF", "VW", "LWT-NL"),
}


# GGA kinetic functionals: name -> (device id, default parameters), the
# registry of kedf/gga.py:577-635 (ids: OFDFT_GGA_* of the C ABI)
_G = dict(LKTVW=0, LKT=1, DK=2, LLP=3, OL=4, THAK=5, B86A=6, B86B=7, DK87=8, PW86=9, PW91O=10, PW91=11, LG94=12,
          P92=13, PBE=14, P82=15, VJKS00=16, VT84F=17, TFVW=18, STV=19, PBE2M=20, PG=21)
GGA_KEDF_LIST = {
    "LKT-VW": (_G["LKTVW"], [1.3]), "LKT": (_G["LKT"], [1.3, 1.0]),
    "DK": (_G["DK"], [0.95, 14.28111, -19.5762, -0.05, 9.99802, 2.96085]),
    "LLP": (_G["LLP"], [0.0044188, 0.0253]), "LLP91": (_G["LLP"], [0.0044188, 0.0253]),
    "OL1": (_G["OL"], [0.00677, 0.0]), "OL": (_G["OL"], [0.00677, 0.0]), "OL2": (_G["OL"], [0.0887, 1.0]),
    "T92": (_G["THAK"], [0.0055, 0.0253, 0.072]), "THAK": (_G["THAK"], [0.0055, 0.0253, 0.072]),
    "B86A": (_G["B86A"], [0.0039, 0.004]), "B86": (_G["B86A"], [0.0039, 0.004]),
    "B86B": (_G["B86B"], [0.00403, 0.007]),
    "DK87": (_G["DK87"], [7.0 / 324.0 / (18.0 * np.pi ** 4) ** (1.0 / 3.0), 0.861504, 0.044286]),
    "PW86": (_G["PW86"], [1.296, 14.0, 0.2]),
    "PW91O": (_G["PW91O"], [0.093907, 0.26608, 0.0809615, 100.0, 76.320, 0.57767e-4]),
    "PW91": (_G["PW91"], [0.19645, 0.2747, 0.1508, 100.0, 7.7956, 0.004]),
    "PW91K": (_G["PW91"], [0.19645, 0.2747, 0.1508, 100.0, 7.7956, 0.004]),
    "LG94": (_G["LG94"], [(1e-8 + 0.1234) / 0.024974, 29.790, 22.417, 12.119, 1570.1, 55.944, 0.024974]),
    "E00": (_G["P92"], [135.0, 28.0, 5.0, 3.0]), "P92": (_G["P92"], [1.0, 88.3960, 16.3683, 88.2108]),
    "PBE2": (_G["PBE"], [2.0309 / 0.2942, 2.0309]), "PBE3": (_G["PBE"], [-3.7425 / 4.1355, -3.7425, 50.258]),
    "PBE4": (_G["PBE"], [-7.2333 / 1.7107, -7.2333, 61.645, -93.683]),
    "P82": (_G["P82"], [5.0 / 27.0]), "TW02": (_G["PBE"], [0.8438, 0.2319]),
    "APBE": (_G["PBE"], [0.804, 0.23889]), "APBEK": (_G["PBE"], [0.804, 0.23889]),
    "REVAPBEK": (_G["PBE"], [1.245, 0.23889]), "REVAPBE": (_G["PBE"], [1.245, 0.23889]),
    "RPBE": (_G["PBE"], [1.9632, 0.01979]), "VJKS00": (_G["VJKS00"], [0.8944, 0.6511, 0.0431]),
    "LC94": (_G["PW91"], [0.093907, 0.26608, 0.0809615, 100.0, 76.32, 0.000057767]),
    "VT84F": (_G["VT84F"], [2.777028126, 2.777028126 - 40.0 / 27.0]),
    "TF": (_G["TFVW"], [1.0, 0.0]), "VW": (_G["TFVW"], [0.0, 1.0]), "X_TF_Y_VW": (_G["TFVW"], [1.0, 1.0]),
    "TFVW": (_G["TFVW"], [1.0, 1.0]), "STV": (_G["STV"], [1.0, 1.0, 0.01, 1.0]),
    "PBE2M": (_G["PBE2M"], [1.0, 0.2942, 2.0309]), "PG": (_G["PG"], [0.75, 1.0]),
}


def gga_spec(functional, params=None):
    """(device id, parameters) with the reference's parameter completion
    (kedf/gga.py:1263-1274)."""
    key = str(functional).upper()
    if key not in GGA_KEDF_LIST:
        raise AttributeError("%s GGA KEDF to be implemented" % functional)
    gid, p0 = GGA_KEDF_LIST[key]
    if params is None:
        return gid, list(p0)
    p = [params] if isinstance(params, (float, int)) else list(params)
    if len(p) > len(p0):
        p = p[:len(p0)]
    elif len(p) < len(p0):
        p = p + list(p0[len(p):])
    return gid, [float(x) for x in p]


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
        self.gga = None
        if key == "GGA" or key.startswith("GGA_"):
            # GGA with k_str (kedf/__init__.py:193-212 name parsing)
            k_str = kwargs.get("k_str") if key == "GGA" else self.name[4:]
            if k_str is None:
                k_str = "LKT"
            self.gga = gga_spec(k_str, kwargs.get("params"))
            self.gga_name = str(k_str).upper()
            key = "GGA"
            if kwargs.get("sigma") is not None or kwargs.get("gga_remove_vw") is not None:
                raise AttributeError("GGA sigma smoothing and gga_remove_vw are not on the device path")
        elif key not in _KEDF_NAMES:
            raise AttributeError("%s KEDF to be implemented" % self.name)
        self.options = dict(x=1.0, y=1.0, alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0=None)
        self.options.update({k: v for k, v in kwargs.items() if k in ("x", "y", "alpha", "beta", "rho0", "sigma")
                             and v is not None})
        if self.options.get("sigma") is not None:
            raise AttributeError("vW smoothing (sigma) is not on the device path")
        self.options["kedf"] = self.name
        self.engines = ("GGA",) if self.gga else _KEDF_NAMES[key]
        self.lwt = None
        if "LWT-NL" in self.engines:
            # LWT, and LMGP / LMGPA / LMGPG = LWT with the MGP kernel tables (lwt.py:569-579)
            lo = {k: v for k, v in kwargs.items() if k in lwt_mod.DEFAULTS}
            lo.update(alpha=self.options["alpha"], beta=self.options["beta"])
            if key != "LWT":
                lo["kerneltype"] = key[1:]
            self.lwt = lwt_mod.LwtTerm(x=self.options["x"], **lo)
        self.ke_kernel_saved = _KernelCache()
        self.energies = {}

    def _device_terms(self):
        o = self.options
        return dict(tf="TF" in self.engines, vw="VW" in self.engines, wt="WT-NL" in self.engines,
                    x=o["x"], y=o["y"], alpha=o["alpha"], beta=o["beta"], rho0=o["rho0"], owner=self,
                    gga=self.gga, lwt=self.lwt)

    def _total(self):
        if getattr(self, "_tot", None) is None:
            self._tot = TotalFunctional(KEDF=self)
        return self._tot

    def compute(self, density, calcType=("E", "V"), split=False, **kwargs):
        tot = self._total()
        if split:
            # keyed by engine name as kedf/__init__.py:158-172 (name2functions)
            out = tot.get_energy_potential(density, calcType=calcType, split=True)
            nl = "WT-NL" if self.lwt is None else self.name.upper() + "-NL"
            names = {"KEDF-TF": "TF", "KEDF-VW": "VW", "KEDF-NL": nl, "KEDF-GGA": "GGA"}
            return {names[k]: v for k, v in out.items() if k in names}
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
    if t["lwt"] is not None:
        # kedf/__init__.py:214 (name2functions over the stress engines, utils.py:289-302)
        raise AttributeError("'LWT' has not implemented.")
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
    if t["gga"] is not None:
        # GGA 'S' branch (kedf/gga.py:1312-1326): p = E - sum pot1 rho - sum_i vkin2 (d_i rho)^2
        gid, prm = t["gga"]
        with torch._C.DisableTorchFunctionSubclass():
            g = ev.P.gga(gid, prm, rho=r)
        pair = sym6(g[4:10])
        pg = g[0] - g[3] - np.trace(pair)
        out["GGA"] = (eye * pg - pair) / vol
    return out, parts


class DeviceEvaluator:
    """The fused evaluation of a TotalFunctional on one grid and precision:
    holds the device plan with the merged functional switches and V_loc."""

    def __init__(self, total, grid, dtype, only=None):
        from .plan import Plan
        self.total = total
        self.grid = grid
        # an evaluator owns its plan: the functional switches live in the plan
        self.P = Plan(tuple(int(x) for x in grid.nrR), grid.lattice,
                      precision="fp64" if dtype == torch.float64 else "fp32", full=grid.full,
                      device=grid.get_device())
        self.dtype = dtype
        self.terms = total._merged_terms()
        if only is not None:
            # one term of the sum on its own (split=True potentials)
            for flag in ("tf", "vw", "wt", "lda", "hartree", "pseudo"):
                self.terms[flag] = self.terms[flag] and flag == only
            for side in ("gga", "lwt"):
                if side != only:
                    self.terms[side] = None
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
        out = {k: float(sums[i]) for k, i, flag in _SPLIT_KEYS if self.terms[flag]}
        if self.terms["gga"] is not None:
            out["KEDF-GGA"] = float(sums[L.NSUM])
        if self.terms["lwt"] is not None:
            out["KEDF-NL"] = float(sums[L.NSUM])
        return out

    def _with_gga(self, sums, **kw):
        """Append the GGA (ofdft_gga) or LWT (ofdft_lwt) kinetic term to the
        fused evaluation's sums: its energy in slot NSUM, its V rho and
        line-search derivative added to S_VRHO and S_DTHETA; V (when given)
        already holds the sum."""
        if self.terms["gga"] is not None:
            gid, prm = self.terms["gga"]
            g = self.P.gga(gid, prm, **kw)
        elif self.terms["lwt"] is not None:
            g = self.terms["lwt"].evaluate(self.P, self.grid, **kw)
        else:
            return sums
        out = np.concatenate([sums, [g[0]]])
        out[L.S_VRHO] += g[2]
        out[L.S_DTHETA] += g[1]
        return out

    def eval_rho(self, rho, V=None):
        if self.terms["wt"] and self.terms.get("rho0_fixed") is None:
            with torch._C.DisableTorchFunctionSubclass():
                self.ensure_kernel(float(rho.to(torch.float64).mean()))
        return self._with_gga(self.P.eval(rho, V), rho=rho, V=V)

    def eval_phi(self, phi, p, c, s, V, aux):
        sums = self.P.eval_phi(phi, p, c, s, V, aux)
        if self.terms["wt"] and self.terms.get("rho0_fixed") is None:
            if self.ensure_kernel(sums[L.S_RHO] / self.grid.volume):
                sums = self.P.eval_phi(phi, p, c, s, V, aux)
        return self._with_gga(sums, phi=phi, p=p, c=c, s=s, V=V)


def _on_device(o, rho):
    """A plugin term's FunctionalOutput with its fields as device
    DirectFields of rho's grid, dtype and device (a plugin may return numpy
    arrays, torch tensors or DirectFields)."""
    if not isinstance(o, FunctionalOutput):
        raise TypeError("a functional must return a FunctionalOutput (got %s)" % type(o).__name__)
    r = FunctionalOutput(name=o.name)
    for k, v in o:
        if k in ("potential", "energydensity", "v2rho2", "v3rho3", "v4rho4") and v is not None:
            t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
            with torch._C.DisableTorchFunctionSubclass():
                t = t.to(device=rho.device, dtype=rho.dtype).reshape(rho.shape)
            v = DirectField(grid=rho.grid, data=t, dtype=rho.dtype)
        elif k == "energy":
            v = float(v)
        setattr(r, k, v)
    return r


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

    def plugin_terms(self):
        """(key, functional) of the terms without a device implementation:
        any AbstractFunctional the caller wrote (the reference's plugin
        contract, total_functional.py:69-77).  They are evaluated by their
        own compute() on the DirectField and summed with the fused device
        terms."""
        return [(k, f) for k, f in self.funcDict.items() if f._device_terms() is None]

    def _merged_terms(self):
        t = dict(tf=False, vw=False, wt=False, lda=False, hartree=False, pseudo=False, x=1.0, y=1.0,
                 alpha=5.0 / 6.0, beta=5.0 / 6.0, rho0_fixed=None, pseudo_obj=None, gga=None, lwt=None)
        seen = set()
        for key, f in self.funcDict.items():
            d = f._device_terms()
            if d is None:
                continue  # a plugin term (plugin_terms)
            for flag in ("tf", "vw", "wt", "lda", "hartree", "pseudo"):
                if d.get(flag):
                    if flag in seen:
                        raise ValueError("term %s appears twice in the TotalFunctional" % flag)
                    seen.add(flag)
                    t[flag] = True
            for side in ("gga", "lwt"):
                if d.get(side) is not None:
                    if t["gga"] is not None or t["lwt"] is not None:
                        raise ValueError("two GGA / LWT kinetic terms in the TotalFunctional")
                    t[side] = d[side]
            if f.type == "KEDF":
                t.update(x=d["x"], y=d["y"], alpha=d["alpha"], beta=d["beta"], rho0_fixed=d["rho0"])
                if d["wt"] and abs(d["alpha"] - d["beta"]) > 1e-9:
                    raise AttributeError("the device WT path requires alpha == beta")
            if d.get("pseudo"):
                t["pseudo_obj"] = d["obj"]
        return t

    def has_device_terms(self):
        return any(f._device_terms() is not None for f in self.funcDict.values())

    def evaluator(self, grid, dtype=torch.float64, only=None):
        """The fused device evaluation of this sum's device terms on ``grid``
        (``only``: one term alone, by its switch name: tf, vw, wt, lda,
        hartree, pseudo, gga, lwt)."""
        # a restarted pseudopotential (new ions) needs a new evaluator
        ver = tuple(getattr(f, "_version", 0) for f in self.funcDict.values())
        key = (tuple(int(x) for x in grid.nrR), grid.lattice.tobytes(), bool(grid.full), dtype, ver, only)
        ev = self._evaluators.get(key)
        if ev is None:
            ev = DeviceEvaluator(self, grid, dtype, only=only)
            self._evaluators[key] = ev
        return ev

    def compute(self, rho, calcType=("E", "V"), **kwargs):
        return self.get_energy_potential(rho, calcType=calcType, **kwargs)

    def Energy(self, rho, **kwargs):
        return self.compute(rho, calcType=("E",)).energy

    def get_energy_potential(self, rho, calcType=("E", "V"), split=False, **kwargs):
        """Sum of every term (total_functional.py:115-146).  The device terms
        run fused (one evaluation for all of them); with ``split`` and a
        potential or energy density requested, each device term runs on its
        own so that its FunctionalOutput carries its own potential, as the
        reference returns them.  Plugin terms are called with the same
        calcType and keyword arguments.  calcType: 'E', 'V' and 'D' (the
        energy density of the kinetic terms: TF C_TF rho^(5/3), vW V_vW rho,
        WT V_NL rho / (2 alpha), kedf/tf.py:66-75, vw.py:125-126,
        wt.py:123-129)."""
        if not isinstance(rho, DirectField):
            raise TypeError("rho must be a DirectField")
        if rho.dim() != 3:
            raise NotImplementedError("spin-polarised densities are not on the device path")
        calc = set(calcType)
        plugins = self.plugin_terms()
        device = self.has_device_terms()
        if device:
            bad = calc - {"E", "V", "D"}
            if bad:
                raise AttributeError("calcType %s is not on the device path (E, V and D are)" % sorted(bad))
        parts = {}
        if device:
            parts = self._device_parts(rho, calc, split)
        for key, f in plugins:
            if split and f.type == "KEDF":
                res = f(rho, calcType=calcType, split=True, **kwargs)
                for k2, o in res.items():
                    k = "KEDF-" + k2.split("-")[-1]
                    parts[k] = _on_device(o, rho) + parts[k] if k in parts else _on_device(o, rho)
            else:
                o = _on_device(f(rho, calcType=calcType, **kwargs), rho)
                k = f.type if split else key
                parts[k] = o + parts[k] if k in parts else o
        ew = self.ewald
        total = FunctionalOutput(name="TOTAL", energy=0.0)
        for o in parts.values():
            total = total + o
        if ew:
            ii = FunctionalOutput(name="II", energy=ew.energy)
            total = total + ii
        total.name = "TOTAL"
        if not split:
            return total
        out = {"TOTAL": total}
        out.update(parts)
        if ew:
            out["II"] = ii
        return out

    def _device_parts(self, rho, calc, split):
        """FunctionalOutputs of the device terms: {"DEVICE": sum} when not
        split and no energy density is asked, else one per term keyed like the
        reference's split output."""
        grid = rho.grid
        ev = self.evaluator(grid, rho.dtype)
        per_term = ("V" in calc and split) or "D" in calc
        if not per_term:
            V = DirectField(grid=grid, dtype=rho.dtype) if "V" in calc else None
            with torch._C.DisableTorchFunctionSubclass():
                sums = ev.eval_rho(rho.contiguous(), V)
            energies = ev.energies(sums)
            if not split:
                o = FunctionalOutput(name="DEVICE", energy=float(sum(energies.values())))
                if V is not None:
                    o.potential = V
                return {"DEVICE": o}
            return {k: FunctionalOutput(name=k, energy=e) for k, e in energies.items()}
        t = ev.terms
        if "D" in calc and (t["gga"] is not None or t["lwt"] is not None):
            raise AttributeError("the energy density of the GGA / LWT kinetic terms is not on the device path")
        out = {}
        for key, idx, flag in _SPLIT_KEYS + [("KEDF-GGA", None, "gga"), ("KEDF-NL", None, "lwt")]:
            active = (t[flag] is not None) if flag in ("gga", "lwt") else bool(t[flag])
            if not active:
                continue
            one = self.evaluator(grid, rho.dtype, only=flag)
            V = DirectField(grid=grid, dtype=rho.dtype)
            with torch._C.DisableTorchFunctionSubclass():
                sums = one.eval_rho(rho.contiguous(), V)
            e = float(sum(one.energies(sums).values()))
            o = FunctionalOutput(name=key, energy=e)
            if "V" in calc:
                o.potential = V
            if "D" in calc and key.startswith("KEDF-"):
                with torch._C.DisableTorchFunctionSubclass():
                    r = rho.contiguous()
                    if flag == "tf":
                        # C_TF cbrt(rho^5), not scaled by x (kedf/tf.py:66-75)
                        ctf = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)
                        ed = ctf * r * torch.pow(r * r, 1.0 / 3.0)
                    elif flag == "vw":
                        ed = V * r
                    else:
                        ed = V * r / (2.0 * t["alpha"])
                o.energydensity = DirectField(grid=grid, data=ed, dtype=rho.dtype)
            out[key] = o
        if not split:
            tot = None
            for o in out.values():
                tot = o if tot is None else tot + o
            tot.name = "DEVICE"
            return {"DEVICE": tot}
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
            key = "KEDF-" + k.split("-")[-1] if k in ("TF", "VW", "WT-NL", "GGA") else k
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
