"""Host side of the LWT / LMGP nonlocal kinetic term (kedf/lwt.py): the
kernel tables in eta (built once, as the reference's ke_kernel_saved), the
kf range that only grows (guess_kf_bound), the kf node list, and the call
of the device term ``ofdft_lwt`` on a plan.

Options and defaults are those of LWT (lwt.py:506-527): interp 'linear'
(or 'newton'; 'hermite'), kerneltype 'WT' / 'MGP' / 'MGPA' / 'MGPG',
symmetrization, lumpfactor (MGP kinetic-electron term), alpha, beta,
etamax, maxpoints, kdd, ratio, nsp, delta, neta, order (1 or 3 here),
kfmin, kfmax, ldw."""
import ctypes

import numpy as np
import torch
from scipy.interpolate import splrep, splev

from . import _lib as L

DEFAULTS = dict(interp="linear", kerneltype="WT", symmetrization=None, lumpfactor=None, alpha=5.0 / 6.0,
                beta=5.0 / 6.0, etamax=50.0, maxpoints=1000, kdd=3, ratio=1.15, nsp=None, delta=None, neta=50000,
                order=3, kfmin=None, kfmax=None, ldw=None)

# large-eta series of the inverse Lindhard function in 1/eta^2 (kedf/kernel.py:80-120)
_LARGE = (-0.13714285714285712, -6.39999999999999875e-2, -3.77825602968460128e-2, -2.51824061652633074e-2,
          -1.80879839616166146e-2, -1.36715733124818332e-2, -1.07236045520990083e-2, -8.65192783339199453e-3,
          -7.1372762502456763e-3, -5.9945117538835746e-3, -5.10997527675418131e-3, -4.41060829979912465e-3,
          -3.84763737842981233e-3, -3.38745061493813488e-3, -3.00624946457977689e-3)


def lindhard(eta, lbda, mu):
    """Inverse Lindhard function for the host tables (LindhardFunction,
    kedf/kernel.py:14-136): the normal form, then the eta -> 0 and eta -> 1
    limits within 1e-10, then the large-eta series above 3.65."""
    eta = np.asarray(eta, dtype=np.float64)
    out = eta.copy()
    tol, tol2 = 1e-10, 1e-11
    sel = (eta > tol2) & (np.abs(eta - 1.0) > tol2)
    e = eta[sel]
    out[sel] = 1.0 / (0.5 + 0.25 * (1.0 - e * e) * np.log((1.0 + e) / np.abs(1.0 - e)) / e) - 3.0 * mu * e * e - lbda
    sel = eta < tol
    e = eta[sel]
    out[sel] = 1.0 - lbda + e * e * (1.0 / 3.0 - 3.0 * mu)
    sel = np.abs(eta - 1.0) < tol
    out[sel] = 2.0 - lbda - 3.0 * mu + 20.0 * (eta[sel] - 1.0)
    sel = eta > 3.65
    e = eta[sel]
    t = 1.0 / (e * e)
    acc = np.full_like(e, _LARGE[-1])
    for c in _LARGE[-2::-1]:
        acc = c + t * acc
    out[sel] = 3.0 * (1.0 - mu) * e * e - lbda - 0.6 + t * acc
    return out


def wt_table(eta, x=1.0):
    """WTKernelTable (kedf/kernel.py:344-353): (4/5) C_TF F(eta; x, 1)."""
    return lindhard(eta, x, 1.0) * (4.0 / 5.0 * 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0))


def mgp_table(eta, maxpoints=1000, symmetrization=None):
    """MGPKernelTable (kedf/kernel.py:267-321) on a 1-D eta table: the
    t-integral of F(eta / t^(1/3)) as the sum over t = i / maxpoints, F
    splined on a log grid."""
    dt = 1.0 / maxpoints
    coe = 4.0 / 5.0 * (0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)) * 5.0 / 6.0 * dt
    g = np.logspace(np.log10(eta[1] / 10), np.log10(eta[-1] + 2), num=eta.size)
    tck = splrep(g, lindhard(g, 1.0, 1.0))
    k = np.zeros_like(eta)
    for i in range(1, maxpoints + 1):
        t = i * dt
        k += splev(eta / np.cbrt(t), tck, ext=3) / (t ** (-2.0 / 3.0) if symmetrization == "Geometric"
                                                     else t ** (1.0 / 6.0))
    if symmetrization == "Arithmetic":
        return 0.5 * (k * coe + wt_table(eta))
    return k * (2.0 * coe if symmetrization == "Geometric" else coe)


_TABLES = {}


class Tables:
    """The kernel table and its derivative table -eta dK/deta on
    linspace(0, etamax, neta) as the device takes them (lwt.py:533-563):
    cubic splines (order 3) or plain arrays read at the nearest entry
    (order 1)."""

    def __init__(self, kerneltype, x, etamax, neta, order, maxpoints, symmetrization):
        eta = np.linspace(0, etamax, neta)
        if kerneltype == "WT":
            tab = wt_table(eta, x)
        elif kerneltype == "MGP":
            tab = mgp_table(eta, maxpoints, symmetrization)
        elif kerneltype == "MGPA":
            tab = mgp_table(eta, maxpoints, "Arithmetic")
        elif kerneltype == "MGPG":
            tab = mgp_table(eta, maxpoints, "Geometric")
        else:
            raise AttributeError("LWT kerneltype %s is not on the device path" % kerneltype)
        dx = etamax / (neta - 1)
        self.keep = []
        if order > 1:
            if order != 3:
                raise AttributeError("LWT order %d: the device evaluates cubic (order 3) splines" % order)
            K = splrep(eta, tab, k=3)
            D = splrep(eta, splev(eta, K, der=1) * (-1.0 * eta), k=3)
            self.tabs = [self._spline(K, dx, etamax), self._spline(D, dx, etamax)]
        else:
            deriv = splev(eta, splrep(eta, tab, k=3), der=1) * (-1.0 * eta)
            self.tabs = [self._array(tab, dx, etamax), self._array(deriv, dx, etamax)]

    def _spline(self, tck, dx, etamax):
        t = np.ascontiguousarray(tck[0], dtype=np.float64)
        c = np.ascontiguousarray(tck[1], dtype=np.float64)
        self.keep += [t, c]
        lim = float(splev(etamax, tck))
        dp = ctypes.POINTER(ctypes.c_double)
        return L.LwtTable(0, t.ctypes.data_as(dp), c.ctypes.data_as(dp), t.size, 0.0, dx, etamax, lim)

    def _array(self, v, dx, etamax):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.keep.append(v)
        dp = ctypes.POINTER(ctypes.c_double)
        return L.LwtTable(1, None, v.ctypes.data_as(dp), v.size, 0.0, dx, etamax, float(v[-1]))

    def upload(self, P):
        """Set the tables on plan P unless they are already there."""
        if getattr(P, "_lwt_tables", None) is not self:
            P.lwt_set_kernel(self.tabs[0], self.tabs[1])
            P._lwt_tables = self


def tables(kerneltype="WT", x=1.0, etamax=50.0, neta=50000, order=3, maxpoints=1000, symmetrization=None):
    key = (kerneltype, float(x), float(etamax), int(neta), int(order), int(maxpoints), symmetrization)
    if key not in _TABLES:
        _TABLES[key] = Tables(*key)
    return _TABLES[key]


def guess_kf_bound(kf_lo, kf_hi, kfmin=None, kfmax=None, kftol=1e-3, saved=None):
    """lwt.py:36-80 from the smallest and largest kf: the range of the
    node list, which only grows between calls (kfmin down by the factors
    of the table, kfmax up in 0.5 steps); kfmin at least kftol."""
    prev_min = saved.get("kfmin") if saved is not None else None
    prev_max = saved.get("kfmax") if saved is not None else None
    if kfmin is not None and kfmax is not None:
        return [kfmin, kfmax]
    if prev_min is None:
        prev_min = 10
        if kfmin is None:
            kfmin = kf_lo
    if prev_max is None:
        prev_max = 1.0
        if kfmax is None:
            kfmax = kf_hi
    if kfmin is None or kfmin > prev_min:
        kfmin = prev_min
    if kfmax is None or kfmax < prev_max:
        kfmax = prev_max
    if kfmin > kf_lo:
        steps = [1e-5, 1e-4, 5e-3, 1e-3, 5e-3, 1e-2, 5e-2, 0.1, 0.5, 1.0]
        r = kf_lo / kfmin
        for i in range(len(steps) - 1, 0, -1):
            if r > steps[i]:
                kfmin *= steps[i]
                break
    if kfmin < kftol:
        kfmin = kftol
    if kfmax < kf_hi:
        kfmax += (np.round((kf_hi - kfmax) / 0.5) + 1) * 0.5
    if saved is not None:
        saved["kfmin"] = kfmin
        saved["kfmax"] = kfmax
    return [kfmin, kfmax]


def kf_nodes(kfmin, kfmax, ratio=1.15, nsp=None, delta=None):
    """The kf node list (lwt.py:148-165): nsp uniform nodes, or uniform with
    spacing delta, or geometric with ``ratio``; widened by 1e-16 at both
    ends."""
    if nsp is not None:
        if kfmax - kfmin < 1e-3:
            kfl = np.array([kfmin, kfmax], dtype=np.float64)
        else:
            kfl = kfmin + (kfmax - kfmin) / (nsp - 1) * np.arange(nsp)
    elif delta is not None:
        kfl = np.linspace(kfmin, kfmax, int(np.ceil((kfmax - kfmin) / delta)) + 1)
    else:
        kfl = kfmin * ratio ** np.arange(int(np.ceil(np.log(kfmax / kfmin) / np.log(ratio))) + 1)
    kfl = np.array(kfl, dtype=np.float64)
    kfl[0] -= 1e-16
    kfl[-1] += 1e-16
    return kfl


def _kf(rho):
    return 2.0 * np.cbrt(3.0 * np.pi ** 2 * rho)


class LwtTerm:
    """One LWT / LMGP term: options, the saved kf range and kernel state of
    its KEDF (ke_kernel_saved), and its evaluation on a plan."""

    def __init__(self, x=1.0, **opts):
        o = dict(DEFAULTS)
        o.update({k: v for k, v in opts.items() if k in DEFAULTS and v is not None})
        if o["interp"] not in ("linear", "newton", "hermite"):
            raise AttributeError("LWT interp %s is not on the device path" % o["interp"])
        if int(o["kdd"]) not in (1, 2, 3):
            raise AttributeError("LWT kdd must be 1, 2 or 3")
        if int(o["order"]) not in (0, 1, 3):
            raise AttributeError("LWT order %s: the device evaluates order 3 splines or plain tables" % o["order"])
        self.o = o
        self.x = float(x)
        self.saved = {"kfmin": None, "kfmax": None, "shape": None, "rho0": 0.0, "mgpe": None}

    def _tables(self):
        o = self.o
        return tables(o["kerneltype"], self.x, o["etamax"], o["neta"], o["order"], o["maxpoints"],
                      o["symmetrization"])

    def _mgpe(self, ne):
        """MGPOmegaE parameters (kernel.py:408-440): a, b, c."""
        lf = self.o["lumpfactor"]
        c, b = 1.0, None
        if isinstance(lf, (list, tuple, np.ndarray)):
            a = lf[0]
            if len(lf) > 1:
                b = lf[1]
            if len(lf) > 2:
                c = lf[2]
        else:
            a = lf
        if b is None:
            a = float(a / ne ** (2.0 / 3.0))
            b = a
        return float(a), float(b), float(c)

    def evaluate(self, P, grid, rho=None, phi=None, p=None, c=1.0, s=0.0, V=None):
        """E, 2 sum V phi dphi and sum V rho (x dV) of the term at rho or
        (c phi + s p)^2; V += V_LWT when V is given (LWT, lwt.py:506-566 and
        LWTPotentialEnergy :83-341)."""
        o = self.o
        sv = self.saved
        shape = tuple(int(n) for n in grid.nrR)
        if shape != sv["shape"]:
            # the reference rebuilds its tables per grid shape, with rho0 of that call
            with torch._C.DisableTorchFunctionSubclass():
                if rho is not None:
                    rho0 = float(rho.double().mean())
                else:
                    ph = phi.double() * c + (p.double() * s if p is not None else 0.0)
                    rho0 = float((ph * ph).mean())
            sv["shape"], sv["rho0"] = shape, rho0
            sv["mgpe"] = self._mgpe(rho0 * grid.volume) if o["lumpfactor"] is not None else None
        self._tables().upload(P)
        lo, hi = P.lwt_range(rho=rho, phi=phi, p=p, c=c, s=s)
        kf_lo, kf_hi = _kf(lo), _kf(hi)
        if abs(kf_hi - kf_lo) < 1e-8:
            # uniform gas: no energy, no potential (lwt.py:117-126)
            return np.zeros(3)
        kfmin, kfmax = guess_kf_bound(kf_lo, kf_hi, o["kfmin"], o["kfmax"], saved=sv)
        kfl = kf_nodes(kfmin, kfmax, o["ratio"], o["nsp"], o["delta"])
        ldw = o["ldw"] if o["ldw"] is not None else 1.0 / 6.0
        ld = max(0.1, ldw)
        mg = sv["mgpe"]
        prm = L.LwtParams(float(o["alpha"]), float(o["beta"]), float(kfmax), float(ld), float(hi ** ld),
                          int(o["kdd"]), 1 if o["interp"] == "hermite" else 0, 1 if mg is not None else 0,
                          *(mg if mg is not None else (0.0, 0.0, 0.0)))
        return P.lwt(prm, kfl, rho=rho, phi=phi, p=p, c=c, s=s, V=V)
