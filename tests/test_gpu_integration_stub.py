"""The reference-side ctypes binding of the integration notes (the first
code block of their section 4), run as written against the built library.  Its
GPUTotalFunctional (WT + TF + vW + LDA + Hartree + local PP through
ofdft_plan_create / ofdft_set_functional / ofdft_set_vloc / ofdft_eval)
must give the oracle's energy and potential on the same density.

DFTpy itself is not importable here, so the three names the stub imports
from it (AbstractFunctional, FunctionalOutput, DirectField) are supplied by
this package's drop-in classes of the same interface, and the stub's
CDLL("libofdft_amd.so") is pointed at the in-tree library.  Tolerances: the
FP64 bar of the other parity tests (energies 1e-10 relative, potential
1e-9 Ha max-abs).
"""
import ctypes
import os
import re
import sys
import types

import numpy as np
import pytest
import torch

import oracle as O
import dftpy_amd as D
from dftpy_amd import _lib as L
from dftpy_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stub_source():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as fh:
        text = fh.read()
    sec = text[text.index("## 4."):text.index("## 5.")]
    return re.findall(r"```python\n(.*?)```", sec, re.S)[0]


def test_integration_stub_matches_oracle(monkeypatch):
    print("case integration-notes stub, al4_32", flush=True)
    names = {"dftpy": {}, "dftpy.functional": {},
             "dftpy.functional.abstract_functional": {"AbstractFunctional": D.AbstractFunctional},
             "dftpy.functional.functional_output": {"FunctionalOutput": D.FunctionalOutput},
             "dftpy.field": {"DirectField": D.DirectField}}
    for name, attrs in names.items():
        mod = types.ModuleType(name)
        for k, v in attrs.items():
            setattr(mod, k, v)
        monkeypatch.setitem(sys.modules, name, mod)
    real_cdll = ctypes.CDLL
    monkeypatch.setattr(ctypes, "CDLL",
                        lambda name, *a, **k: real_cdll(L.LIB_PATH if name == "libofdft_amd.so" else name, *a, **k))
    ns = {"__name__": "gpu_total"}
    exec(compile(_stub_source(), "integration-notes-stub", "exec"), ns)

    wl = W.build("al4_32")
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {"Al": O.empty_core_table(3.0, 1.12)}, PME=True)
    tot = O.TotalFunctional(og, kedf="WT", pseudo=opp)
    rho = np.ascontiguousarray(wl["rho_ini"], dtype=np.float64)

    f = ns["GPUTotalFunctional"](grid, opp.vreal, rho0=float(rho.mean()))
    out = f.compute(rho)
    E, V, parts = tot.compute(rho, split=True)
    e_ref = sum(v for k, v in parts.items() if k != "II")
    assert abs(out.energy - e_ref) <= 1e-10 * abs(e_ref), (out.energy, e_ref)
    with torch._C.DisableTorchFunctionSubclass():
        vg = out.potential.detach().cpu().numpy()
    dv = float(np.abs(vg - V).max())
    assert dv < 1e-9, dv
    print("stub E = %.12f Ha (oracle %.12f), max|dV| = %.2e" % (out.energy, e_ref, dv), flush=True)
