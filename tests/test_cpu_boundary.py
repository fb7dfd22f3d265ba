"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every function the header declares, the Python constants agree with the
header, the product has no CPU fallback and never touches the oracle, and
the host-side helpers (workload generator, grid sizing) behave as the
reference's do.  No kernel is launched here."""
import ast
import ctypes
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ofdft_amd.h")
PKG = os.path.join(ROOT, "dftpy_amd")


def _defines():
    out = {}
    for m in re.finditer(r"#define\s+(OFDFT_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


def test_library_exports_every_header_function():
    from dftpy_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built (run python -m dftpy_amd.build_lib)")
    lib = ctypes.CDLL(L.LIB_PATH)
    declared = L.exported_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # every declared function has a ctypes signature and vice versa
    assert sorted(L._SIGS) == declared
    assert L.load().ofdft_version() > 0


def test_python_constants_match_header():
    from dftpy_amd import _lib as L
    d = _defines()
    assert (d["OFDFT_F64"], d["OFDFT_F32"]) == (L.F64, L.F32)
    assert d["OFDFT_NSUM"] == L.NSUM
    slots = ["OFDFT_E_TF", "OFDFT_E_VW", "OFDFT_E_NL", "OFDFT_E_XC", "OFDFT_E_H", "OFDFT_E_PP", "OFDFT_S_VRHO",
             "OFDFT_S_RHO", "OFDFT_S_DTHETA"]
    assert [d[s] for s in slots] == [L.E_TF, L.E_VW, L.E_NL, L.E_XC, L.E_H, L.E_PP, L.S_VRHO, L.S_RHO, L.S_DTHETA]
    assert [d[s] for s in ("OFDFT_CONV_NONE", "OFDFT_CONV_HARTREE", "OFDFT_CONV_GG", "OFDFT_CONV_WT")] == \
        [L.CONV_NONE, L.CONV_HARTREE, L.CONV_GG, L.CONV_WT]
    tn = ["OFDFT_TN_CONV", "OFDFT_TN_WARN", "OFDFT_TN_FAILED", "OFDFT_TN_NOTREDUCE", "OFDFT_TN_SMALL",
          "OFDFT_TN_MAXFUN"]
    assert [d[s] for s in tn] == [L.TN_CONV, L.TN_WARN, L.TN_FAILED, L.TN_NOTREDUCE, L.TN_SMALL, L.TN_MAXFUN]
    from dftpy_amd.plan import Plan
    assert d["OFDFT_NPROF"] == len(Plan.PROFILE_CLASSES)


def test_missing_library_fails_loudly(monkeypatch):
    from dftpy_amd import _lib as L
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libofdft_amd.so"))
    with pytest.raises(ImportError):
        L.load()


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the no-GPU error path")
def test_no_cpu_fallback_without_gpu():
    import dftpy_amd as D
    from dftpy_amd.plan import Plan
    with pytest.raises(RuntimeError):
        Plan((8, 8, 8), np.eye(3) * 5.0)
    grid = D.DirectGrid(np.eye(3) * 5.0, nr=(8, 8, 8), full=False)
    ev = D.TotalFunctional(KEDF=D.KEDF("TFVW"), XC=D.XC("LDA"), HARTREE=D.Hartree())
    with pytest.raises(RuntimeError):
        ev.compute(D.DirectField(grid=grid, data=np.full((8, 8, 8), 0.01)))


def test_product_never_imports_the_oracle():
    for dirpath, _, files in os.walk(PKG):
        for f in files:
            if not f.endswith(".py"):
                continue
            tree = ast.parse(open(os.path.join(dirpath, f)).read())
            for node in ast.walk(tree):
                if isinstance(node, ast.Import):
                    names = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom):
                    names = [node.module or ""]
                else:
                    continue
                assert not any(n.split(".")[0] == "oracle" for n in names), (f, names)


def test_workload_generator_is_seeded_and_sized():
    from dftpy_amd import workloads as W
    a = W.build("al4_32")
    b = W.build("al4_32")
    assert np.array_equal(a["rho_ini"], b["rho_ini"])
    assert a["nr"] == (32, 32, 32) and len(a["symbols"]) == 4
    # integral of the initial density is the electron count
    vol = abs(np.linalg.det(a["cell"]))
    assert abs(a["rho_ini"].sum() * vol / a["rho_ini"].size - 12.0) < 1e-10
    c = W.build("almg256_128")
    assert c["nr"] == (128, 128, 128) and len(c["symbols"]) == 256
    assert sum(1 for s in c["symbols"] if s == "Mg") == 51
    vol = abs(np.linalg.det(c["cell"]))
    assert abs(c["rho_ini"].sum() * vol / c["rho_ini"].size - 717.0) < 1e-8


def test_grid_sizing_matches_oracle():
    import oracle as O
    from dftpy_amd import math_utils as M
    for n in (17, 30, 47, 97, 190, 191, 250):
        assert M.best_fft_size(n) == O.best_fft_size(n)
    cell = np.eye(3) * 7.653390809640
    for ecut in (10.0, 22.9, 40.0):
        assert tuple(M.ecut2nr(ecut=ecut, lattice=cell)) == tuple(O.ecut2nr(cell, ecut=ecut))
    # known answers of the reference's tests/unit/test_math_utils.py:4-12
    assert M.ecut2spacing(1) == pytest.approx(2.221441469079183)
    assert M.ecut2spacing(2) == pytest.approx(1.5707963267948966)


def test_plan_size_limits_raise_before_any_device_call():
    """The per-axis LDS bound (160 KiB per workgroup): FP64 n0 <= 601,
    FP32 n0 <= 1203, checked on the host before anything touches a GPU."""
    import ctypes
    from dftpy_amd import _lib as L
    lib = L.load()
    lat = (ctypes.c_double * 9)(*[10.0, 0, 0, 0, 10.0, 0, 0, 0, 10.0])
    h = ctypes.c_void_p()
    for prec, n in ((L.F64, 640), (L.F32, 1250)):
        rc = lib.ofdft_plan_create(ctypes.byref(h), n, 32, 32, lat, prec, 0, None)
        assert rc != 0
        assert "LDS" in lib.ofdft_last_error().decode()
    # a length with a prime factor above 13 is refused on the host too
    rc = lib.ofdft_plan_create(ctypes.byref(h), 34, 32, 32, lat, L.F64, 0, None)
    assert rc != 0 and "products of 2, 3, 5, 7, 11 and 13" in lib.ofdft_last_error().decode()
