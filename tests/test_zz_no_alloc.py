"""ofdft_ewald_noalloc (the allocation-free variant of ofdft_ewald, staging
and PME grid in a static per-GPU array) against ofdft_ewald, called through
the C ABI: the same real and reciprocal sums on two calls in a row, the
second with a smaller input (fewer atoms, smaller grid), with and without
PME; and a request larger than the static array rejected before any launch."""
import ctypes

import numpy as np
import pytest
from scipy import special

import dftpy_amd as D
from dftpy_amd import _lib as L
from dftpy_amd import workloads as W
from dftpy_amd.plan import Plan

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _args(name):
    wl = W.build(name)
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    ions.set_charges({s: W.EMPTY_CORE[s][0] for s in wl["tables"]})
    ew = D.Ewald(ions=ions, grid=grid, PME=True)
    rmax = special.erfcinv(ew.precision / 3.0) / np.sqrt(ew.eta)
    pos = np.ascontiguousarray(ions.positions, dtype=np.float64)
    q = np.ascontiguousarray(ions.charges, dtype=np.float64)
    return wl, grid, pos, q, float(ew.eta), float(rmax)


def _call(fn, P, pos, q, eta, rmax, pme):
    er, ek = ctypes.c_double(), ctypes.c_double()
    P._stream()
    rc = fn(P._h, len(q), pos.ctypes.data_as(_dp), q.ctypes.data_as(_dp), eta, rmax, pme, 10, ctypes.byref(er),
            ctypes.byref(ek))
    return rc, er.value, ek.value


def test_ewald_noalloc_matches_ewald():
    lib = L.load()
    for name, pme in (("al108_96", 1), ("al4_32", 1), ("al4_32", 0), ("al108_96", 0)):
        print("case ewald noalloc %s pme=%d" % (name, pme), flush=True)
        wl, grid, pos, q, eta, rmax = _args(name)
        P = Plan(tuple(wl["nr"]), wl["cell"], "fp64")
        rc0, er0, ek0 = _call(lib.ofdft_ewald, P, pos, q, eta, rmax, pme)
        L.check(rc0)
        for rep in range(2):
            rc1, er1, ek1 = _call(lib.ofdft_ewald_noalloc, P, pos, q, eta, rmax, pme)
            L.check(rc1)
            print("  real %.15e / %.15e  rec %.15e / %.15e" % (er0, er1, ek0, ek1))
            assert abs(er1 - er0) <= 1e-13 * abs(er0)
            assert abs(ek1 - ek0) <= 1e-13 * abs(ek0)


def test_ewald_noalloc_rejects_oversized_request():
    lib = L.load()
    wl, grid, pos, q, eta, rmax = _args("almg256_128")  # a 128^3 PME grid is 16 MiB
    P = Plan(tuple(wl["nr"]), wl["cell"], "fp64")
    rc, _, _ = _call(lib.ofdft_ewald_noalloc, P, pos, q, eta, rmax, 1)
    assert rc < 0
    msg = lib.ofdft_last_error().decode()
    print(msg)
    assert "scratch" in msg and "bytes" in msg
    # the exact sum of the same cell needs no grid and fits
    rc0, er0, ek0 = _call(lib.ofdft_ewald, P, pos, q, eta, rmax, 0)
    rc1, er1, ek1 = _call(lib.ofdft_ewald_noalloc, P, pos, q, eta, rmax, 0)
    L.check(rc0)
    L.check(rc1)
    assert abs(er1 - er0) <= 1e-13 * abs(er0) and abs(ek1 - ek0) <= 1e-13 * abs(ek0)
