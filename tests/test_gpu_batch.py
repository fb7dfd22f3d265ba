"""Batching: independent cells solved concurrently on several streams give
exactly the energies of the same solves run one at a time, and stay within
1e-6 eV/atom of FP64 solves."""
import numpy as np
import pytest

from dftpy_amd import batch as B
from dftpy_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_concurrent_batch_matches_one_at_a_time():
    cells = [c for c in W.batch_cells(12) if c["nr"][0] <= 64][:4]
    print("case batch of", [(len(c["symbols"]), c["nr"][0]) for c in cells], flush=True)
    res, wall = B.solve_batch(cells, "mixed", workers=3, econv=1e-8)
    for c, r in zip(cells, res):
        one = B.solve_cell(c, "mixed", econv=1e-8)
        f64 = B.solve_cell(c, "fp64", econv=1e-8)
        err = abs(r["energy"] - f64["energy"]) * W.HARTREE / r["nat"]
        print(c["name"], r["energy"], one["energy"], f64["energy"], "%.2e eV/atom" % err)
        assert r["converged"] == 0
        assert r["energy"] == one["energy"]
        assert err < 1e-6
