"""Batching (independent Al/Mg cells, 32-500 atoms, 48^3-108^3 grids at
0.2 A, Mg fraction U[0, 0.3], sigma = 0.05 A, mixed precision): the full
seeded 64-cell batch solved concurrently gives exactly the energies of the
same solves run one at a time, every cell within 1e-6 eV/atom of its FP64
solve, and one cell of every grid size within 1e-6 eV/atom of the oracle's
converged minimum (tests/golden/oracle_converged.json)."""
import json
import os

import numpy as np
import pytest

from dftpy_amd import batch as B
from dftpy_amd import workloads as W

pytestmark = pytest.mark.gpu


def _golden():
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_converged.json")) as fh:
        return json.load(fh)


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


def test_full_64_cell_batch():
    cells = W.batch_cells(64)
    sizes = sorted({c["nr"][0] for c in cells})
    print("case full batch: 64 cells, grids", sizes, flush=True)
    assert sizes == [48, 64, 90, 108]
    gold = _golden()
    res, wall = B.solve_batch(cells, "mixed", workers=8, econv=1e-8)
    print("batch wall %.3f s (%.1f solves/s)" % (wall, len(cells) / wall), flush=True)
    worst = 0.0
    anchored = 0
    for c, r in zip(cells, res):
        assert r["converged"] == 0, c["name"]
        one = B.solve_cell(c, "mixed", econv=1e-8)
        f64 = B.solve_cell(c, "fp64", econv=1e-8)
        assert f64["converged"] == 0, c["name"]
        assert r["energy"] == one["energy"], c["name"]
        err = abs(r["energy"] - f64["energy"]) * W.HARTREE / r["nat"]
        worst = max(worst, err)
        assert err < 1e-6, (c["name"], err)
        if c["name"] in gold:
            g = gold[c["name"]]["energy_Ha"]
            eo = abs(r["energy"] - g) * W.HARTREE / r["nat"]
            e64 = abs(f64["energy"] - g) * W.HARTREE / r["nat"]
            print("%s %d atoms %d^3: mixed %.10f FP64 %.10f oracle %.10f Ha (%.2e / %.2e eV/atom)"
                  % (c["name"], r["nat"], c["nr"][0], r["energy"], f64["energy"], g, eo, e64))
            assert eo < 1e-6 and e64 < 1e-6
            anchored += 1
    print("largest mixed-vs-FP64 error %.2e eV/atom" % worst)
    assert anchored == 4
