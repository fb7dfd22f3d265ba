"""The batched TN inner loop (ofdft_batch_tn_direction: the inner CG loops
of many cells of different grid lengths at once, in its graph form -- per
length stage launches replayed as one hipGraph per CG iteration -- and its
persistent form -- one cooperative launch) against the one-cell device
loop (ofdft_tn_direction) on the same inputs, and the lockstep batch
solver (batch.solve_batch_persistent) against one-at-a-time solves, FP64
solves and the oracle's converged minima
(tests/golden/oracle_converged.json).

Bar: the same loop length and status as the one-cell loop, and the same
direction up to the summation order of the sums (FP64 1e-9, FP32 1e-4
relative); converged energies within 1e-6 eV/atom of FP64 and of the
oracle."""
import json
import os

import numpy as np
import pytest
import torch

import dftpy_amd as D
from dftpy_amd import batch as B
from dftpy_amd import workloads as W
from dftpy_amd.optimization import Optimization

pytestmark = pytest.mark.gpu


def _prepared(cell, prec):
    """An Optimization of one cell after its first evaluation (the state the
    TN inner loop starts from), working in `prec`."""
    grid, total = B.build_cell(cell)
    total.ewald.energy
    opt = Optimization(EnergyEvaluator=total, optimization_method="TN",
                       optimization_options={"precision": prec, "tn_graph": False})
    opt.verbose = False
    opt._begin(guess_rho=D.DirectField(grid=grid, data=cell["rho_ini"]))
    return opt


def _cells():
    cells = W.batch_cells(12)
    pick, seen = [], set()
    for c in cells:
        if c["nr"][0] not in seen and c["nr"][0] <= 90:
            seen.add(c["nr"][0])
            pick.append(c)
    return pick


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_batched_direction_matches_one_cell_loop(prec, graph):
    cells = _cells()
    print("case batched TN direction", prec, "graph" if graph else "persistent", [c["nr"][0] for c in cells],
          flush=True)
    opts = [_prepared(c, prec) for c in cells]
    works = [o._st.W for o in opts]
    mus = [o._st.mu for o in opts]
    # the one-cell device loop first (it reads phi, res, W0 and writes d)
    ref = []
    for o, w, mu in zip(opts, works, mus):
        n, stat, pap = w.P.tn_direction(w.phi, w.res, w.W0, w.tn()["d"], w.tn()["work"], 50, 4)
        ref.append((n, stat, pap, w.tn()["d"].double().cpu().numpy().copy()))
    bt = B.TnBatch([w.P for w in works], graph=graph)
    niter, status, pap = bt.direction(works, 50)
    tol = 1e-9 if prec == "fp64" else 1e-4
    for k, (w, r) in enumerate(zip(works, ref)):
        d = w.tn()["d"].double().cpu().numpy()
        rel = np.abs(d - r[3]).max() / np.abs(r[3]).max()
        print("cell %d (%d^3): loop %d/%d status %d/%d  max rel |d - d_one| %.2e" % (
            k, cells[k]["nr"][0], niter[k], r[0], status[k], r[1], rel))
        assert niter[k] == r[0] and status[k] == r[1]
        assert rel < tol
    # the first cell left out of a second call: it reports no iterations and
    # leaves its direction alone; the others give the same directions again
    d0 = works[0].tn()["d"].clone()
    n2, s2, p2 = bt.direction([None] + works[1:], 50)
    assert n2[0] == 0 and torch.equal(works[0].tn()["d"], d0)
    for k, (w, r) in enumerate(zip(works[1:], ref[1:]), start=1):
        d = w.tn()["d"].double().cpu().numpy()
        assert n2[k] == r[0] and s2[k] == r[1]
        assert np.abs(d - r[3]).max() / np.abs(r[3]).max() < tol


@pytest.mark.parametrize("graph", [True, False])
def test_lockstep_batch_matches_one_at_a_time_and_fp64(graph):
    cells = _cells()
    print("case lockstep batch", "graph" if graph else "persistent", [(len(c["symbols"]), c["nr"][0]) for c in cells],
          flush=True)
    res, wall = B.solve_batch_persistent(cells, econv=1e-8, graph=graph)
    for c, r in zip(cells, res):
        one = B.solve_cell(c, "mixed", econv=1e-8, options={"mixed_phase1": False})
        f64 = B.solve_cell(c, "fp64", econv=1e-8)
        err = abs(r["energy"] - f64["energy"]) * W.HARTREE / r["nat"]
        print(c["name"], r["energy"], one["energy"], f64["energy"], "%.2e eV/atom" % err, r["iterations"],
              one["iterations"])
        assert r["converged"] == 0
        assert abs(r["energy"] - one["energy"]) * W.HARTREE / r["nat"] < 1e-7
        assert err < 1e-6


@pytest.mark.parametrize("graph", [True, False])
def test_full_64_cell_batch_persistent(graph):
    with open(os.path.join(os.path.dirname(__file__), "golden", "oracle_converged.json")) as fh:
        gold = json.load(fh)
    cells = W.batch_cells(64)
    print("case lockstep full batch: 64 cells,", "graph" if graph else "persistent", flush=True)
    res, wall = B.solve_batch_persistent(cells, econv=1e-8, graph=graph)
    print("lockstep batch wall %.3f s (%.1f solves/s)" % (wall, len(cells) / wall), flush=True)
    anchored = 0
    worst = 0.0
    for c, r in zip(cells, res):
        assert r["converged"] == 0, c["name"]
        if c["name"] in gold:
            g = gold[c["name"]]["energy_Ha"]
            eo = abs(r["energy"] - g) * W.HARTREE / r["nat"]
            print("%s %d atoms %d^3: %.10f oracle %.10f (%.2e eV/atom)" % (c["name"], r["nat"], c["nr"][0],
                                                                        r["energy"], g, eo))
            assert eo < 1e-6
            anchored += 1
        else:
            f64 = B.solve_cell(c, "fp64", econv=1e-8)
            err = abs(r["energy"] - f64["energy"]) * W.HARTREE / r["nat"]
            worst = max(worst, err)
            assert err < 1e-6, (c["name"], err)
    print("largest error against FP64 %.2e eV/atom" % worst)
    assert anchored == 4
