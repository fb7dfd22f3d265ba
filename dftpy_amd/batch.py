"""Many independent cells solved concurrently on one GPU.

Each worker thread owns one HIP stream (torch.cuda.Stream, default
priority) and solves its share of the cells with the ordinary
``Optimization``; every plan a worker creates is bound to that worker's
stream, so the kernels of different cells overlap on the device while the
host-side control of one cell (line search, TN control) waits for its own
stream only.  The ctypes calls release the GIL.  Cells are handed out
largest first (cost ~ n^3 log n) to balance the workers.

This is the batched counterpart of looping ``optimize_rho`` over cells
(the reference has no batching; each cell is one DFTpy run).
"""
import math
import threading
import time

import torch

from .field import DirectField
from .functional import KEDF, XC, Hartree, TotalFunctional
from .grid import DirectGrid
from .ions import Ions
from .optimization import Optimization
from .pseudo import LocalPseudo, empty_core
from .workloads import EMPTY_CORE


def build_cell(cell, pme=True):
    """Device objects of one cell description (see workloads.batch_cells)."""
    grid = DirectGrid(cell["cell"], nr=cell["nr"], full=False)
    ions = Ions(cell["symbols"], cell["positions"], cell["cell"])
    pp = LocalPseudo(grid=grid, ions=ions, PP_list={s: empty_core(*EMPTY_CORE[s]) for s in cell["tables"]}, PME=pme)
    ke = KEDF(cell["kedf"], **cell.get("kedf_options", {}))
    total = TotalFunctional(KEDF=ke, XC=XC("LDA"), HARTREE=Hartree(), PSEUDO=pp)
    return grid, total


def solve_cell(cell, precision="mixed", econv=1e-8, options=None):
    """One TN solve of one cell on the current stream; returns a result dict."""
    t0 = time.perf_counter()
    grid, total = build_cell(cell)
    total.ewald.energy
    rho = DirectField(grid=grid, data=cell["rho_ini"])
    # launch mode for the TN loop: a graph capture on one worker's stream
    # would make the other workers' synchronous copies (ion set-up) fail
    opts = {"econv": econv, "precision": precision, "maxiter": 200, "tn_graph": False}
    if options:
        opts.update(options)
    opt = Optimization(EnergyEvaluator=total, optimization_method="TN", optimization_options=opts)
    opt.verbose = False
    t1 = time.perf_counter()
    opt.optimize_rho(guess_rho=rho)
    torch.cuda.current_stream().synchronize()
    t2 = time.perf_counter()
    return dict(name=cell.get("name"), energy=opt.energy, converged=opt.converged, nat=len(cell["symbols"]),
                nr=tuple(cell["nr"]), iterations=sum(1 for l in opt.log[1:] if l["prec"] != "switch->fp64"),
                setup_s=t1 - t0, solve_s=t2 - t1)


def solve_batch(cells, precision="mixed", workers=4, econv=1e-8, options=None):
    """Solve all cells with ``workers`` concurrent streams.  Returns the
    per-cell results (input order) and the wall time of the whole batch."""
    if workers < 1:
        raise ValueError("workers must be positive")
    cost = [c["nr"][0] * c["nr"][1] * c["nr"][2] * math.log(max(c["nr"][0], 2)) for c in cells]
    order = sorted(range(len(cells)), key=lambda i: -cost[i])
    results = [None] * len(cells)
    errors = []
    lock = threading.Lock()
    nxt = [0]
    device = torch.cuda.current_device()

    def work():
        try:
            torch.cuda.set_device(device)
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                while True:
                    with lock:
                        if nxt[0] >= len(order) or errors:
                            return
                        i = order[nxt[0]]
                        nxt[0] += 1
                    results[i] = solve_cell(cells[i], precision, econv, options)
        except BaseException as e:  # surfaced to the caller below
            with lock:
                errors.append(e)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    threads = [threading.Thread(target=work, daemon=True) for _ in range(min(workers, len(cells)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return results, wall
