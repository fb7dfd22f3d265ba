"""Many independent cells solved concurrently on one GPU.

Two forms.  ``solve_batch_persistent`` runs the cells in lockstep: every TN
outer iteration runs the inner CG loops of all cells in ONE persistent
cooperative launch (``ofdft_batch_tn_direction``: each stage of a CG
iteration is a work queue over (cell, tile) items of every cell, stages
separated by grid barriers), then each cell's line search and update on
worker threads and streams.
``solve_batch`` hands the cells to worker threads instead: each worker thread owns one HIP stream (torch.cuda.Stream, default
priority) and solves its share of the cells with the ordinary
``Optimization``; every plan a worker creates is bound to that worker's
stream, so the kernels of different cells overlap on the device while the
host-side control of one cell (line search, TN control) waits for its own
stream only.  The ctypes calls release the GIL.  Cells are handed out
largest first (cost ~ n^3 log n) to balance the workers.

This is the batched counterpart of looping ``optimize_rho`` over cells
(the reference has no batching; each cell is one DFTpy run).
"""
import ctypes
import math
import threading
import time

import torch

from . import _lib as L

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


class TnBatch:
    """ofdft_batch over the inner-loop plans of several cells (one
    precision, one stream).  graph=True: the graph form (per-length stage
    launches, one captured hipGraph per CG iteration of all cells);
    graph=False: the persistent form (one cooperative launch)."""

    def __init__(self, plans, graph=True):
        self.lib = L.load()
        self.plans = list(plans)
        n = len(self.plans)
        arr = (ctypes.c_void_p * n)(*[P._h.value for P in self.plans])
        self._h = ctypes.c_void_p()
        L.check(self.lib.ofdft_batch_create(ctypes.byref(self._h), arr, n))
        L.check(self.lib.ofdft_batch_set_graph(self._h, 1 if graph else 0))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.ofdft_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def direction(self, works, maxfun):
        """get_direction_TN of every cell (work buffers of optimization._Work:
        phi, res, W0 in, tn()["d"] out) in one batched call; a None work
        leaves its cell out.  Returns (niter, status, pAp_last) lists."""
        n = len(works)
        if n != len(self.plans):
            raise ValueError("one work per plan of the batch")

        def arr(ts):
            return (ctypes.c_void_p * n)(*[t.data_ptr() if t is not None else None for t in ts])

        tn = [w.tn() if w is not None else {"d": None, "work": None} for w in works]
        works = [w if w is not None else _Absent for w in works]
        niter, status, pap = (ctypes.c_int * n)(), (ctypes.c_int * n)(), (ctypes.c_double * n)()
        L.check(self.lib.ofdft_batch_tn_direction(self._h, arr([w.phi for w in works]), arr([w.res for w in works]),
                                                  arr([w.W0 for w in works]), arr([t["d"] for t in tn]),
                                                  arr([t["work"] for t in tn]), int(maxfun), niter, status, pap))
        return list(niter), list(status), list(pap)


class _Absent:
    """Stands for a cell left out of a batched call (null pointers)."""
    phi = res = W0 = None


def solve_batch_persistent(cells, econv=1e-8, options=None, workers=8, graph=True):
    """Solve all cells in lockstep, mixed precision (FP64 outer iterations,
    FP32 inner CG with the exact Hessian product): per TN outer iteration,
    one batched call runs the inner loops of every unconverged cell (graph
    form by default, the single cooperative launch with graph=False), then
    each cell takes its line search and update (the reference's optimize_rho
    per cell, optimization.py:283-525).

    The per-cell outer work (line search, update, residual) runs on
    ``workers`` host threads, each with its own HIP stream: the outer plans
    of cell k are bound to stream k mod workers, the inner (FP32) plans of
    every cell to the calling stream, on which the batched launch runs.  The
    streams meet at a device synchronisation before each batched launch (the
    launch reads the outer fields) and at the launch's own completion.
    Returns the per-cell results (input order) and the wall time."""
    from concurrent.futures import ThreadPoolExecutor
    from .optimization import _finite
    if workers < 1:
        raise ValueError("workers must be positive")
    opts_common = {"econv": econv, "precision": "mixed", "mixed_phase1": False, "maxiter": 200, "hvp": "exact",
                   "tn_graph": False}
    if options:
        opts_common.update(options)
    if opts_common["mixed_phase1"]:
        raise ValueError("the lockstep batch runs FP64 outer iterations from the start (mixed_phase1=False)")
    device = torch.cuda.current_device()
    main = torch.cuda.current_stream()
    nw = max(1, min(workers, len(cells)))
    streams = [torch.cuda.Stream() for _ in range(nw)]
    groups = [list(range(g, len(cells), nw)) for g in range(nw)]
    built = [None] * len(cells)

    def on_workers(fn, ids=None):
        """fn(i) for every cell i (of ``ids``), cell i on its worker stream."""
        want = None if ids is None else set(ids)

        def run(g):
            torch.cuda.set_device(device)
            out = []
            with torch.cuda.stream(streams[g]):
                for i in groups[g]:
                    if want is None or i in want:
                        out.append((i, fn(i)))
            return out

        res = {}
        for part in pool.map(run, range(nw)):
            res.update(part)
        return res

    def build(i):
        c = cells[i]
        grid, total = build_cell(c)
        total.ewald.energy
        rho = DirectField(grid=grid, data=c["rho_ini"])
        opt = Optimization(EnergyEvaluator=total, optimization_method="TN", optimization_options=dict(opts_common))
        opt.verbose = False
        built[i] = (c, opt, rho)

    with ThreadPoolExecutor(nw) as pool:
        on_workers(build)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        on_workers(lambda i: built[i][1]._begin(guess_rho=built[i][2]))
        # the inner plans of every cell go on the batch (calling) stream
        for c, opt, rho in built:
            opt._st.Wi.P.bind_stream(True)
        maxiter = opts_common["maxiter"]
        maxfun = opts_common.get("maxfun", 50)
        active = list(range(len(built)))
        # one batch over every cell; converged cells are left out of later calls
        batch = TnBatch([c[1]._st.Wi.P for c in built], graph=graph)
        it = 0
        while active and it < maxiter - 1:
            it += 1
            torch.cuda.synchronize()  # the outer fields of every cell are final
            for i in active:
                st = built[i][1]._st
                built[i][1]._tn_inner_inputs(st.W, st.Wi)
            live = set(active)
            niter, status, pap = batch.direction(
                [c[1]._st.Wi if i in live else None for i, c in enumerate(built)], maxfun)
            nd = {}
            for i in active:
                _finite(pap[i], "p.Ap in the truncated-Newton loop")
                nd[i] = niter[i]

            def step(i):
                opt = built[i][1]
                st = opt._st
                st.W.P.convert_from(st.Wi.tn()["d"], st.W.p)
                opt.nhvp += nd[i]
                return opt._advance(st, it, nd[i])

            done = on_workers(step, active)
            active = [i for i in active if not done[i]]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        on_workers(lambda i: built[i][1]._end())
    torch.cuda.synchronize()
    results = []
    for c, opt, rho in built:
        results.append(dict(name=c.get("name"), energy=opt.energy, converged=opt.converged, nat=len(c["symbols"]),
                            nr=tuple(c["nr"]), iterations=len(opt.log) - 1))
    return results, wall


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
