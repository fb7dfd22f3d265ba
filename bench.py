#!/usr/bin/env python3
"""Benchmark: truncated-Newton density optimization of the 864-atom fcc Al
cell (6x6x6 conventional cells, sigma = 0.05 A seeded displacements) on a
192^3 grid with WT + TF + vW + LDA + Hartree + empty-core local PP, on one
MI355X.

A step is one complete TN solve in mixed precision, from the seeded
uniform-plus-noise density to convergence: FP32 outer iterations (FP32
fields and transforms, FP64 reductions) until |dE| < 1e-8 |E|, then FP64
outer iterations (FP64 energies, potential, line search, electron-count
constraint) with the TN inner CG kept in FP32, to econv.  ``value`` is TN outer
iterations per second over the timed solves.  The same solve in FP64 is run
after the timed region (one untimed warm-up, then one timed solve) for the
mixed/FP64 ratio and the energy error;
both energies are also compared with the oracle's converged minimum of the
workload (tests/golden/oracle_converged.json).  Also outside the timed
region of ``value``: the same mixed solves with the inner CG's spectra
stored as bf16 between the transform passes (``cg_bf16_spectra``: its own
rate, energy error and ratio to FP64; every transform still computes in
FP32), and with the CG iterations replayed from a hipGraph (``tn_graph``).

After that, outside the timed region: the batching comparison (64 uneven
independent cells, lockstep in the graph and persistent forms and
concurrent streams, each cell's error against its own FP64 solve and, where
committed, the oracle),
and the CPU baseline: the first TN outer iteration of the oracle on the same
workload with one FFT thread and with the usable cores, next to the full
oracle solve timed in the build container.

Prints one JSON line (rank 0 of one GPU).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the timed region times one launch in PROFILE_STRIDE of the dominant class;
# odd, so that classes whose launches alternate between two kernels (the
# forward and inverse axis passes) are sampled alike
PROFILE_STRIDE = 5

# HBM traffic per launch of a kernel class: a calibrated constant, not
# measured by this script.  It comes from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script on the named workload, scaled by what the
# same counters read for a plain device copy (FETCH x 2.00, WRITE x 1.00);
# the summary, the commands and the source revision they ran on are in
# profiles/axis_c2c_fp32_al864_192.md.
MEASURED_TRAFFIC = {("al864_192", "fp32", "axis_c2c"): 171.75e6}
TRAFFIC_SOURCE = ("calibrated constant: rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, copy-kernel scaled, "
                  "measured at commit 801a045 on the split-row axis pass (profiles/axis_c2c_fp32_al864_192.md)")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="al864_192")
    ap.add_argument("--econv", type=float, default=1e-8)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="FFT threads of the all-cores CPU baseline sample (default: the usable cores, at most 16)")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-phase1", dest="phase1", action="store_false",
                    help="mixed mode without the FP32 outer phase (FP64 outer loop from the start)")
    ap.add_argument("--batch", type=int, default=0, metavar="NCELLS",
                    help="batch mode only: solve NCELLS seeded independent Al/Mg cells concurrently, report solves/s")
    ap.add_argument("--batch-cells", type=int, default=64,
                    help="cells of the batching comparison added to the main line (0: skip it)")
    ap.add_argument("--workers", type=int, default=8, help="concurrent streams in batch mode")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="only run a plain device copy of MB megabytes 5 times (counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(mb):
    """A plain copy kernel of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch calibrate the counters against the
    bytes actually moved (mb * 1e6 read and written per copy)."""
    import torch
    n = mb * 1_000_000 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": 4 * n, "dispatches": 5}), flush=True)


def build_system(name, D, W):
    wl = W.build(name)
    grid = D.DirectGrid(wl["cell"], nr=wl["nr"], full=False)
    ions = D.Ions(wl["symbols"], wl["positions"], wl["cell"])
    pp = D.LocalPseudo(grid=grid, ions=ions, PP_list={s: D.empty_core(*W.EMPTY_CORE[s]) for s in wl["tables"]},
                       PME=True)
    ke = D.KEDF(wl["kedf"], **wl["kedf_options"])
    ev = D.TotalFunctional(KEDF=ke, XC=D.XC("LDA"), HARTREE=D.Hartree(), PSEUDO=pp)
    return wl, grid, ions, ev


def solve(D, ev, rho, precision, econv, verbose, phase1=False, hvp="exact", tn_graph=False, cg_spectra="plan"):
    """One TN solve from ``rho`` (a DirectField already resident on the GPU;
    optimize_rho does not modify its guess)."""
    opt = D.Optimization(EnergyEvaluator=ev, optimization_method="TN",
                         optimization_options={"econv": econv, "precision": precision, "maxiter": 200,
                                               "mixed_phase1": phase1, "hvp": hvp, "tn_graph": tn_graph,
                                               "cg_spectra": cg_spectra})
    opt.verbose = verbose
    opt.optimize_rho(guess_rho=rho)
    niter = sum(1 for l in opt.log[1:] if l["prec"] != "switch->fp64")
    return opt, niter


def host_cores():
    """(cores this process may run on, cores of the host)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def cpu_baseline(wl, threads):
    """The oracle (numpy FP64 restatement of the reference's optimize_rho, TN
    with its forward-difference Hessian product) timed on the host cores:
    the first TN outer iteration (initial evaluation, inner CG, line search)
    from the same start on the same workload, ``threads`` FFT worker threads
    (the pointwise numpy work is single-threaded either way).  Set-up (V_loc,
    Ewald, the WT kernel table) is outside the sample."""
    import oracle as O
    from dftpy_amd import workloads as W
    O.set_fft_workers(threads)
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {s: O.empty_core_table(*W.EMPTY_CORE[s]) for s in wl["tables"]}, PME=True)
    tot = O.TotalFunctional(og, kedf=wl["kedf"], kedf_options=wl["kedf_options"], pseudo=opp,
                            ewald=O.Ewald(og, oi, PME=True))
    tot.ewald.energy
    opp.vreal
    tot.compute(wl["rho_ini"])  # builds the kernel table
    opt = O.Optimization(lambda r, c, p: tot.compute(r, c, phi=p), og, "TN", {"econv": 1e-9, "maxiter": 2})
    t0 = time.perf_counter()
    opt.optimize_rho(wl["rho_ini"])
    dt = time.perf_counter() - t0
    it = len(opt.log) - 1
    O.set_fft_workers(1)
    return dict(value=it / dt, unit="TN iterations/s", cores=threads, kind="port",
                sample="first TN outer iteration of the oracle (%d functional evaluations: Nd=%d, Nls=%d) on the "
                       "same %s workload (%d^3, %s) from the same start, %d FFT thread(s): %.1f s"
                       % (opt.nfeval, opt.log[1][4], opt.log[1][5], wl["name"], wl["nr"][0], wl["kedf"], threads, dt))


def committed_oracle_solve(name):
    """The full oracle TN solve of the workload timed in the build container
    (tests/golden/oracle_converged.json, written by
    tests/golden/make_converged.py): energy, iterations, wall time, CPU."""
    path = os.path.join(ROOT, "tests", "golden", "oracle_converged.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        g = json.load(fh).get(name)
    if not g:
        return None
    return {"energy_Ha": g["energy_Ha"], "econv_Ha": g["econv_Ha"], "iterations": g["iterations"],
            "evaluations": g["evaluations"], "wall_s": g["wall_s"], "value": g["tn_iterations_per_s"],
            "unit": "TN iterations/s", "fft_threads": g["fft_threads"], "cpu": g["cpu"], "host_cores": g["cpu_count"]}


def batch_compare(ncells, econv, workers):
    """The batching configuration (``ncells`` independent cubic Al/Mg cells,
    32-500 atoms, grids 48^3-108^3, mixed precision) solved three ways, all
    lockstep forms running the inner CG loops of every cell in one batched
    call per TN outer iteration: ``graph`` (per-length stage launches, one
    captured hipGraph per CG iteration of all cells), ``persistent`` (one
    cooperative launch per TN outer iteration), and ``streams`` (worker
    threads, one HIP stream each, cells solved independently).  Every cell's
    energy is compared with an FP64 solve of that cell on its own, and the
    cells with a committed oracle minimum (one per grid size) with it."""
    import numpy as np
    from dftpy_amd import batch as B
    from dftpy_amd import workloads as W
    cells = W.batch_cells(ncells)
    seen = {}
    for c in cells:
        seen.setdefault(c["nr"], c)
    warm = list(seen.values())
    # warm-up: one cell of every grid size in each form (first launches, kernel tables)
    B.solve_batch_persistent(warm, econv=econv, graph=True)
    B.solve_batch_persistent(warm, econv=econv, graph=False)
    B.solve_batch(warm, "mixed", workers=len(warm), econv=econv)
    res_g, wall_g = B.solve_batch_persistent(cells, econv=econv, graph=True)
    res_p, wall_p = B.solve_batch_persistent(cells, econv=econv, graph=False)
    res_s, wall_s = B.solve_batch(cells, "mixed", workers=workers, econv=econv)
    t0 = time.perf_counter()
    ref = [B.solve_cell(c, "fp64", econv=econv) for c in cells]
    t64 = time.perf_counter() - t0
    gold = {}
    gpath = os.path.join(ROOT, "tests", "golden", "oracle_converged.json")
    if os.path.exists(gpath):
        with open(gpath) as fh:
            gold = json.load(fh)

    def errors(res):
        e64 = np.array([abs(a["energy"] - b["energy"]) * W.HARTREE / a["nat"] for a, b in zip(res, ref)])
        eo = [abs(a["energy"] - gold[c["name"]]["energy_Ha"]) * W.HARTREE / a["nat"]
              for a, c in zip(res, cells) if c["name"] in gold]
        return {"vs_fp64_max": float(e64.max()), "vs_fp64_mean": float(e64.mean()),
                "vs_oracle_max": float(max(eo)) if eo else None, "oracle_anchored_cells": len(eo),
                "all_converged": all(r["converged"] == 0 for r in res)}

    sizes = {}
    for c in cells:
        k = "%d^3" % c["nr"][0]
        sizes[k] = sizes.get(k, 0) + 1
    return {"metric": "independent-cell TN solves/sec, mixed precision", "cells": len(cells), "grids": sizes,
            "econv_Ha": econv, "data": "synthetic: workloads.batch_cells(%d, seed=11)" % len(cells),
            "graph": {"value": len(cells) / wall_g, "unit": "solves/s", "wall_s": wall_g,
                      "tn_iterations_total": int(sum(r["iterations"] for r in res_g)), "error_eV_per_atom":
                      errors(res_g)},
            "persistent": {"value": len(cells) / wall_p, "unit": "solves/s", "wall_s": wall_p,
                           "tn_iterations_total": int(sum(r["iterations"] for r in res_p)), "error_eV_per_atom":
                           errors(res_p)},
            "streams": {"value": len(cells) / wall_s, "unit": "solves/s", "wall_s": wall_s, "workers": workers,
                        "tn_iterations_total": int(sum(r["iterations"] for r in res_s)), "error_eV_per_atom":
                        errors(res_s)},
            "fp64_one_at_a_time": {"value": len(cells) / t64, "unit": "solves/s", "wall_s": t64}}


def batch_main(args):
    """Batching configuration only: prints its own JSON line."""
    out = batch_compare(args.batch, args.econv, args.workers)
    out.update({"value": out["graph"]["value"], "unit": "solves/s", "n_gpus": 1, "higher_is_better": True})
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.batch:
        batch_main(args)
        return
    import numpy as np
    import torch
    import dftpy_amd as D
    from dftpy_amd import workloads as W

    verbose = not args.quiet
    wl, grid, ions, ev = build_system(args.workload, D, W)
    nat = len(wl["symbols"])
    # set-up (V_loc, Ewald, plans) outside the timed region
    ev.ewald.energy
    ev32 = ev.evaluator(grid, torch.float32)
    ev64 = ev.evaluator(grid, torch.float64)
    plans = {"fp32": ev32.P, "fp64": ev64.P}
    # the initial density is copied to HBM once, before any timed region
    rho_ini = D.DirectField(grid=grid, data=wl["rho_ini"])
    # warm-up; the last warm-up solve is timed per kernel class (all classes)
    for w in range(max(args.warmup, 1)):
        if w == max(args.warmup, 1) - 1:
            for P in plans.values():
                P.set_profiling(True)
        solve(D, ev, rho_ini, "mixed", args.econv, False, args.phase1)
    breakdown = {k: P.get_profile() for k, P in plans.items()}
    # dominant kernel class: largest total device time
    prec_d, cls_d = max(((pk, ck) for pk, pr in breakdown.items() for ck, v in pr.items() if v[1]),
                        key=lambda t: breakdown[t[0]][t[1]][0])
    # timed region: events around the dominant class only (two per timed
    # launch), one launch in every PROFILE_STRIDE of that class: the event
    # pair adds dispatch overhead (about 10 us per timed launch), so timing
    # a uniform sample keeps most of it out of the timed solves
    for pk, P in plans.items():
        P.set_profiling(pk == prec_d, classes=[cls_d], stride=PROFILE_STRIDE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, energies, logs = 0, [], []
    for _ in range(args.steps):
        opt, n = solve(D, ev, rho_ini, "mixed", args.econv, verbose and not logs, args.phase1)
        iters += n
        energies.append(opt.energy)
        logs.append(opt.log)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_d, cnt_d, by_d = plans[prec_d].get_profile()[cls_d]
    for P in plans.values():
        P.set_profiling(False)
    value = iters / elapsed
    avg_s = ms_d / cnt_d / 1e3
    bytes_per_launch = by_d / cnt_d
    achieved = bytes_per_launch / avg_s / 1e9
    total_dev_ms = sum(v[0] for pr in breakdown.values() for v in pr.values())
    total_bytes = sum(v[2] for pr in breakdown.values() for v in pr.values())

    out = {
        "metric": "TN iterations/sec at 192^3 grid, mixed vs FP64 GPU, at 1e-6 eV/atom",
        "value": value,
        "unit": "TN iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: seeded fcc Al 6x6x6 supercell, sigma=0.05 A displacements, uniform+1% noise density",
        "config": {"workload": args.workload, "grid": list(wl["nr"]), "atoms": nat,
                   "kedf": "WT+TF+vW" if wl["kedf"] == "WT" else "TF+vW(y=%.4g)" % wl["kedf_options"]["y"],
                   "xc": "LDA(PZ)",
                   "pp": ", ".join("empty-core %s (Z=%g, rc=%g bohr)" % ((k,) + W.EMPTY_CORE[k]) for k in wl["tables"])
                         + ", PME order 10",
                   "optimizer": "TN, exact HVP", "econv_Ha": args.econv, "precision": ("mixed: FP32 outer then FP64 outer, FP32 inner CG" if args.phase1 else "mixed: FP64 outer, FP32 inner CG"),
                   "tn_iterations_per_solve": iters / args.steps},
        "roofline": {"bound": "hbm", "kernel": "%s (%s)" % (cls_d, prec_d), "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get((args.workload, prec_d, cls_d)),
                     "traffic_source": TRAFFIC_SOURCE if (args.workload, prec_d, cls_d) in MEASURED_TRAFFIC else None,
                     "launches": cnt_d, "sampled": "1 launch in %d of the class, events on the plan stream" % PROFILE_STRIDE,
                     "avg_us": avg_s * 1e6, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "all_kernels_GBps": total_bytes / (total_dev_ms / 1e3) / 1e9 if total_dev_ms else None},
        # algorithmic bytes of every kernel of one solve over the solve's wall time
        "whole_solve": {"algorithmic_bytes": total_bytes, "seconds": elapsed / args.steps,
                        "GBps": total_bytes / (elapsed / args.steps) / 1e9,
                        "frac": total_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS,
                        "device_busy_frac": total_dev_ms / 1e3 / (elapsed / args.steps)},
        # one warm-up solve, every class timed: [total ms, launches, us per launch, GB/s]
        "profile": {p: {k: [round(v[0], 3), v[1], round(v[0] / v[1] * 1e3, 1),
                            round(v[2] / v[1] / (v[0] / v[1] / 1e3) / 1e9, 0)] for k, v in pr.items() if v[1]}
                    for p, pr in breakdown.items()},
    }
    # the same solves with the device TN loop replaying one captured hipGraph
    # per CG iteration (ofdft_plan_set_graph) instead of its kernel launches
    solve(D, ev, rho_ini, "mixed", args.econv, False, args.phase1, tn_graph=True)  # capture outside the timing
    torch.cuda.synchronize()
    tg = time.perf_counter()
    giters, genergy = 0, None
    for _ in range(args.steps):
        og, ng = solve(D, ev, rho_ini, "mixed", args.econv, False, args.phase1, tn_graph=True)
        giters += ng
        genergy = og.energy
    torch.cuda.synchronize()
    tg = time.perf_counter() - tg
    out["tn_graph"] = {"value": giters / tg, "unit": "TN iterations/s", "ms_per_step": tg / args.steps * 1e3,
                       "same_energy": bool(genergy == energies[-1])}
    # the same solves with the inner CG's spectra stored as bf16 between the
    # transform passes (cg_spectra="bf16"; every transform, multiplier and
    # pointwise step still in FP32), reported beside the FP32-spectra value
    solve(D, ev, rho_ini, "mixed", args.econv, False, args.phase1, cg_spectra="bf16")  # first launches untimed
    bf_tiles = plans["fp32"].cg_spectra_tiles
    torch.cuda.synchronize()
    tb = time.perf_counter()
    biters, benergy = 0, None
    for _ in range(args.steps):
        ob, nb = solve(D, ev, rho_ini, "mixed", args.econv, False, args.phase1, cg_spectra="bf16")
        biters += nb
        benergy = ob.energy
    torch.cuda.synchronize()
    tb = time.perf_counter() - tb
    out["cg_bf16_spectra"] = {"value": biters / tb, "unit": "TN iterations/s", "ms_per_step": tb / args.steps * 1e3,
                              "tn_iterations_per_solve": biters / args.steps, "energy_Ha": benergy,
                              "converged": ob.converged, "tiles_axis_pencil": list(bf_tiles),
                              "storage": "bf16 spectra between the passes of the inner CG's Hessian products; "
                                         "FP32 transforms, FP64 sums, FP64 outer refinement"}
    gold = committed_oracle_solve(args.workload)
    if gold is not None:
        out["oracle_energy_Ha"] = gold["energy_Ha"]
        out["mixed_energy_error_vs_oracle_eV_per_atom"] = abs(energies[-1] - gold["energy_Ha"]) * W.HARTREE / nat
        out["cg_bf16_spectra"]["energy_error_vs_oracle_eV_per_atom"] = abs(benergy - gold["energy_Ha"]) * W.HARTREE / nat
    if not args.no_fp64:
        # untimed warm-up of the FP64 solves: their TN-loop kernels are not
        # used by the mixed solve, and first launches are not timed
        solve(D, ev, rho_ini, "fp64", args.econv, False)
        solve(D, ev, rho_ini, "fp64", args.econv, False, hvp="fd")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        o64, n64 = solve(D, ev, rho_ini, "fp64", args.econv, verbose)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t1
        fp64_rate = n64 / t64
        err = abs(energies[-1] - o64.energy) * W.HARTREE / nat
        out["fp64_gpu"] = {"value": fp64_rate, "unit": "TN iterations/s", "iterations": n64, "seconds": t64,
                           "energy_Ha": o64.energy, "converged": o64.converged}
        if gold is not None:
            out["fp64_gpu"]["energy_error_vs_oracle_eV_per_atom"] = abs(o64.energy - gold["energy_Ha"]) * W.HARTREE / nat
        out["mixed_over_fp64"] = value / fp64_rate
        out["cg_bf16_spectra"]["over_fp64"] = out["cg_bf16_spectra"]["value"] / fp64_rate
        out["cg_bf16_spectra"]["energy_error_vs_fp64_eV_per_atom"] = abs(benergy - o64.energy) * W.HARTREE / nat
        out["time_to_solution_ratio"] = t64 / (elapsed / args.steps)
        out["mixed_energy_error_eV_per_atom"] = err
        out["mixed_energy_Ha"] = energies[-1]
        # the reference's own TN on the FP64 kernels: the Hessian product as
        # a forward difference of the potential (optimization.py:106-114)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ofd, nfd = solve(D, ev, rho_ini, "fp64", args.econv, False, hvp="fd")
        torch.cuda.synchronize()
        tfd = time.perf_counter() - t1
        out["fp64_fd_gpu"] = {"value": nfd / tfd, "unit": "TN iterations/s", "iterations": nfd, "seconds": tfd,
                              "energy_Ha": ofd.energy, "converged": ofd.converged,
                              "hvp": "forward difference of V (the reference's TN), FP64 kernels"}
        out["mixed_over_fp64_fd"] = value / (nfd / tfd)
        out["time_to_solution_ratio_fd"] = tfd / (elapsed / args.steps)
    if args.batch_cells:
        out["batch"] = batch_compare(args.batch_cells, args.econv, args.workers)
    usable, total = host_cores()
    out["host_cores"] = {"usable": usable, "total": total}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, 1)
        nthr = args.cpu_threads or min(usable, 16)
        if nthr > 1:
            out["cpu_baseline_all_cores"] = cpu_baseline(wl, nthr)
        if gold is not None:
            out["cpu_full_solve_committed"] = gold
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
