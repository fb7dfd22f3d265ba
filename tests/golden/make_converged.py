"""Writes tests/golden/oracle_converged.json: converged TN minima of the
FP64 oracle (numpy restatement of the reference's optimize_rho with its
forward-difference Hessian product) for the configurations whose GPU solves
are anchored to it:

* al864_192: fcc Al 6x6x6 (864 atoms), 192^3, WT + TF + vW + LDA + Hartree +
  empty-core PP (PME order 10), seeded uniform-plus-noise start;
* batch cells: the first cell of every grid size in the seeded 64-cell batch
  (workloads.batch_cells(64)): 48^3, 64^3, 90^3, 108^3.

For each: the total energy, the per-term energies at the converged density,
the chemical potential, the iteration log (E, Nd, Nls), the wall time of the
whole solve, the number of functional evaluations and the CPU it ran on.
The wall time is the CPU baseline of a complete reference-style solve
(single-threaded pocketfft, as the reference's threads=1; --fft-workers N
times the same solve with N-thread transforms).

Run in the build container (minutes per case at 192^3):
    python tests/golden/make_converged.py [--cases al864_192,batch] [--fft-workers N] [--out FILE]
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from dftpy_amd import workloads as W  # noqa: E402

ECONV = 1e-9


def cpu_name():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def solve(wl, econv=ECONV):
    og = O.Grid(wl["cell"], wl["nr"], full=False)
    oi = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    opp = O.LocalPseudo(og, oi, {s: O.empty_core_table(*W.EMPTY_CORE[s]) for s in wl["tables"]}, PME=True)
    otot = O.TotalFunctional(og, kedf=wl["kedf"], kedf_options=wl["kedf_options"], pseudo=opp,
                             ewald=O.Ewald(og, oi, PME=True))
    otot.ewald.energy
    opp.vreal
    opt = O.Optimization(lambda r, c, p: otot.compute(r, c, phi=p), og, "TN", {"econv": econv, "maxiter": 200})
    t0 = time.perf_counter()
    opt.optimize_rho(wl["rho_ini"])
    wall = time.perf_counter() - t0
    E, _, parts = otot.compute(opt.rho, ("E",), split=True)
    nat = len(wl["symbols"])
    return {
        "nat": nat, "nr": list(wl["nr"]), "nelec": float(wl["nelec"]), "econv_Ha": econv,
        "energy_Ha": float(opt.energy), "energy_at_rho_Ha": float(E),
        "terms_Ha": {k: float(v) for k, v in parts.items()},
        "mu_Ha": float(opt.mu), "converged": int(opt.converged),
        "iterations": len(opt.log) - 1, "evaluations": int(opt.nfeval),
        "log": [[int(l[0]), float(l[1]), int(l[4]), int(l[5])] for l in opt.log],
        "wall_s": wall, "tn_iterations_per_s": (len(opt.log) - 1) / wall,
        "fft_threads": O.fft_workers(), "cpu": cpu_name(), "cpu_count": os.cpu_count(),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", default="al864_192,batch")
    ap.add_argument("--fft-workers", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                   "oracle_converged.json"))
    args = ap.parse_args()
    O.set_fft_workers(args.fft_workers)
    res = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            res = json.load(f)
    cases = args.cases.split(",")
    key_sfx = "" if args.fft_workers == 1 else "@fft%d" % args.fft_workers
    todo = []
    if "al864_192" in cases:
        todo.append(("al864_192", W.build("al864_192")))
    if "batch" in cases:
        seen = set()
        for c in W.batch_cells(64):
            if c["nr"][0] not in seen:
                seen.add(c["nr"][0])
                todo.append((c["name"], c))
    for name, wl in todo:
        print("solving", name, wl["nr"], len(wl["symbols"]), "atoms", flush=True)
        r = solve(wl)
        print(json.dumps({k: v for k, v in r.items() if k != "log"}), flush=True)
        res[name + key_sfx] = r
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
