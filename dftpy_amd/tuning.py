"""Autotuner of the column-transform tile widths.

The axis-1 pass and the axis-0 pencils (forward transform, reciprocal
multiplier, inverse transform) of the compile-time lengths run W columns per
workgroup, W in {8, 16, 32} (FP32) or {8, 16} (FP64): W sets the threads per
workgroup (W x N2 of the four-step split N = N1 x N2), the LDS per workgroup,
the width of each coalesced row segment and the number of workgroups.  The
best W depends on the length, the precision and the GPU, so it is measured:
``autotune`` times every candidate on a full evaluation (three convolution
fields, as in the solver) with the plan's per-class kernel timers and keeps
the fastest per kernel class.  ``python -m dftpy_amd.tuning`` measures the
compile-time lengths and writes the table (``tuning_gfx950.json``, committed
from a run on the GPU); ``Plan`` applies the table's choice for its lengths
at creation.  The bf16 spectra of the TN inner loop (``Plan.set_cg_spectra``)
have their own entries, measured by ``autotune_bf16`` on the device loop.
"""
import argparse
import json
import os

import numpy as np

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")
CANDIDATES = {"fp32": (8, 16, 32), "fp64": (8, 16)}
FIXED_LENGTHS = (32, 48, 64, 90, 96, 108, 128, 160, 192)
DEFAULT = 16

_table = None


def table():
    """The committed table: {precision: {length: {"axis": W, "pencil": W, ...}}}."""
    global _table
    if _table is None:
        _table = {}
        if os.path.exists(TABLE_PATH):
            with open(TABLE_PATH) as fh:
                _table = json.load(fh).get("tiles", {})
    return _table


def bf16_tiles_for(nr):
    """(tile_axis, tile_pencil) of the bf16 spectra of the TN inner loop
    (Plan.set_cg_spectra): the axis pass tuned on the axis-1 length, the
    pencils on the axis-0 length; 0 (the widest dividing n2/2) without an
    entry or where the tuned width does not divide n2/2."""
    t = table().get("bf16", {})
    nb = int(nr[2]) // 2
    ax = int(t.get(str(int(nr[1])), {}).get("axis", 0))
    pe = int(t.get(str(int(nr[0])), {}).get("pencil", 0))
    return (ax if ax and nb % ax == 0 else 0, pe if pe and nb % pe == 0 else 0)


def tiles_for(nr, precision):
    """(tile_axis, tile_pencil) for a plan: the axis pass is tuned on the
    axis-1 length, the pencils on the axis-0 length; lengths without an entry
    keep the default."""
    t = table().get(precision, {})
    ax = t.get(str(int(nr[1])), {}).get("axis", DEFAULT)
    pe = t.get(str(int(nr[0])), {}).get("pencil", DEFAULT)
    return int(ax), int(pe)


def autotune(n, precision, repeats=12):
    """Time every candidate tile width on evaluations of an n^3 grid (TF, vW,
    WT, LDA and Hartree: three convolution fields) and return
    {"axis": W, "pencil": W, "us": {W: [axis_us, pencil_us]}}."""
    import torch
    from .plan import Plan
    P = Plan((n, n, n), np.eye(3) * (0.25 * n), precision=precision)
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=False, rho0=0.02)
    rng = np.random.default_rng(0)
    rho = torch.tensor(0.02 * (1.0 + 0.1 * rng.random((n, n, n))), dtype=P.dtype, device=P.device)
    V = P.empty()
    us = {}
    for w in CANDIDATES[precision]:
        P.set_tiles(w, w)
        P.eval(rho, V)  # warm-up (first launch of this shape)
        P.set_profiling(True, classes=["axis_c2c", "pencil_conv"])
        for _ in range(repeats):
            P.eval(rho, V)
        prof = P.get_profile()
        P.set_profiling(False)
        us[w] = [prof[c][0] / prof[c][1] * 1e3 for c in ("axis_c2c", "pencil_conv")]
    best_axis = min(us, key=lambda w: us[w][0])
    best_pencil = min(us, key=lambda w: us[w][1])
    P.set_tiles(DEFAULT, DEFAULT)
    return {"axis": best_axis, "pencil": best_pencil, "us": {str(w): v for w, v in us.items()}}


def autotune_bf16(n, repeats=6):
    """Time every tile width of the bf16 spectra on the device TN loop of an
    n^3 FP32 plan (TF, vW, WT, LDA, Hartree: three convolution fields) and
    return {"axis": W, "pencil": W, "us": {W: [axis_us, pencil_us]}}."""
    import torch
    from .plan import Plan
    P = Plan((n, n, n), np.eye(3) * (0.25 * n), precision="fp32")
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=False, rho0=0.02)
    rng = np.random.default_rng(0)
    phi = torch.tensor(np.sqrt(0.02 * (1.0 + 0.1 * rng.random((n, n, n)))), dtype=P.dtype, device=P.device)
    V = P.empty()
    aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
    P.eval_phi(phi, None, 1.0, 0.0, V, aux)
    res, W0 = P.empty(), P.empty()
    P.hvp_prepare(phi, V, aux, 0.3, res, W0)
    d = P.zeros()
    work = torch.zeros((4, n, n, n), dtype=P.dtype, device=P.device)
    us = {}
    for w in CANDIDATES["fp32"]:
        if (n // 2) % w:
            continue
        P.set_cg_spectra("bf16", (w, w))
        P.tn_direction(phi, res, W0, d, work, 8, 8)  # warm-up (first launch of this shape)
        P.set_profiling(True, classes=["axis_c2c", "pencil_conv"])
        for _ in range(repeats):
            P.tn_direction(phi, res, W0, d, work, 8, 8)
        prof = P.get_profile()
        P.set_profiling(False)
        us[w] = [prof[c][0] / prof[c][1] * 1e3 for c in ("axis_c2c", "pencil_conv")]
    P.set_cg_spectra("plan")
    if not us:
        return None
    best_axis = min(us, key=lambda w: us[w][0])
    best_pencil = min(us, key=lambda w: us[w][1])
    return {"axis": best_axis, "pencil": best_pencil, "us": {str(w): v for w, v in us.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default=",".join(str(n) for n in FIXED_LENGTHS))
    ap.add_argument("--precisions", default="fp32,fp64,bf16")
    ap.add_argument("--out", default=TABLE_PATH)
    args = ap.parse_args()
    # the measured precisions replace their entries; the others are kept
    tiles = {k: dict(v) for k, v in table().items()} if os.path.exists(args.out) and args.out == TABLE_PATH else {}
    for prec in args.precisions.split(","):
        tiles.setdefault(prec, {})
        for n in (int(x) for x in args.sizes.split(",")):
            r = autotune_bf16(n) if prec == "bf16" else autotune(n, prec)
            print(prec, n, json.dumps(r), flush=True)
            if r is not None:
                tiles[prec][str(n)] = r
    import torch
    out = {"device": torch.cuda.get_device_name(0), "method": "dftpy_amd.tuning.autotune: mean kernel time over "
           "12 timed evaluations per candidate, three convolution fields; bf16 (autotune_bf16): mean kernel "
           "time over 6 device TN loops of 8 CG iterations per candidate", "tiles": tiles}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
