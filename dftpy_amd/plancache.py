"""Device plans shared per (grid, precision): the analogue of the reference's
global FFT plan cache (math_utils.py:20-25, :94-155)."""
import torch

from .plan import Plan

_PLANS = {}


def plan_for(grid, dtype):
    prec = "fp64" if dtype in (torch.float64, torch.complex128) else "fp32"
    key = (tuple(int(x) for x in grid.nrR), grid.lattice.tobytes(), bool(grid.full), prec, str(grid.get_device()))
    P = _PLANS.get(key)
    if P is None:
        P = Plan(tuple(int(x) for x in grid.nrR), grid.lattice, precision=prec, full=grid.full,
                 device=grid.get_device())
        _PLANS[key] = P
    return P


def clear():
    _PLANS.clear()
