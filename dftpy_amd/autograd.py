"""torch.autograd view of the functional: energy, potential and forces as
differentiable functions of the density and of the ionic positions.

* ``energy(rho, total)`` returns E[rho] (the TotalFunctional energy, ion-ion
  constant included) as an FP64 scalar tensor.  Its gradient with respect
  to rho is V * dV (E is a grid sum times dV, V = dE/drho, as
  ``TotalFunctional.compute`` returns it, total_functional.py:107-146).
  That gradient is itself differentiable: its backward is the exact
  Hessian product of ``ofdft_potential_jvp``, so forward-over-backward
  (``torch.autograd.grad`` with ``create_graph=True``) gives H u without the
  forward difference of get_direction_TN (optimization.py:106-114).
* ``potential(rho, total)`` returns V[rho] with the same Hessian backward.
* ``pseudo_energy(rho, positions, pseudo)`` returns the local-PP energy
  (pseudo/__init__.py:219-234); its gradient with respect to the positions
  is minus the local-PP forces (the structure-factor adjoint,
  pseudo/__init__.py:260-270, 505-571), with respect to rho V_loc * dV.

The WT kernel is held at its cached rho0 (wt.py:64-72): no derivative of the
kernel with respect to the mean density, as in the reference.  Third
derivatives are not provided.  Everything runs on the GPU kernels; there is
no CPU path.
"""
import numpy as np
import torch

from .field import DirectField


def _plain(t):
    return t.as_subclass(torch.Tensor) if isinstance(t, DirectField) else t


def _grid_of(rho, grid):
    g = grid if grid is not None else getattr(rho, "grid", None)
    if g is None:
        raise TypeError("rho must be a DirectField or a grid must be given")
    return g


def _hessian_product(ev, phi, V, aux, u):
    """(dV/drho)[u] at rho = phi^2: p = u / (2 phi) turns the phi-space
    product of the kernels into the density-space one."""
    P = ev.P
    if ev.terms.get("gga") is not None or ev.terms.get("lwt") is not None:
        raise NotImplementedError("the Hessian product of the GGA / LWT kinetic terms is not on the device path")
    with torch._C.DisableTorchFunctionSubclass():
        u = u.to(phi.dtype).contiguous()
        p = torch.where(phi > 0, u / (2.0 * torch.where(phi > 0, phi, torch.ones_like(phi))), torch.zeros_like(u))
        out = torch.empty_like(phi)
        P.potential_jvp(phi, p.contiguous(), V, aux, out)
    return out


def _device_only(total):
    if total.plugin_terms():
        raise NotImplementedError("autograd covers the device terms only; %s are plugin terms"
                                  % [k for k, _ in total.plugin_terms()])


class _Potential(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, total, grid):
        _device_only(total)
        ev = total.evaluator(grid, rho.dtype)
        with torch._C.DisableTorchFunctionSubclass():
            r = rho.detach().contiguous()
            phi = torch.sqrt(torch.clamp(r, min=0.0)).contiguous()
            V = torch.empty_like(r)
            aux = torch.empty(2 * ev.P.N, dtype=r.dtype, device=r.device)
            ev.eval_phi(phi, None, 1.0, 0.0, V, aux)
        ctx.save_for_backward(phi, V, aux)
        ctx.ev = ev
        return V

    @staticmethod
    def backward(ctx, gV):
        phi, V, aux = ctx.saved_tensors
        # the Jacobian dV_i/drho_j = (1/dV) d2E/drho_i drho_j is symmetric: vjp = jvp
        return _hessian_product(ctx.ev, phi, V, aux, gV), None, None


class _Energy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, total, grid):
        _device_only(total)
        ev = total.evaluator(grid, rho.dtype)
        with torch._C.DisableTorchFunctionSubclass():
            sums = ev.eval_rho(rho.detach().contiguous(), None)
        e = float(sum(ev.energies(sums).values()))
        ew = total.ewald
        if ew:
            e += ew.energy
        ctx.save_for_backward(rho)
        ctx.total, ctx.grid = total, grid
        return torch.tensor(e, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, g):
        (rho,) = ctx.saved_tensors
        V = _Potential.apply(rho, ctx.total, ctx.grid)
        return g.to(V.dtype) * V * ctx.grid.dV, None, None


def energy(rho, total, grid=None):
    """Differentiable total energy E[rho] (Ha)."""
    g = _grid_of(rho, grid)
    return _Energy.apply(_plain(rho), total, g)


def potential(rho, total, grid=None):
    """Differentiable total potential V[rho] (Ha), a plain tensor on the grid."""
    g = _grid_of(rho, grid)
    return _Potential.apply(_plain(rho), total, g)


def _pseudo_at(pseudo, positions):
    """The LocalPseudo of ``pseudo``'s species and tables at ``positions``."""
    from .ions import Ions
    from .pseudo import LocalPseudo
    pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    if pos.shape == pseudo.ions.positions.shape and np.array_equal(pos, pseudo.ions.positions):
        return pseudo
    ions = Ions(pseudo.ions.symbols, pos, pseudo.ions.cell)
    return LocalPseudo(grid=pseudo.grid, ions=ions, PP_list=pseudo.pps, PME=pseudo.PME,
                       BsplineOrder=pseudo.BsplineOrder)


class _PseudoEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, positions, pseudo, grid):
        pp = _pseudo_at(pseudo, positions.detach().cpu().numpy())
        r = DirectField(grid=grid, data=rho.detach(), dtype=torch.float64)
        e = pp.compute(r, calcType=("E",)).energy
        ctx.pp, ctx.grid = pp, grid
        ctx.save_for_backward(rho)
        ctx.pos_meta = (positions.dtype, positions.device)
        return torch.tensor(e, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, g):
        (rho,) = ctx.saved_tensors
        pp, grid = ctx.pp, ctx.grid
        grad_rho = grad_pos = None
        if ctx.needs_input_grad[0]:
            with torch._C.DisableTorchFunctionSubclass():
                grad_rho = g.to(rho.dtype) * pp.vreal_device(rho.dtype) * grid.dV
        if ctx.needs_input_grad[1]:
            F = pp.forces(DirectField(grid=grid, data=rho.detach(), dtype=torch.float64))
            dtype, device = ctx.pos_meta
            grad_pos = -g.to(torch.float64).cpu() * torch.as_tensor(F)
            grad_pos = grad_pos.to(dtype=dtype, device=device)
        return grad_rho, grad_pos, None, None


def pseudo_energy(rho, positions, pseudo, grid=None):
    """Differentiable local-PP energy; d/dpositions = -forces."""
    g = _grid_of(rho, grid)
    if not torch.is_tensor(positions):
        positions = torch.as_tensor(np.asarray(positions, dtype=np.float64))
    return _PseudoEnergy.apply(_plain(rho), positions, pseudo, g)
