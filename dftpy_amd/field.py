"""DirectField / ReciprocalField: device tensors that carry their grid.

The reference's fields are numpy ndarray subclasses (field.py:11-192,
:727-804); here they are torch.Tensor subclasses resident on the GPU.
Arithmetic, slicing and in-place assignment behave as on tensors; results of
grid shape keep the grid.  ``fft`` / ``ifft`` call the HIP transforms.
"""
import numpy as np
import torch

from .grid import DirectGrid, ReciprocalGrid


def _find_grid(objs):
    for o in objs:
        if isinstance(o, (DirectField, ReciprocalField)):
            g = getattr(o, "_grid", None)
            if g is not None:
                return o.__class__, g
        elif isinstance(o, (list, tuple)):
            r = _find_grid(o)
            if r[1] is not None:
                return r
    return None, None


def _wrap(out, cls, grid):
    if isinstance(out, torch.Tensor) and grid is not None and out.dim() >= 3:
        target = tuple(grid.nrR) if cls is DirectField else tuple(grid.nrG)
        if tuple(out.shape[-3:]) == target:
            o = out.as_subclass(cls)
            o._grid = grid
            return o
    if isinstance(out, torch.Tensor) and type(out) is not torch.Tensor and out.dim() < 3:
        return out.as_subclass(torch.Tensor)
    if isinstance(out, (tuple, list)):
        return type(out)(_wrap(o, cls, grid) for o in out)
    return out


class _GridTensor(torch.Tensor):
    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        kwargs = kwargs or {}
        fcls, grid = _find_grid(list(args) + list(kwargs.values()))
        with torch._C.DisableTorchFunctionSubclass():
            out = func(*args, **kwargs)
        if fcls is None:
            return out
        return _wrap(out, fcls, grid)

    @property
    def grid(self):
        return self._grid

    @property
    def rank(self):
        return 1 if self.dim() == 3 else int(self.shape[0])

    def asnumpy(self):
        return self.detach().cpu().numpy()


class DirectField(_GridTensor):
    """Real-space field on a DirectGrid (field.py:141-192)."""

    @staticmethod
    def __new__(cls, grid, memo="", rank=1, data=None, griddata_3d=None, griddata_F=None, dtype=None, **kwargs):
        if hasattr(grid, "get_direct") and not isinstance(grid, DirectGrid):
            grid = grid.get_direct()
        if not isinstance(grid, DirectGrid):
            raise TypeError("the grid argument is not an instance of DirectGrid")
        dev = grid.get_device()
        shape = tuple(int(x) for x in grid.nrR) if rank == 1 else (rank,) + tuple(int(x) for x in grid.nrR)
        src = data if data is not None else griddata_3d
        if griddata_F is not None:
            src = np.ascontiguousarray(np.asarray(griddata_F).reshape(shape, order="F"))
        if dtype is None:
            dtype = src.dtype if isinstance(src, torch.Tensor) and src.is_floating_point() else torch.float64
        if src is None:
            t = torch.zeros(shape, dtype=dtype, device=dev)
        else:
            t = torch.as_tensor(src, dtype=dtype, device=dev).reshape(shape).contiguous()
        obj = t.as_subclass(cls)
        obj._grid = grid
        return obj

    def integral(self):
        """sum * dV (field.py:186-192), FP64 accumulation."""
        with torch._C.DisableTorchFunctionSubclass():
            if self.dim() == 3:
                return float(self.to(torch.float64).sum()) * self._grid.dV
            return (self.to(torch.float64).sum(dim=(1, 2, 3)) * self._grid.dV).cpu().numpy()

    @property
    def N(self):
        return self.integral()

    def amean(self):
        with torch._C.DisableTorchFunctionSubclass():
            return float(self.to(torch.float64).mean())

    def fft(self):
        """DirectField.fft (field.py:413-430): rfftn (full=False) or fftn
        (full=True) times dV, through the HIP R2C transform."""
        from .plancache import plan_for
        grid = self._grid
        P = plan_for(grid, self.dtype)
        with torch._C.DisableTorchFunctionSubclass():
            F = P.r2c(self.contiguous())
            F = F * grid.dV
            if grid.full:
                F = _hermitian_expand(F, tuple(int(x) for x in grid.nrR))
        out = F.as_subclass(ReciprocalField)
        out._grid = grid.get_reciprocal()
        return out


class ReciprocalField(_GridTensor):
    """Reciprocal-space field (field.py:727-804)."""

    @staticmethod
    def __new__(cls, grid, memo="", rank=1, data=None, griddata_3d=None, dtype=torch.complex128, **kwargs):
        if hasattr(grid, "get_reciprocal") and not isinstance(grid, ReciprocalGrid):
            grid = grid.get_reciprocal()
        dev = grid.get_direct().get_device()
        shape = tuple(int(x) for x in grid.nrG)
        src = data if data is not None else griddata_3d
        t = torch.zeros(shape, dtype=dtype, device=dev) if src is None else \
            torch.as_tensor(src, dtype=dtype, device=dev).reshape(shape).contiguous()
        obj = t.as_subclass(cls)
        obj._grid = grid
        return obj

    def ifft(self, check_real=False, force_real=True):
        """ReciprocalField.ifft (field.py:779-804): irfftn (full=False) or
        Re(ifftn) (full=True) divided by dV, through the HIP C2R transform."""
        from .plancache import plan_for
        rgrid = self._grid
        dgrid = rgrid.get_direct()
        nr = tuple(int(x) for x in dgrid.nrR)
        with torch._C.DisableTorchFunctionSubclass():
            X = self
            if dgrid.full:
                X = _hermitian_project(X, nr)
            rdtype = torch.float64 if X.dtype == torch.complex128 else torch.float32
            P = plan_for(dgrid, rdtype)
            f = P.c2r(X.contiguous()) / dgrid.dV
        out = f.as_subclass(DirectField)
        out._grid = dgrid
        return out


def _neg_index(n, device):
    return (-torch.arange(n, device=device)) % n


def _hermitian_expand(Fh, nr):
    """Full fftn spectrum from the rfftn half (X[-k] = conj X[k])."""
    n0, n1, n2 = nr
    nh = n2 // 2 + 1
    full = torch.empty((n0, n1, n2), dtype=Fh.dtype, device=Fh.device)
    full[:, :, :nh] = Fh
    if n2 > nh:
        k2 = torch.arange(nh, n2, device=Fh.device)
        src = Fh[_neg_index(n0, Fh.device)][:, _neg_index(n1, Fh.device)][:, :, (n2 - k2)]
        full[:, :, nh:] = torch.conj(src)
    return full


def _hermitian_project(X, nr):
    """Half spectrum of the Hermitian part (X[k] + conj X[-k])/2 of a full
    spectrum, so that irfftn of it equals Re(ifftn(X))."""
    n0, n1, n2 = nr
    nh = n2 // 2 + 1
    dev = X.device
    Xm = X[_neg_index(n0, dev)][:, _neg_index(n1, dev)][:, :, _neg_index(n2, dev)]
    return (0.5 * (X + torch.conj(Xm)))[:, :, :nh].contiguous()
