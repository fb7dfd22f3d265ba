"""DirectGrid / ReciprocalGrid with the reference's attribute names
(grid.py:26-57, :291-415, :467-735).  Lattice rows are lattice vectors in
bohr.  Reciprocal-space arrays (g, gg, q, invgg, mask) are device tensors
built on first use; the hot path never reads them (the kernels generate G on
the fly)."""
import numpy as np
import torch

from .math_utils import ecut2nr, spacing2ecut


class BaseGrid:
    def __init__(self, lattice, nr=None, origin=(0.0, 0.0, 0.0), full=True, ecut=None, device=None, **kwargs):
        self.lattice = np.asarray(lattice, dtype=np.float64).reshape(3, 3).copy()
        if nr is None:
            nr = ecut2nr(ecut=ecut, lattice=self.lattice, **kwargs)
        self._nrR = np.array(nr, dtype=np.int32)
        self._nnrR = int(np.prod(self._nrR))
        self.volume = abs(float(np.linalg.det(self.lattice)))
        self._dV = self.volume / self._nnrR
        self._full = bool(full)
        self._ecut = ecut
        self.origin = np.asarray(origin, dtype=np.float64)
        self.device = torch.device(device) if device is not None else None

    @property
    def nrR(self):
        return self._nrR

    @property
    def nr(self):
        return self._nrR

    @property
    def nnrR(self):
        return self._nnrR

    @property
    def nnr(self):
        return self._nnrR

    @property
    def dV(self):
        return self._dV

    @property
    def full(self):
        return self._full

    @property
    def nrG(self):
        g = self._nrR.copy()
        if not self._full:
            g[-1] = g[-1] // 2 + 1
        return g

    @property
    def nnrG(self):
        return int(np.prod(self.nrG))

    def lengths(self):
        return np.sqrt((self.lattice ** 2).sum(axis=1))

    @property
    def spacings(self):
        return self.lengths() / self._nrR

    def get_device(self):
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        return self.device

    def __eq__(self, other):
        return np.allclose(self.lattice, other.lattice) and np.allclose(self.nrR, other.nrR)

    def __hash__(self):
        return hash((tuple(np.round(self.lattice.ravel(), 12)), tuple(self.nrR)))


class DirectGrid(BaseGrid):
    """Real-space grid (grid.py:291-465).  ``full`` defaults to True as in
    the reference's DirectGrid."""

    def __init__(self, lattice, nr=None, origin=(0.0, 0.0, 0.0), full=True, **kwargs):
        super().__init__(lattice, nr=nr, origin=origin, full=full, **kwargs)
        self._rgrid = None

    def get_reciprocal(self):
        if self._rgrid is None:
            bg = (2.0 * np.pi * np.linalg.inv(self.lattice)).T
            self._rgrid = ReciprocalGrid(bg, nr=self.nrR, full=self.full, direct=self, device=self.device)
        return self._rgrid

    def guess_ecut(self):
        sp2 = self.lengths() / (self._nrR - 1)
        return spacing2ecut((0.5 * (self.spacings + sp2)).max())

    @property
    def ecut(self):
        return self._ecut if self._ecut is not None else self.guess_ecut()

    def get_direct(self):
        return self


class ReciprocalGrid(BaseGrid):
    """Reciprocal grid (grid.py:467-735); lattice rows are b_j = 2 pi inv(L)^T."""

    def __init__(self, lattice, nr=None, full=True, direct=None, **kwargs):
        super().__init__(lattice, nr=nr, full=full, **kwargs)
        self._direct = direct
        self._g = None

    def get_direct(self):
        return self._direct

    @property
    def ecut(self):
        return self._direct.ecut

    @property
    def g2max(self):
        return 2.0 * self.ecut

    @property
    def g(self):
        if self._g is None:
            dev = self._direct.get_device()
            ax = []
            for i in range(3):
                n = int(self.nrR[i])
                if i == 2 and not self.full:
                    f = np.fft.rfftfreq(n, d=1.0 / n)
                else:
                    f = np.fft.fftfreq(n, d=1.0 / n)
                ax.append(torch.tensor(f, dtype=torch.float64, device=dev))
            S = torch.meshgrid(*ax, indexing="ij")
            bg = torch.tensor(self.lattice, dtype=torch.float64, device=dev)
            self._g = torch.stack([S[0] * bg[0, k] + S[1] * bg[1, k] + S[2] * bg[2, k] for k in range(3)])
        return self._g

    @property
    def gg(self):
        g = self.g
        return (g * g).sum(0)

    @property
    def q(self):
        return torch.sqrt(self.gg)

    @property
    def invgg(self):
        gg = self.gg
        inv = torch.where(gg > 0, 1.0 / torch.where(gg > 0, gg, torch.ones_like(gg)), torch.zeros_like(gg))
        return inv
