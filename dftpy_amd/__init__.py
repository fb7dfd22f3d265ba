"""MI355X (gfx950) orbital-free DFT density-optimization hot path with
DFTpy's Python API: TotalFunctional.compute(rho, calcType) and
Optimization(...).optimize_rho(guess_rho) on device-resident DirectFields.

The compute runs in hand-written HIP kernels (``libofdft_amd.so``, C ABI in
include/ofdft_amd.h).  There is no CPU fallback: the functional and
optimizer classes raise when the library or the GPU is missing.
"""
from .grid import DirectGrid, ReciprocalGrid
from .field import DirectField, ReciprocalField
from .ions import Ions
from .functional import (FunctionalOutput, AbstractFunctional, KEDF, XC, Hartree, TotalFunctional, Functional,
                         ZeroFunctional)
from .pseudo import LocalPseudo, Ewald, RadialPseudo, read_recpot, empty_core
from .optimization import Optimization
from .math_utils import ecut2nr, best_fft_size

__version__ = "0.1.0"
