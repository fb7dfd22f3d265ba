"""FP64 CPU oracle for the orbital-free DFT density-optimization hot path.

TEST INFRASTRUCTURE ONLY.  This package is a numpy/scipy restatement of the
reference algorithm (DFTpy, ``src/dftpy``), written from the reference source
text.  It exists to check the HIP implementation in ``dftpy_amd``.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and there only as the checker or as the timed CPU
baseline -- never as the thing measured or shipped.  ``dftpy_amd`` never
imports it and has no CPU fallback.

Pinning status of every function is listed in ``oracle/ofdft_oracle.py``'s
module docstring.
"""
from .ofdft_oracle import *  # noqa: F401,F403
