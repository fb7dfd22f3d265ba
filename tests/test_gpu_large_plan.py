"""A plan too large for 32-bit element offsets: FP64, 600 x 600 x 256, whose
3-field spectrum (3 x 600 x 600 x 144 complex, 2.49 GB) exceeds 2^31 bytes,
so every transform takes the runtime-schedule kernels with 64-bit addressing
-- also the rows of length 256, which on smaller plans run the compile-time
kernels.

The density is one plane wave along the cell diagonal,
rho = rho0 (1 + a cos u), u = 2 pi (i/600 + j/600 + k/256).  Every energy
and potential then follows from functions of u alone, evaluated on the
19200-point grid of u that the 3D grid samples (19200 = lcm(600, 256); each
of its points appears 4800 times on the 3D grid):

* Hartree (hartree.py:24-45): E = pi Omega rho0^2 a^2 / |G0|^2 and
  V = 4 pi rho0 a cos(u) / |G0|^2, closed forms;
* TF (tf.py:66-80): C_TF Omega <rho^(5/3)> and (5/3) C_TF rho^(2/3);
* vW with y = 1 (vw.py:115-131): (1/2) |G0|^2 Omega <phi'(u)^2> and
  -(1/2) |G0|^2 phi''(u) / phi(u), phi = sqrt(rho), with the derivatives
  of the 1D spectral expansion (the harmonics of phi fall below 1e-15
  before the axis-2 Nyquist index 128, so the 3D transforms see the same
  series without aliasing).

Bar: the FP64 one of the hot path (energies 1e-10 relative, potential 1e-9
Ha max-abs)."""
import numpy as np
import pytest
import torch

import dftpy_amd as D

pytestmark = pytest.mark.gpu

NR = (600, 600, 256)
L = (60.0, 60.0, 25.6)
RHO0, AMP = 0.02, 0.3
M = 19200  # lcm(600, 256)


def _u_index():
    """Index m of each 3D grid point on the 1D grid u_m = 2 pi m / M."""
    i = torch.arange(NR[0], dtype=torch.int64, device="cuda") * (M // NR[0])
    j = torch.arange(NR[1], dtype=torch.int64, device="cuda") * (M // NR[1])
    k = torch.arange(NR[2], dtype=torch.int64, device="cuda") * (M // NR[2])
    return (i[:, None, None] + j[None, :, None] + k[None, None, :]) % M


def _expected_1d():
    u = 2 * np.pi * np.arange(M) / M
    rho = RHO0 * (1 + AMP * np.cos(u))
    phi = np.sqrt(rho)
    q = np.fft.fftfreq(M, 1.0 / M)  # harmonic number
    ph = np.fft.fft(phi)
    # harmonics above 64 (below 1e-50 of the mean) are zeroed: their
    # rounding noise times q^2 (up to 9.2e7) would swamp the derivatives
    ph[np.abs(q) > 64] = 0.0
    d1 = np.fft.ifft(1j * q * ph).real
    d2 = np.fft.ifft(-(q ** 2) * ph).real
    return u, rho, phi, d1, d2


def test_plan_beyond_32_bit_offsets_matches_closed_forms():
    print("case large plan FP64", NR, flush=True)
    lattice = np.diag(L)
    omega = float(np.prod(L))
    g0sq = sum((2 * np.pi / l) ** 2 for l in L)
    grid = D.DirectGrid(lattice, nr=NR, full=False)
    u, rho1, phi1, d1, d2 = _expected_1d()
    idx = _u_index()
    rho = torch.tensor(rho1, dtype=torch.float64, device="cuda")[idx]
    assert rho.shape == NR
    tot = D.TotalFunctional(KEDF=D.KEDF("TFVW", x=1.0, y=1.0), HARTREE=D.Hartree())
    field = D.DirectField(grid=grid, data=rho)
    parts = tot.get_energy_potential(field, split=True)
    full = tot.compute(field)

    ctf = 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0)
    e_ref = {"HARTREE": np.pi * omega * RHO0 ** 2 * AMP ** 2 / g0sq,
             "KEDF-TF": ctf * omega * np.mean(rho1 ** (5.0 / 3.0)),
             "KEDF-VW": 0.5 * g0sq * omega * np.mean(d1 ** 2)}
    for k, e in e_ref.items():
        got = parts[k].energy
        print("%s %.15e closed form %.15e (rel %.1e)" % (k, got, e, abs(got - e) / abs(e)))
        assert abs(got - e) <= 1e-10 * abs(e), k
    v1 = (4 * np.pi * RHO0 * AMP * np.cos(u) / g0sq + 5.0 / 3.0 * ctf * rho1 ** (2.0 / 3.0)
          - 0.5 * g0sq * d2 / phi1)
    v_ref = torch.tensor(v1, dtype=torch.float64, device="cuda")[idx]
    with torch._C.DisableTorchFunctionSubclass():
        dv = float((full.potential - v_ref).abs().max())
    print("max |V - V_closed| = %.2e Ha" % dv)
    assert dv < 1e-9
    assert abs(full.energy - sum(e_ref.values())) <= 1e-10 * abs(sum(e_ref.values()))
