"""Grid-size helpers (integer host arithmetic): math_utils.py:271-298 and
mpi/mp_serial.py:17-83 of the reference."""
import numpy as np


def spacing2ecut(spacing):
    return np.pi ** 2 / (2 * spacing ** 2)


def ecut2spacing(ecut):
    return np.sqrt(np.pi ** 2 / ecut * 0.5)


def best_fft_size(N, max_prime=13, scale=0.99, even=True):
    """Smallest 2^a 3^b 5^c 7^d (x 11 or 13) size >= ceil(scale*N) - 1
    (mp_serial.py:17-83).  Every such size is a product of 2, 3, 5, 7, 11
    and 13, which the device transforms accept (up to the per-axis LDS
    bound of the plan: FP64 601 on axes 0 and 1)."""
    a = int(np.ceil(np.log2(N))) + 1
    b = int(np.ceil(np.log(N) / np.log(3))) + 1
    c = int(np.ceil(np.log(N) / np.log(5))) + 1
    d = int(np.ceil(np.log(N) / np.log(7))) + 1
    istart = 1 if even else 0
    mg = np.mgrid[istart:a, :b, :c, :d].reshape(4, -1)
    arr0 = 2 ** mg[0] * 3 ** mg[1] * 5 ** mg[2] * 7 ** mg[3]
    if N < 100:
        arr1 = arr0[np.logical_and(arr0 > N / (max_prime + 1), arr0 < 2 * N)]
    else:
        arr1 = arr0[np.logical_and(arr0 > N / (max_prime + 1), arr0 < 1.2 * N)]
    allv = list(arr1)
    for item in (11, 13):
        if item <= max_prime:
            allv.extend(arr1 * item)
    allv = np.asarray(allv)
    return int(np.min(allv[allv > np.ceil(scale * N) - 1]))


def ecut2nr(ecut=None, lattice=None, optfft=True, spacing=None, max_prime=13, **kwargs):
    """math_utils.py:282-298."""
    sp = np.ones(3) * (ecut2spacing(ecut) if spacing is None else spacing)
    metric = np.dot(lattice, lattice.T)
    nr = np.zeros(3, dtype=np.int32)
    for i in range(3):
        nr[i] = int(np.sqrt(metric[i, i]) / sp[i])
        nr[i] += nr[i] % 2
        if optfft:
            nr[i] = best_fft_size(nr[i], max_prime=max_prime)
    return nr
