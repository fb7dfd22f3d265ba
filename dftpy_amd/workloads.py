"""Seeded synthetic workloads: fcc Al / Al-Mg supercells with empty-core
local pseudopotentials and a uniform-plus-noise initial density.

Every array returned here is host (numpy) input data; the GPU path consumes
it through the functional classes.  Seeds are fixed so that the GPU run, the
oracle and the CPU baseline see identical inputs.

Workloads (name -> cells per edge m, grid n, KEDF, Mg fraction):
  al4_32     1,  32, TF+vW(1/9)+LDA+H+PP, uniform+noise, no displacement
  al108_96   3,  96, WT+LDA+H+PP, sigma=0.05 A displacements
  al864_192  6, 192, WT+LDA+H+PP, sigma=0.05 A displacements
  almg256_128 4, 128, WT+LDA+H+PP, 20% Mg, sigma=0.05 A displacements
"""
import numpy as np

BOHR = 0.5291772105638411  # Angstrom per bohr (ase.units.Bohr, CODATA 2014)
HARTREE = 27.21138602436724  # eV per hartree (ase.units.Ha)
A_AL = 4.05  # Angstrom

EMPTY_CORE = {"Al": (3.0, 1.12), "Mg": (2.0, 1.39)}  # Z, r_c (bohr)
FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])

WORKLOADS = {
    "al4_32": dict(m=1, n=32, kedf="TFVW", y=1.0 / 9.0, sigma=0.0, mg=0.0),
    "al108_96": dict(m=3, n=96, kedf="WT", y=1.0, sigma=0.05, mg=0.0),
    "al864_192": dict(m=6, n=192, kedf="WT", y=1.0, sigma=0.05, mg=0.0),
    "almg256_128": dict(m=4, n=128, kedf="WT", y=1.0, sigma=0.05, mg=0.2),
}


def empty_core_table(Z, rc, qmax=30.0, npts=2001):
    """Radial table of v(q) = -4 pi Z cos(q rc)/q^2 on linspace(0, qmax, npts)
    with the finite q=0 value 2 pi Z rc^2 (recpot convention: the Coulomb
    tail is kept, the q=0 entry holds the non-Coulomb limit)."""
    q = np.linspace(0.0, qmax, npts)
    v = np.empty_like(q)
    v[1:] = -4.0 * np.pi * Z * np.cos(q[1:] * rc) / q[1:] ** 2
    v[0] = 2.0 * np.pi * Z * rc ** 2
    return q, v


def fcc_supercell(m, a=A_AL, sigma=0.0, mg_fraction=0.0, seed=1):
    """(symbols, positions [bohr], cell [bohr]) of an m x m x m conventional
    fcc supercell; Gaussian displacements sigma (Angstrom) per Cartesian
    component from default_rng(seed); Mg on round(f*nat) sites chosen by
    default_rng(seed+1).choice."""
    L = a / BOHR
    cell = np.eye(3) * (m * L)
    grid = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], dtype=np.float64)
    frac = (grid[:, None, :] + FCC_BASIS[None, :, :]).reshape(-1, 3) / m
    pos = frac @ cell
    nat = len(pos)
    if sigma > 0:
        pos = pos + np.random.default_rng(seed).normal(0.0, sigma / BOHR, size=(nat, 3))
    symbols = ["Al"] * nat
    nmg = int(round(mg_fraction * nat))
    if nmg:
        for i in np.random.default_rng(seed + 1).choice(nat, nmg, replace=False):
            symbols[i] = "Mg"
    return symbols, pos, cell


def initial_density(nr, nelec, volume, noise=0.01, seed=0):
    """rho = (Ne/Omega)(1 + noise*xi), xi = default_rng(seed).standard_normal(nr),
    renormalised to integrate to Ne."""
    xi = np.random.default_rng(seed).standard_normal(tuple(nr))
    rho = (nelec / volume) * (1.0 + noise * xi)
    dV = volume / np.prod(nr)
    rho *= nelec / (rho.sum() * dV)
    return rho


def even_fft_size(n):
    """Smallest even 2^a 3^b 5^c >= n."""
    m = max(2, int(np.ceil(n)))
    while True:
        if m % 2 == 0:
            k = m
            for f in (2, 3, 5):
                while k % f == 0:
                    k //= f
            if k == 1:
                return m
        m += 1


def batch_cells(n=64, seed=11, spacing=0.2, sigma=0.05, mg_max=0.3):
    """Seeded batch of independent cubic fcc Al/Mg cells: atom counts drawn
    log-uniform in [32, 500] and snapped (in log space) to 4 m^3, m = 2..5;
    Mg fraction U[0, mg_max]; grid = smallest even 2^a 3^b 5^c >= L/spacing
    (spacing in Angstrom: 48, 64, 90, 108 points for m = 2..5 at 0.2 A);
    displacements sigma (Angstrom).  Cell i uses seeds seed*1000 + i."""
    rng = np.random.default_rng(seed)
    ms = np.array([2, 3, 4, 5])
    out = []
    for i in range(n):
        target = np.exp(rng.uniform(np.log(32.0), np.log(500.0)))
        m = int(ms[np.argmin(np.abs(np.log(4.0 * ms ** 3) - np.log(target)))])
        mg = float(rng.uniform(0.0, mg_max))
        symbols, pos, cell = fcc_supercell(m, sigma=sigma, mg_fraction=mg, seed=seed * 1000 + i)
        ng = even_fft_size(m * A_AL / spacing)
        nr = (ng,) * 3
        tables = {s: empty_core_table(*EMPTY_CORE[s]) for s in sorted(set(symbols))}
        zval = {s: EMPTY_CORE[s][0] for s in tables}
        nelec = sum(zval[s] for s in symbols)
        volume = abs(np.linalg.det(cell))
        out.append(dict(name="batch%02d" % i, symbols=symbols, positions=pos, cell=cell, nr=nr, kedf="WT",
                        kedf_options=dict(x=1.0, y=1.0), tables=tables, zval=zval, nelec=nelec,
                        rho_ini=initial_density(nr, nelec, volume, seed=seed * 1000 + i)))
    return out


def build(name):
    """Host description of a workload: dict with symbols, positions, cell,
    nr, kedf options, PP tables, zval, nelec, rho_ini."""
    w = WORKLOADS[name]
    symbols, pos, cell = fcc_supercell(w["m"], sigma=w["sigma"], mg_fraction=w["mg"])
    nr = (w["n"],) * 3
    tables = {s: empty_core_table(*EMPTY_CORE[s]) for s in sorted(set(symbols))}
    zval = {s: EMPTY_CORE[s][0] for s in tables}
    nelec = sum(zval[s] for s in symbols)
    volume = abs(np.linalg.det(cell))
    return dict(name=name, symbols=symbols, positions=pos, cell=cell, nr=nr, kedf=w["kedf"],
                kedf_options=dict(x=1.0, y=w["y"]), tables=tables, zval=zval, nelec=nelec,
                rho_ini=initial_density(nr, nelec, volume))
