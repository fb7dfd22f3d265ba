"""Parity of the HIP transform and evaluation kernels with numpy / the oracle.

Tolerances: FP64 transforms 1e-12 relative to max|F|; FP32 transforms 2e-6;
FP64 per-term energies 1e-10 relative, potentials 1e-9 Ha max-abs (the
FP64 bar of the hot path); FP32 energies 1e-5 relative.
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
from dftpy_amd import workloads as W
from tests.conftest import DATA

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (20, 30, 20), (36, 36, 72), (48, 40, 30), (96, 96, 96), (24, 50, 18), (15, 9, 20), (20, 18, 15),
          (160, 24, 192), (108, 90, 48), (90, 64, 108),
          # factors 7, 11, 13 (grids of the reference's max_prime = 13 sizing)
          (20, 30, 22), (14, 22, 26), (7, 11, 13), (143, 12, 10), (28, 44, 52), (12, 10, 143),
          # column workgroups above 64 KiB of LDS (FP64 n0, n1 > 240): up to the 160 KiB bound
          (256, 250, 16), (24, 600, 18)]


def _plan(nr, lattice, prec="fp64", full=False):
    from dftpy_amd.plan import Plan
    return Plan(nr, lattice, precision=prec, full=full)


@pytest.mark.parametrize("nr", SHAPES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_r2c_c2r_match_numpy(nr, prec):
    print("case r2c/c2r", nr, prec, flush=True)
    P = _plan(nr, np.eye(3) * 10.0, prec)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(nr)
    ft = torch.tensor(f, dtype=P.dtype, device=P.device)
    F = P.r2c(ft).cpu().numpy()
    ref = np.fft.rfftn(f)
    tol = 1e-12 if prec == "fp64" else 2e-6
    assert np.abs(F - ref).max() <= tol * np.abs(ref).max()
    g = P.c2r(torch.tensor(ref, dtype=P.cdtype, device=P.device)).cpu().numpy()
    assert np.abs(g - f).max() <= tol * np.abs(f).max() * 10
    # c2r must not modify its input
    Ft = torch.tensor(ref, dtype=P.cdtype, device=P.device)
    F0 = Ft.clone()
    P.c2r(Ft)
    assert torch.equal(Ft, F0)


def _al4(full=False, kedf="TFVW", y=1.0 / 9.0, shear=None):
    wl = W.build("al4_32")
    if shear is not None:
        # the same fractional positions in a sheared cell
        frac = np.linalg.solve(wl["cell"].T, wl["positions"].T).T
        wl["cell"] = wl["cell"] @ (np.eye(3) + shear)
        wl["positions"] = frac @ wl["cell"]
    grid = O.Grid(wl["cell"], wl["nr"], full=full)
    ions = O.Ions(wl["symbols"], wl["positions"], wl["cell"])
    tabs = {k: O.empty_core_table(*W.EMPTY_CORE[k]) for k in wl["tables"]}
    pp = O.LocalPseudo(grid, ions, tabs, PME=True)
    return wl, grid, ions, pp


@pytest.mark.parametrize("kind", [1, 2])
def test_convolve_matches_oracle(kind):
    wl, grid, ions, pp = _al4()
    P = _plan(wl["nr"], wl["cell"])
    rho = wl["rho_ini"]
    out = P.convolve(kind, torch.tensor(rho, device=P.device)).cpu().numpy()
    mult = grid.invgg * 4 * np.pi if kind == 1 else grid.gg
    ref = np.fft.irfftn(np.fft.rfftn(rho) * mult, s=wl["nr"], axes=(0, 1, 2))
    assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()


def _check_eval(P, grid, tot, rho, tol_e=1e-10, tol_v=1e-9):
    E, V, parts = tot.compute(rho, split=True)
    Vg = torch.empty(P.nr, dtype=P.dtype, device=P.device)
    sums = P.eval(torch.tensor(rho, dtype=P.dtype, device=P.device), Vg)
    names = {"KEDF-TF": 0, "KEDF-VW": 1, "KEDF-NL": 2, "XC": 3, "HARTREE": 4, "PSEUDO": 5}
    for k, e in parts.items():
        if k == "II":
            continue
        got = sums[names[k]]
        print(k, got, e, flush=True)
        assert abs(got - e) <= tol_e * max(abs(e), 1e-3), (k, got, e)
    dv = np.abs(Vg.cpu().numpy() - V).max()
    print("max|dV|", dv)
    assert dv <= tol_v


def test_eval_tfvw_al4_fp64():
    wl, grid, ions, pp = _al4()
    tot = O.TotalFunctional(grid, kedf="TFVW", kedf_options=dict(x=1.0, y=1.0 / 9.0), pseudo=pp)
    P = _plan(wl["nr"], wl["cell"])
    P.set_functional(tf=True, vw=True, lda=True, hartree=True, pseudo=True, x=1.0, y=1.0 / 9.0)
    P.set_vloc(torch.tensor(pp.vreal, device=P.device))
    _check_eval(P, grid, tot, wl["rho_ini"])


def test_eval_wt_al4_fp64():
    wl, grid, ions, pp = _al4()
    rho = wl["rho_ini"]
    tot = O.TotalFunctional(grid, kedf="WT", pseudo=pp)
    P = _plan(wl["nr"], wl["cell"])
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=True, rho0=rho.mean())
    P.set_vloc(torch.tensor(pp.vreal, device=P.device))
    _check_eval(P, grid, tot, rho)


def test_eval_wt_fixture_full_grid():
    """Al_fde_rho.pp: non-orthogonal cell, full=True; KEDF WT known answer
    2.916818700014412 (test_functional.py:27)."""
    ions, rho, grid = O.read_qepp(os.path.join(DATA, "Al_fde_rho.pp"), full=True)
    tot = O.TotalFunctional(grid, kedf="WT", xc=True, hartree=True)
    P = _plan(tuple(grid.nr), grid.lattice, full=True)
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, rho0=rho.mean())
    _check_eval(P, grid, tot, rho)
    sums = P.eval(torch.tensor(rho, device=P.device))
    assert abs(sums[0] + sums[1] + sums[2] - 2.916818700014412) < 1e-10


def test_eval_fp32_al4():
    wl, grid, ions, pp = _al4()
    rho = wl["rho_ini"]
    tot = O.TotalFunctional(grid, kedf="WT", pseudo=pp)
    E, V, parts = tot.compute(rho, split=True)
    P = _plan(wl["nr"], wl["cell"], "fp32")
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=True, rho0=rho.mean())
    P.set_vloc(torch.tensor(pp.vreal, dtype=torch.float32, device=P.device))
    Vg = torch.empty(P.nr, dtype=P.dtype, device=P.device)
    sums = P.eval(torch.tensor(rho, dtype=torch.float32, device=P.device), Vg)
    for i, k in enumerate(["KEDF-TF", "KEDF-VW", "KEDF-NL", "XC", "HARTREE", "PSEUDO"]):
        assert abs(sums[i] - parts[k]) <= 1e-5 * max(abs(parts[k]), 1e-2), (k, sums[i], parts[k])
    assert np.abs(Vg.cpu().numpy() - V).max() <= 1e-4


def test_hvp_matches_central_difference():
    """Exact HVP vs central difference of (V(phi+hp) - mu)(phi+hp) (FP64)."""
    wl, grid, ions, pp = _al4()
    rho = wl["rho_ini"]
    tot = O.TotalFunctional(grid, kedf="WT", pseudo=pp)
    P = _plan(wl["nr"], wl["cell"])
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=True, rho0=rho.mean())
    P.set_vloc(torch.tensor(pp.vreal, device=P.device))
    phi = np.sqrt(rho)
    p = np.random.default_rng(5).standard_normal(rho.shape) * phi * 0.1
    mu = 0.3
    t = lambda a: torch.tensor(a, device=P.device)
    V = P.empty()
    aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
    P.eval_phi(t(phi), None, 1.0, 0.0, V, aux)
    Ap = P.empty()
    pAp = P.hvp(t(phi), t(p), V, aux, mu, Ap)
    h = 1e-5
    def F(ph):
        _, Vv = tot.compute(ph * ph, ("V",))
        return (Vv - mu) * ph
    fd = (F(phi + h * p) - F(phi - h * p)) / (2 * h)
    A = Ap.cpu().numpy()
    err = np.abs(A - fd).max() / np.abs(fd).max()
    print("hvp rel err", err)
    assert err < 1e-7
    assert abs(pAp - np.sum(p * A)) < 1e-9 * abs(pAp)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_hvp_fast_equals_hvp(prec):
    """hvp_prepare + hvp_fast (precomputed local part) == ofdft_hvp."""
    wl, grid, ions, pp = _al4()
    rho = wl["rho_ini"]
    P = _plan(wl["nr"], wl["cell"], prec)
    P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=True, rho0=rho.mean())
    P.set_vloc(torch.tensor(pp.vreal, dtype=P.dtype, device=P.device))
    t = lambda a: torch.tensor(a, dtype=P.dtype, device=P.device)
    phi = np.sqrt(rho)
    p = np.random.default_rng(5).standard_normal(rho.shape) * phi * 0.1
    V = P.empty()
    aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
    P.eval_phi(t(phi), None, 1.0, 0.0, V, aux)
    mu = 0.31
    A1, A2, res, W0 = P.empty(), P.empty(), P.empty(), P.empty()
    q1 = P.hvp(t(phi), t(p), V, aux, mu, A1)
    r2 = P.hvp_prepare(t(phi), V, aux, mu, res, W0)
    q2 = P.hvp_fast(t(phi), t(p), W0, A2)
    tol = 1e-12 if prec == "fp64" else 2e-5
    a1, a2 = A1.cpu().double().numpy(), A2.cpu().double().numpy()
    assert np.abs(a1 - a2).max() <= tol * np.abs(a1).max()
    assert abs(q1 - q2) <= tol * abs(q1) * 10
    ref = ((V.cpu().double().numpy() - mu) * phi)
    assert np.abs(res.cpu().double().numpy() - ref).max() <= tol * np.abs(ref).max() * 10
    assert abs(r2 - np.sum(ref ** 2)) <= 1e-4 * r2


def test_hvp_fast_fp32_matches_fp64_with_signed_phi():
    """The FP32 product (sqrt(rho) = |phi| cancelled in the functors) against
    the FP64 one on a phi with negative and zero entries, as a line search
    can produce: both are the reference's formula on rho = phi^2."""
    wl, grid, ions, pp = _al4()
    rho = wl["rho_ini"]
    rng = np.random.default_rng(8)
    phi = np.sqrt(rho) * np.where(rng.random(rho.shape) < 0.1, -1.0, 1.0)
    phi.flat[rng.integers(0, phi.size, 20)] = 0.0
    p = rng.standard_normal(rho.shape) * np.abs(phi).mean() * 0.1
    out = {}
    for prec in ("fp64", "fp32"):
        P = _plan(wl["nr"], wl["cell"], prec)
        P.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, pseudo=True, rho0=rho.mean())
        P.set_vloc(torch.tensor(pp.vreal, dtype=P.dtype, device=P.device))
        t = lambda a: torch.tensor(a, dtype=P.dtype, device=P.device)
        V = P.empty()
        aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
        P.eval_phi(t(np.abs(phi)), None, 1.0, 0.0, V, aux)
        res, W0, A = P.empty(), P.empty(), P.empty()
        P.hvp_prepare(t(phi), V, aux, 0.31, res, W0)
        q = P.hvp_fast(t(phi), t(p), W0, A)
        A1 = P.empty()
        P.hvp(t(phi), t(p), V, aux, 0.31, A1)
        out[prec] = (A.cpu().double().numpy(), q, A1.cpu().double().numpy())
    a64, q64, h64 = out["fp64"]
    a32, q32, h32 = out["fp32"]
    assert np.abs(a64 - h64).max() <= 1e-12 * np.abs(a64).max()
    err = np.abs(a32 - a64).max() / np.abs(a64).max()
    print("FP32 hvp_fast relative error %.3e" % err)
    assert err < 2e-5
    assert abs(q32 - q64) <= 2e-5 * abs(q64)


def _tn_host(hvp, res0, maxfun):
    """get_direction_TN (optimization.py:92-148) on the host, step for step,
    with Ap from the given product; `Best` aliases `direction` until the
    first copy, as in the reference."""
    direction = np.zeros_like(res0)
    res = -res0.copy()
    p = res.copy()
    r0 = np.sum(res * res)
    rconv = r0 * 0.1
    best, alias = None, True
    rl = [1e10]
    stat, pAp = "NOTCONV", 0.0
    for it in range(maxfun):
        Ap = hvp(p)
        pAp = np.sum(p * Ap)
        if pAp < 0.0:
            if it == 0:
                direction = r0 / pAp * p
                stat = "WARN"
            else:
                stat = "FAILED"
            break
        alpha = r0 / pAp
        direction = direction + alpha * p
        res = res - alpha * Ap
        r1 = np.sum(res * res)
        if r1 < min(rl):
            best, alias = direction.copy(), False
        rl.append(r1)
        if r1 < rconv:
            stat = "CONV"
            break
        elif r1 > 1000 * min(rl[:-1]):
            stat = "WARN : Not reduce"
            direction = direction if alias else best
            break
        elif it > 10 and abs(r0 - r1) < 0.1 * r0:
            stat = "WARN : Change too small"
            direction = direction if alias else best
            break
        beta = r1 / r0
        r0 = r1
        p = res + beta * p
    return direction, it + 1, stat, pAp


def _tn_case(mu, maxfun, chunk, local=False, full=False, shear=None, prec="fp64"):
    """Device TN loop against the host restatement on the 4-atom cell; with
    `local` only TF + LDA + PP (no convolution fields, so no forward row pass
    carries the CG update)."""
    from dftpy_amd import _lib as L
    print("case tn_direction", mu, maxfun, chunk, local, full, shear is not None, flush=True)
    wl, grid, ions, pp = _al4(full=full, shear=shear)
    rho = wl["rho_ini"]
    P = _plan(wl["nr"], wl["cell"], prec, full=full)
    P.set_functional(tf=True, vw=not local, wt=not local, lda=True, hartree=not local, pseudo=True, rho0=rho.mean())
    P.set_vloc(torch.tensor(pp.vreal, dtype=P.dtype, device=P.device))
    t = lambda a: torch.tensor(a, dtype=P.dtype, device=P.device)
    phi = t(np.sqrt(rho))
    V = P.empty()
    aux = torch.empty(2 * P.N, dtype=P.dtype, device=P.device)
    P.eval_phi(phi, None, 1.0, 0.0, V, aux)
    res, W0 = P.empty(), P.empty()
    P.hvp_prepare(phi, V, aux, mu, res, W0)
    Ap = P.empty()

    def hvp(pv):
        P.hvp_fast(phi, t(pv), W0, Ap)
        return Ap.cpu().numpy()

    dref, nref, sref, pref = _tn_host(hvp, res.cpu().numpy(), maxfun)
    d = P.zeros()
    work = torch.full((4,) + tuple(P.nr), 7.0, dtype=P.dtype, device=P.device)  # stale scratch must not matter
    n, st, pAp = P.tn_direction(phi, res, W0, d, work, maxfun, chunk)
    assert (n, L.TN_STATUS[st]) == (nref, sref)
    if mu > 1.0:
        assert sref == "WARN" and n == 1
    # FP32: the device's p.Ap (spectral sums of the FP32 transforms plus the
    # carried pointwise part) and the host's sum over the FP32 product differ
    # by rounding
    tol_p, tol_d = (1e-10, 1e-9) if prec == "fp64" else (1e-5, 1e-4)
    assert abs(pAp - pref) <= tol_p * abs(pref)
    dd = d.cpu().numpy()
    assert np.abs(dd - dref).max() <= tol_d * np.abs(dref).max()
    # a second call on the same plan starts from a clean state
    n2, st2, _ = P.tn_direction(phi, res, W0, d, work, maxfun, chunk)
    assert (n2, st2) == (n, st)
    assert np.abs(d.cpu().numpy() - dd).max() <= 1e-12 * np.abs(dd).max()
    return n, L.TN_STATUS[st]


# (mu, maxfun, branch the loop ends in); mu inside the spectrum of the
# Hessian makes the CG system indefinite and reaches the other exits
TN_CASES = [(0.31, 50, "CONV"), (0.31, 2, "NOTCONV"), (0.31, 1, "NOTCONV"), (40.0, 50, "WARN"),
            (0.0, 50, "WARN : Change too small"), (0.9, 50, "WARN : Not reduce"), (0.75, 50, "FAILED")]


@pytest.mark.parametrize("mu,maxfun,branch", TN_CASES)
@pytest.mark.parametrize("chunk", [1, 3, 16])
def test_tn_direction_matches_host_loop(mu, maxfun, branch, chunk):
    """The device-resident CG loop takes the branches of the reference loop
    (convergence, maxfun, negative curvature at the first and at a later
    iteration, `Best` restored after a stall or a growing residual) and
    returns its direction and iteration count."""
    n, status = _tn_case(mu, maxfun, chunk)
    assert status == branch


@pytest.mark.parametrize("full", [False, True])
def test_tn_direction_sheared_cell(full):
    """p.Ap of the device loop comes from the pencil's spectral sums with the
    Hermitian weights (and, with full=1 frequencies, the Hermitian part of
    the multipliers); on a sheared cell |G(k)| and |G(-k)| differ on the
    Nyquist planes, and the loop still follows the host loop."""
    shear = np.array([[0.0, 0.04, -0.03], [0.02, 0.0, 0.05], [-0.01, 0.03, 0.0]])
    n, status = _tn_case(0.31, 50, 3, full=full, shear=shear)
    assert status == "CONV"


def test_tn_direction_fp32():
    """The FP32 loop (the mixed-precision inner CG) against the host loop
    over the same FP32 products."""
    n, status = _tn_case(0.31, 50, 3, prec="fp32")
    assert status == "CONV"


@pytest.mark.parametrize("mu,branch", [(-0.3, "CONV"), (0.0, "FAILED"), (0.15, "WARN : Not reduce"), (40.0, "WARN")])
def test_tn_direction_local_functional(mu, branch):
    n, status = _tn_case(mu, 50, 3, local=True)
    assert status == branch


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_orth_sums_match_numpy(prec):
    """The five FP64-accumulated sums behind OrthogonalNormalization and the
    theta = 0 derivative (optimization.py:205-226, :370-380)."""
    nr = (20, 18, 15)
    P = _plan(nr, np.eye(3) * 7.0, prec)
    rng = np.random.default_rng(11)
    p, phi, V = (rng.standard_normal(nr) for _ in range(3))
    t = lambda a: torch.tensor(a, dtype=P.dtype, device=P.device)
    got = P.orth_sums(t(p), t(phi), t(V))
    # the sums of the values the kernel actually sees (rounded to the plan precision)
    p, phi, V = (t(a).double().cpu().numpy() for a in (p, phi, V))
    ref = (np.sum(p * phi), np.sum(phi * phi), np.sum(p * p), np.sum(V * phi * p), np.sum(V * phi * phi))
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-12 * max(abs(r), 1.0)


@pytest.mark.parametrize("n", [48, 96, 192])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_column_tile_widths_give_identical_transforms(n, prec):
    """Every tile width the autotuner may pick (8, 16, 32 columns per
    workgroup) computes each column with the same instructions: the
    transforms and a full evaluation agree bit for bit, and match numpy."""
    print("case tiles", n, prec, flush=True)
    nr = (n, n // 2, n // 4 * 2)
    P = _plan(nr, np.eye(3) * 10.0, prec)
    rng = np.random.default_rng(5)
    f = rng.standard_normal(nr)
    ft = torch.tensor(f, dtype=P.dtype, device=P.device)
    ref = np.fft.rfftn(f)
    out = {}
    for w in ((8, 16, 32) if prec == "fp32" else (8, 16)):
        P.set_tiles(w, w)
        F = P.r2c(ft)
        g = P.c2r(F.clone())
        out[w] = (F.cpu().numpy(), g.cpu().numpy())
    tol = 1e-12 if prec == "fp64" else 2e-6
    for w, (F, g) in out.items():
        assert np.array_equal(F, out[16][0]) and np.array_equal(g, out[16][1]), w
    assert np.abs(out[16][0] - ref).max() <= tol * np.abs(ref).max()
    # a fused evaluation (three convolution fields through the pencils)
    P2 = _plan((n, n, n), np.eye(3) * (0.25 * n), prec)
    P2.set_functional(tf=True, vw=True, wt=True, lda=True, hartree=True, rho0=0.02)
    rho = torch.tensor(0.02 * (1.0 + 0.1 * rng.random((n, n, n))), dtype=P2.dtype, device=P2.device)
    sums = {}
    for w in ((8, 16, 32) if prec == "fp32" else (8, 16)):
        P2.set_tiles(w, w)
        V = P2.empty()
        sums[w] = (P2.eval(rho, V), V.cpu().numpy())
    for w in sums:
        assert np.array_equal(sums[w][1], sums[16][1]), w
        assert np.array_equal(sums[w][0][:6], sums[16][0][:6]), w
    with pytest.raises(Exception):
        P2.set_tiles(12, 16)


@pytest.mark.parametrize("stride", [1, 3])
def test_profiling_times_one_launch_in_stride(stride):
    """ofdft_plan_set_profiling: per-class launch counts, time and bytes;
    with a stride k only launches 0, k, 2k, ... of each class are timed."""
    print("case profiling stride", stride, flush=True)
    P = _plan((32, 32, 32), np.eye(3) * 10.0, "fp32")
    f = torch.rand((32, 32, 32), dtype=P.dtype, device=P.device)
    P.set_profiling(True, stride=stride)
    for _ in range(6):
        P.r2c(f)
    prof = P.get_profile()
    P.set_profiling(False)
    # each transform: one row pass, two axis passes
    n_rows, n_axis = 6, 12
    assert prof["rows_fwd"][1] == -(-n_rows // stride)
    assert prof["axis_c2c"][1] == -(-n_axis // stride)
    for k in ("rows_fwd", "axis_c2c"):
        ms, cnt, by = prof[k]
        assert ms > 0 and by > 0
    assert all(v[1] == 0 for k, v in prof.items() if k not in ("rows_fwd", "axis_c2c"))
    with pytest.raises(ValueError):
        P.set_profiling(True, stride=0)
