"""Host side of the LWT / LMGP term on the CPU: the kernel tables handed to
the device (splines of the WT and MGP tables and their derivative tables),
the kf range and the node list, against the oracle's restatement of
kedf/lwt.py and kedf/kernel.py."""
import numpy as np
import pytest
from scipy.interpolate import splev

import oracle as O
from dftpy_amd import lwt as W


def test_host_lindhard_matches_oracle():
    eta = np.concatenate([np.linspace(0, 6, 6001), [1e-12, 1.0 - 1e-11, 1.0 + 5e-11, 3.65, 3.6500001, 40.0]])
    for lb, mu in ((1.0, 1.0), (0.6, 1.0), (1.0, 0.0)):
        a = W.lindhard(eta, lb, mu)
        b = O.lindhard(eta, lb, mu)
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("kt,sym", [("WT", None), ("MGP", None), ("MGPA", None), ("MGPG", None)])
def test_kernel_tables_match_oracle(kt, sym):
    neta, etamax, mp = 5000, 50.0, 100
    t = W.Tables(kt, 1.0, etamax, neta, 3, mp, sym)
    K, D = O.lwt_kernel_tables(kt, 1.0, etamax, neta, 3, mp, sym)
    x = np.linspace(0, etamax * 0.999, 3001)
    for tab, ref in zip(t.tabs, (K, D)):
        n = tab.n
        tk = np.ctypeslib.as_array(tab.knots, shape=(n,))
        ck = np.ctypeslib.as_array(tab.coefs, shape=(n,))
        got = splev(x, (tk, ck, 3))
        want = splev(x, ref)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
        assert abs(tab.limit - splev(etamax, ref)) <= 1e-12 * np.abs(want).max()
        assert tab.dx == etamax / (neta - 1) and tab.x0 == 0.0


def test_plain_tables_for_order_one():
    t = W.Tables("WT", 1.0, 50.0, 2000, 1, 100, None)
    K, D = O.lwt_kernel_tables("WT", 1.0, 50.0, 2000, 1)
    for tab, ref in zip(t.tabs, (K, D)):
        v = np.ctypeslib.as_array(tab.coefs, shape=(tab.n,))
        assert tab.kind == 1 and v.size == ref.size and tab.limit == v[-1]
        # the derivative table amplifies last-digit differences near eta = 1
        assert np.abs(v - ref).max() <= 1e-11 * np.abs(ref).max()


def test_kf_bound_and_nodes_match_oracle():
    rng = np.random.default_rng(3)
    saved_o, saved_h = O.LwtCache(), {"kfmin": None, "kfmax": None}
    for step in range(6):
        rho = 0.02 * np.exp(rng.normal(0.0, 1.0 + step, size=200))
        kf = 2.0 * np.cbrt(3.0 * np.pi ** 2 * rho)
        a = O.guess_kf_bound(kf, None, None, saved=saved_o)
        b = W.guess_kf_bound(kf.min(), kf.max(), None, None, saved=saved_h)
        assert a == b
        for opts in (dict(), dict(ratio=1.05), dict(nsp=7), dict(delta=0.1)):
            assert np.array_equal(O.lwt_kflist(a[0], a[1], **opts), W.kf_nodes(b[0], b[1], **opts))
    assert W.guess_kf_bound(0.5, 2.0, 0.3, 3.0) == [0.3, 3.0]


def test_mgpe_parameters():
    q = np.abs(np.random.default_rng(0).normal(size=(4, 4, 4)))
    q[0, 0, 0] = 0.0
    for lf in (0.2, [0.1, 0.05], [0.1, 0.05, 2.0]):
        a, b, c = W.LwtTerm(lumpfactor=lf)._mgpe(12.0)
        from scipy.special import erf
        gg = np.where(q > 0, q * q, 1.0)
        got = np.where(q > 0, 4 * np.pi * erf(c * gg) ** 2 * a * np.exp(-gg * b) / gg, 0.0)
        assert np.allclose(got, O.mgp_omega_e(q, 12.0, lf), rtol=1e-14, atol=0)


def test_bad_options_are_rejected():
    for bad in (dict(interp="cubic"), dict(kdd=4), dict(order=2)):
        with pytest.raises(AttributeError):
            W.LwtTerm(**bad)


def test_ini_kedf_section_reaches_the_lwt_term(tmp_path):
    """[KEDF] options of the reference's schema (configentries.json), with
    its INI default interp = hermite, arrive in the LMGP term."""
    from dftpy_amd import interface as I
    from dftpy_amd.functional import KEDF
    ini = tmp_path / "c.ini"
    ini.write_text("[KEDF]\nkedf = lmgpa\nmaxpoints = 100\nlumpfactor = 0.1 0.05\nratio = 1.1\nkdd = 2\n")
    conf = I.ReadConf(str(ini))
    kopt = {k: v for k, v in conf["KEDF"].items() if k != "kedf" and v is not None}
    k = KEDF(conf["KEDF"]["kedf"], **kopt)
    o = k.lwt.o
    assert k.engines == ("TF", "VW", "LWT-NL")
    assert o["kerneltype"] == "MGPA" and o["interp"] == "hermite" and o["maxpoints"] == 100
    assert o["lumpfactor"] == [0.1, 0.05] and o["ratio"] == 1.1 and o["kdd"] == 2
    assert o["neta"] == 50000 and o["etamax"] == 50.0 and o["order"] == 3
