"""INI front end (config/config.py:58-149): defaults, typed values,
unknown keys, the spacing/cutoff relation and the checks that keep the
job on the GPU path.  The job itself runs in tests/test_gpu_interface.py."""
import numpy as np
import pytest

from dftpy_amd import interface as I
from dftpy_amd.ions import BOHR


def _ini(tmp_path, text):
    p = tmp_path / "config.ini"
    p.write_text(text)
    return str(p)


def test_defaults_and_types(tmp_path, capsys):
    conf = I.ReadConf(_ini(tmp_path, """
[JOB]
calctype = energy force stress
[GRID]
nr = 32 32 30   # comment
bogus = 3
[KEDF]
kedf = tfvw
y = 1/9
[OPT]
method = tn
econv = 1e-7 * 2
[PP]
al = al.lda.recpot
"""))
    assert conf["JOB"]["task"] == "Optdensity"
    assert conf["JOB"]["calctype"] == ["Energy", "Force", "Stress"]
    assert conf["GRID"]["nr"] == [32, 32, 30]
    assert conf["KEDF"]["kedf"] == "TFVW" and abs(conf["KEDF"]["y"] - 1 / 9) < 1e-16
    assert conf["OPT"]["method"] == "TN" and conf["OPT"]["econv"] == 2e-7
    assert conf["OPT"]["maxiter"] == 100 and conf["OPT"]["c2"] == 0.2 and conf["OPT"]["maxfun"] == 50
    assert conf["PP"] == {"Al": "al.lda.recpot"}
    assert conf["MATH"]["linearie"] is True and conf["GRID"]["gfull"] is False
    assert "bogus" in capsys.readouterr().out


def test_spacing_and_cutoff_determine_each_other(tmp_path):
    c = I.ReadConf(_ini(tmp_path, "[GRID]\necut = 600\n"))
    h = c["GRID"]["spacing"] / BOHR  # bohr
    assert abs(np.pi ** 2 / (2 * h ** 2) * I.HARTREE_EV - 600.0) < 1e-9
    c = I.ReadConf(_ini(tmp_path, "[GRID]\nspacing = 0.2\n"))
    assert abs(c["GRID"]["ecut"] - np.pi ** 2 / (2 * (0.2 / BOHR) ** 2) * I.HARTREE_EV) < 1e-9


@pytest.mark.parametrize("line", ["[DENSITY]\nnspin = 2", "[MATH]\ntwostep = True", "[EXC]\nxc = PBE",
                                  "[DENSITY]\ndensityini = atomic", "[JOB]\ntask = Propagate"])
def test_options_off_the_gpu_path_are_rejected(tmp_path, line):
    conf = I.ReadConf(_ini(tmp_path, line + "\n"))
    with pytest.raises(AttributeError):
        I._check_supported(conf)


def test_stress_unit_conversion():
    # 1 Ha/bohr^3 in GPa from the CODATA-2014 hartree energy (4.359744650e-18 J)
    # and bohr radius (0.52917721067e-10 m)
    assert abs(I.HA_BOHR3_GPA - 4.359744650e-18 / 0.52917721067e-10 ** 3 / 1e9) < 1e-3
