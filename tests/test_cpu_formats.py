"""File formats (formats/io.py, xsf.py, qepp.py, cube.py, snpy.py,
vasp.py) on the host: the reference's data files read to the values the
oracle reads, write/read round trips, and the snpy test of
examples/test/test_snpy.py:19-36."""
import os

import numpy as np
import pytest

import oracle as O
from dftpy_amd import formats as F
from dftpy_amd.ions import BOHR, Ions
from dftpy_amd.grid import DirectGrid
from dftpy_amd.field import DirectField
from tests.conftest import DATA


def _np(f):
    return f.detach().cpu().numpy()


def test_qepp_fixture_matches_oracle_reader():
    ions, rho, info = F.read_qepp(os.path.join(DATA, "Al_fde_rho.pp"), device="cpu")
    oi, orho, og = O.read_qepp(os.path.join(DATA, "Al_fde_rho.pp"), full=True)
    assert np.array_equal(_np(rho), orho)
    assert np.allclose(rho.grid.lattice, og.lattice, rtol=0, atol=1e-14)
    assert np.allclose(ions.positions, oi.positions, rtol=0, atol=1e-12)
    assert ions.symbols == list(oi.symbols)
    assert rho.grid.full and ions.zval["Al"] == 3.0


def test_poscar_fixture_matches_oracle_reader():
    ions = F.read(os.path.join(DATA, "GaAs_random.vasp"), format="vasp")
    oi = O.read_poscar(os.path.join(DATA, "GaAs_random.vasp"))
    assert ions.symbols == list(oi.symbols)
    assert np.allclose(ions.cell, oi.cell, rtol=0, atol=1e-12)
    assert np.allclose(ions.positions, oi.positions, rtol=0, atol=1e-12)
    assert F.guess_format("x/POSCAR") == "vasp" and F.guess_format("a.pp") == "qepp"


def test_snpy_round_trip_of_xsf_fixture(tmp_path):
    """examples/test/test_snpy.py:19-36 on GaAs_random.xsf."""
    ions, field, _ = F.read_all(os.path.join(DATA, "GaAs_random.xsf"), device="cpu")
    assert tuple(field.shape) == (16, 26, 18)  # 17 x 27 x 19 with the periodic image dropped
    assert ions.nat == 10 and ions.symbols[:5] == ["Ga"] * 5
    out = str(tmp_path / "serial.snpy")
    F.write_snpy(out, ions, field)
    ions2, field2, _ = F.read_snpy(out, device="cpu")
    assert np.allclose(ions.cell, ions2.cell)
    assert np.allclose(ions.positions, ions2.positions)
    assert np.array_equal(ions.numbers, ions2.numbers)
    assert np.array_equal(_np(field), _np(field2))
    # the XSF values are in Angstrom^-3, the field in bohr^-3; the grid is the DATAGRID lattice
    raw = open(os.path.join(DATA, "GaAs_random.xsf")).read().split("BEGIN_DATAGRID_3D")[1].split()
    assert abs(float(raw[15]) * BOHR ** 3 - _np(field)[0, 0, 0]) < 1e-16
    assert np.allclose(field.grid.lattice[0], np.array([4.1604539101422, 0, 0]) / BOHR)


@pytest.mark.parametrize("fmt", ["xsf", "qepp", "cube", "snpy"])
def test_write_read_round_trip(fmt, tmp_path):
    cell = np.array([[7.1, 0.2, 0.0], [0.3, 6.5, 0.1], [-0.4, 0.2, 8.0]])
    ions = Ions(symbols=["Ga", "As", "As"], positions=[[0.1, 0.2, 0.3], [2.0, 3.1, 1.7], [5.5, 0.4, 6.2]], cell=cell)
    ions.set_charges({"Ga": 3.0, "As": 5.0})
    grid = DirectGrid(cell, nr=(6, 5, 7), full=False, device="cpu")
    data = np.random.default_rng(2).random((6, 5, 7))
    field = DirectField(grid=grid, data=data)
    out = str(tmp_path / ("f." + {"qepp": "pp"}.get(fmt, fmt)))
    F.write(out, ions=ions, data=field)
    ions2, f2, _ = F.read_all(out, device="cpu")
    tol = 0.0 if fmt in ("snpy",) else 1e-14
    assert np.abs(_np(f2) - data).max() <= tol * np.abs(data).max() + (1e-13 if fmt == "xsf" else 0.0)
    assert np.allclose(f2.grid.lattice, cell, rtol=0, atol=1e-12)
    assert np.allclose(ions2.positions, ions.positions, rtol=0, atol=1e-12)
    assert ions2.symbols == ions.symbols
    # a potential round-trips through the format's units (eV, Ry) as well
    if fmt in ("xsf", "qepp"):
        F.write_potential(out, data=field, ions=ions)
        p = F.read_potential(out, device="cpu")
        assert np.abs(_np(p) - data).max() <= 1e-13 * np.abs(data).max()
