"""Density / potential file formats and the read/write dispatch of
formats/io.py:9-200: XSF (formats/xsf.py), Quantum ESPRESSO pp
(formats/qepp.py), Gaussian cube (formats/cube.py), snpy (formats/snpy.py:
a sequence of .npy records) and VASP POSCAR for ions (formats/vasp.py).

The readers return the reference's tuples: ``kind='ions'`` -> Ions,
``kind='data'`` -> DirectField, ``kind='all'`` -> (ions, field, info).
Fields are built on the grid's device (``device=`` picks it; default the
current GPU).  Host-side text and .npy I/O: nothing here is on the solver's
hot path.  Units follow the reference: XSF and POSCAR are in Angstrom (the
XSF density in Angstrom^-3, potentials in eV); pp and cube are in bohr (pp
potentials in Ry).
"""
import os

import numpy as np
import torch

from .field import DirectField
from .grid import DirectGrid
from .ions import BOHR, Ions, symbol_of

HARTREE = 27.21138602436724  # eV per hartree (ase.units.Ha, CODATA 2014)

__all__ = ["read", "write", "read_all", "read_data", "read_density", "read_potential", "write_all",
           "write_density", "write_potential", "guess_format", "read_xsf", "write_xsf", "read_qepp",
           "write_qepp", "read_cube", "write_cube", "read_snpy", "write_snpy", "read_vasp", "write_vasp"]


def _host(field):
    """(rank, n0, n1, n2) float64 numpy copy of a field."""
    with torch._C.DisableTorchFunctionSubclass():
        a = field.detach().to(torch.float64).cpu().numpy() if isinstance(field, torch.Tensor) else np.asarray(field)
    return a.reshape((1,) + a.shape) if a.ndim == 3 else a


def _field(grid, data):
    data = np.asarray(data, dtype=np.float64)
    rank = data.shape[0] if data.ndim == 4 else 1
    return DirectField(grid=grid, data=np.ascontiguousarray(data), rank=rank)


# ---------------------------------------------------------------------------
# XSF (formats/xsf.py:6-246)
# ---------------------------------------------------------------------------
def _lines(fh):
    for line in fh:
        line = line.strip()
        if line and not line.startswith("#"):
            yield line


def read_xsf(infile, kind="all", full=False, pbc=True, units="angstrom", data_type="density", device=None, **kwargs):
    """read_xsf (formats/xsf.py:14-140): crystal cell, PRIMCOORD atoms and
    the 3D (or 2D) DATAGRID blocks (one per spin); with pbc the repeated
    last point of each axis is dropped."""
    ang = units.lower() == "angstrom"
    with open(infile) as fr:
        fh = _lines(fr)
        if next(fh).upper() != "CRYSTAL":
            raise AttributeError("%s: the XSF reader takes periodic (CRYSTAL) files only" % infile)
        lattice = np.zeros((3, 3))
        line = next(fh).upper()
        if line.startswith("PRIMVEC"):
            lattice = np.array([[float(x) for x in next(fh).split()[:3]] for _ in range(3)])
            line = next(fh).upper()
        if line.startswith("CONVVEC"):
            for _ in range(3):
                next(fh)
            line = next(fh).upper()
        label, pos = [], []
        if line.startswith("PRIMCOORD"):
            nat = int(next(fh).split()[0])
            for _ in range(nat):
                w = next(fh).split()
                label.append(symbol_of(w[0]))
                pos.append([float(x) for x in w[1:4]])
            line = next(fh).upper() if kind != "ions" else ""
        ions = Ions(symbols=label, positions=np.array(pos).reshape(-1, 3), cell=lattice, units="ase" if ang else "au")
        if kind == "ions":
            return ions
        if line.startswith("BEGIN_BLOCK_DATAGRID_"):
            next(fh)  # block name
        blocks = []
        vlat = None
        for line in fh:
            u = line.upper()
            if u.startswith("BEGIN_DATAGRID"):
                nrx = np.ones(3, dtype=int)
                npbc = 3
                if u.startswith("BEGIN_DATAGRID_3D"):
                    nrx[:] = [int(x) for x in next(fh).split()[:3]]
                elif u.startswith("BEGIN_DATAGRID_2D"):
                    nrx[:2] = [int(x) for x in next(fh).split()[:2]]
                    npbc = 2
                else:
                    raise AttributeError("only 2D and 3D XSF data grids are read")
                next(fh)  # origin
                vlat = np.zeros((3, 3))
                for i in range(npbc):
                    vlat[i] = [float(x) for x in next(fh).split()[:3]]
                if npbc == 2:
                    c = np.cross(vlat[0], vlat[1])
                    vlat[2] = c / np.sqrt(np.dot(c, c))
                vals = []
                for line in fh:
                    if line[0] == "E":
                        break
                    vals.extend(float(x) for x in line.split())
                data = np.asarray(vals)[: int(np.prod(nrx))]
                blocks.append(data.reshape(nrx, order="F"))
            if line.strip().upper().startswith("END_BLOCK_DATAGRID"):
                break
    if not blocks:
        raise AttributeError("%s: no DATAGRID block found" % infile)
    shape = blocks[0].shape
    out = []
    for b in blocks:
        if b.shape != shape:
            raise AttributeError("%s: the DATAGRID blocks differ in shape" % infile)
        nr = np.array(b.shape)
        if pbc:
            nr = np.where(nr > 1, nr - 1, nr)
            b = b[: nr[0], : nr[1], : nr[2]]
        out.append(b * BOHR ** 3 if ang else b)
    data_lat = vlat / BOHR if ang else vlat
    grid = DirectGrid(lattice=data_lat, nr=out[0].shape, full=full, device=device)
    arr = np.stack(out) if len(out) > 1 else out[0]
    if data_type == "potential":
        arr = arr / HARTREE
    field = _field(grid, arr)
    return field if kind == "data" else (ions, field, None)


def _fmt_line(values):
    return "\n  " + "".join(str(v) + "    " for v in values)


def write_xsf(filexsf, ions=None, data=None, data_type="density", units="angstrom", title="DFTpy", **kwargs):
    """XSF.write (formats/xsf.py:150-246): values with one periodic image
    appended on each axis, Fortran order, five per line."""
    ang = units.lower() == "angstrom"
    s = BOHR if ang else 1.0
    arr = _host(data)
    grid = data.grid
    with open(filexsf, "w") as fo:
        fo.write("# " + str(title) + "    ")
        fo.write("\n  CRYSTAL \n    ")
        fo.write("\n  PRIMVEC    ")
        for v in ions.cell * s:
            fo.write(_fmt_line(v))
        fo.write("\n  PRIMCOORD    ")
        fo.write(_fmt_line((ions.nat, 1)))
        for sym, p in zip(ions.symbols, ions.positions * s):
            fo.write(_fmt_line([sym] + list(p)))
        fo.write("\n  BEGIN_BLOCK_DATAGRID_3D    ")
        fo.write("\n  3d_datagrid_{}    ".format(data_type))
        nr = arr.shape[1:]
        for ispin, a in enumerate(arr):
            v = np.pad(a, [(0, 1)] * 3, mode="wrap").ravel(order="F")
            if ang:
                v = v / BOHR ** 3
            if data_type == "potential":
                v = v * HARTREE
            fo.write("\n  BEGIN_DATAGRID_3D#{}    ".format(ispin))
            fo.write(_fmt_line((nr[0] + 1, nr[1] + 1, nr[2] + 1)))
            fo.write(_fmt_line(np.asarray(getattr(grid, "origin", np.zeros(3)), dtype=float) * s))
            for row in grid.lattice * s:
                fo.write(_fmt_line(row))
            for i in range(0, v.size, 5):
                fo.write(_fmt_line(v[i:i + 5]))
            fo.write("\n  END_DATAGRID_3D    ")
        fo.write("\n  END_BLOCK_DATAGRID_3D    ")


# ---------------------------------------------------------------------------
# Quantum ESPRESSO pp (formats/qepp.py:7-218)
# ---------------------------------------------------------------------------
def _celldm2at(ibrav, celldm):
    """PP.celldm2at (formats/qepp.py:122-136): ibrav 1 and 2 only."""
    at = np.zeros((3, 3))
    if ibrav == 1:
        at = celldm[0] * np.identity(3)
    elif ibrav == 2:
        at[:, 0] = 0.5 * celldm[0] * np.array([-1.0, 0.0, 1.0])
        at[:, 1] = 0.5 * celldm[0] * np.array([0.0, 1.0, 1.0])
        at[:, 2] = 0.5 * celldm[0] * np.array([-1.0, 1.0, 0.0])
    else:
        raise NotImplementedError("celldm2at is only implemented for ibrav = 0 and ibrav = 1")
    return at


def read_qepp(infile, kind="all", full=True, data_type="density", device=None, **kwargs):
    """PP.read (formats/qepp.py:17-119): header, cell (ibrav 0, 1, 2),
    species with valence charges, atoms, data in Fortran order.  Returns
    (ions, field, cutoffvars) with the charges set from the valences."""
    with open(infile) as fh:
        title = fh.readline()
        w = fh.readline().split()
        nrx = np.array([int(x) for x in w[:3]])
        nat, ntyp = int(w[6]), int(w[7])
        w = fh.readline().split()
        ibrav = int(w[0])
        celldm = np.asarray(w[1:], dtype=float)
        if ibrav == 0:
            at = np.array([[float(x) for x in fh.readline().split()] for _ in range(3)]) * celldm[0]
        else:
            at = _celldm2at(ibrav, celldm)
        gcutm, dual, ecut, plot_num = (float(x) for x in fh.readline().split())
        info = dict(gcutm=gcutm, dual=dual, ecut=ecut, plot_num=int(plot_num), ibrav=ibrav, celldm=celldm,
                    title=title)
        atm, zval = [], {}
        for _ in range(ntyp):
            w = fh.readline().split()
            atm.append(w[1])
            zval[w[1]] = float(w[2])
        label, pos = [], []
        for _ in range(nat):
            w = fh.readline().split()
            pos.append([float(x) for x in w[1:4]])
            label.append(atm[int(w[4]) - 1])
        pos = np.asarray(pos).reshape(-1, 3)
        if ibrav == 0:
            ions = Ions(symbols=label, positions=pos * celldm[0], cell=at)
        else:
            ions = Ions(symbols=label, scaled_positions=pos, cell=at)
        ions.set_charges(zval)
        if kind == "ions":
            return ions
        vals = np.array(fh.read().split(), dtype=np.float64)
    grid = DirectGrid(lattice=at, nr=nrx, full=full, device=device)
    arr = vals[: int(np.prod(nrx))].reshape(nrx, order="F")
    if data_type == "potential":
        arr = arr * 0.5  # Ry -> Ha
    field = _field(grid, arr)
    return field if kind == "data" else (ions, field, info)


def write_qepp(filepp, ions=None, data=None, data_type="density", header=None, information=None, **kwargs):
    """PP.write (formats/qepp.py:138-197), %22.15e, five values per line."""
    info = dict(gcutm=1.0, dual=4.0, ecut=1.0)
    if information:
        info.update(information)
    arr = _host(data)[0]
    nr = arr.shape
    uniq = list(ions.symbols_uniq)
    celldm = info.get("celldm", None)
    if celldm is None:
        celldm = np.zeros(6)
        celldm[0] = np.sqrt((ions.cell[0] ** 2).sum())
    ibrav = int(info.get("ibrav", 0))
    zv = ions.zval

    def w(fo, items, newline):
        if newline:
            fo.write("\n  ")
        for x in (items if isinstance(items, (list, tuple, np.ndarray)) else [items]):
            fo.write(str(x) + "    ")

    with open(filepp, "w") as fo:
        fo.write(header or info.get("title", "DFTpy"))
        w(fo, list(nr), True)
        w(fo, list(nr), False)
        w(fo, [ions.nat, len(uniq)], False)
        w(fo, ibrav, True)
        w(fo, list(celldm), False)
        if ibrav == 0:
            for row in ions.cell / celldm[0]:
                w(fo, list(row), True)
        w(fo, info["gcutm"], True)
        w(fo, info["dual"], False)
        w(fo, info["ecut"], False)
        w(fo, info.get("plot_num", 1 if data_type == "potential" else 0), False)
        for ity, spc in enumerate(uniq):
            w(fo, [ity + 1, spc, zv.get(spc, 0.0)], True)
        tau = ions.positions / celldm[0]
        for iat in range(ions.nat):
            w(fo, iat + 1, True)
            w(fo, list(tau[iat]), False)
            w(fo, uniq.index(ions.symbols[iat]) + 1, False)
        fo.write("\n")
        v = arr.ravel(order="F")
        if data_type == "potential":
            v = v * 2.0  # Ha -> Ry
        for i in range(0, v.size, 5):
            fo.write(" ".join("%22.15e" % x for x in v[i:i + 5]))
            if i + 5 < v.size:
                fo.write("\n")


# ---------------------------------------------------------------------------
# Gaussian cube (formats/cube.py:6-112)
# ---------------------------------------------------------------------------
def read_cube(infile, kind="all", full=False, data_type="density", device=None, **kwargs):
    """read_cube (formats/cube.py:6-56): bohr, atoms by atomic number, data
    with X outer and Z inner (C order), optional spin count after the origin."""
    with open(infile) as fh:
        fh.readline()
        fh.readline()
        w = fh.readline().split()
        nat = int(w[0])
        origin = np.array([float(x) for x in w[1:4]])
        rank = int(w[4]) if len(w) > 4 else 1
        nr, lattice = [], []
        for _ in range(3):
            w = fh.readline().split()
            nr.append(int(w[0]))
            lattice.append([float(x) for x in w[1:4]])
        nr = np.array(nr)
        lattice = np.array(lattice) * nr[:, None]
        numbers, pos = [], []
        for _ in range(nat):
            w = fh.readline().split()
            numbers.append(int(w[0]))
            pos.append([float(x) for x in w[2:5]])
        ions = Ions(numbers=numbers, positions=np.array(pos).reshape(-1, 3), cell=lattice)
        if kind == "ions":
            return ions
        vals = np.array(fh.read().split(), dtype=np.float64)
    grid = DirectGrid(lattice=lattice, nr=nr, full=full, origin=origin, device=device)
    n = int(np.prod(nr))
    r = vals.size // n
    if r > 1:
        arr = vals[: r * n].reshape((-1, r)).T.reshape((r,) + tuple(nr))
    else:
        arr = vals[:n].reshape(nr)
    field = _field(grid, arr)
    return field if kind == "data" else (ions, field, None)


def write_cube(filename, ions=None, data=None, data_type="density", header=None, origin=None, long=True, **kwargs):
    """write_cube (formats/cube.py:58-112).  Every float field is preceded by
    a space, so that a negative value at full width cannot run into the
    field before it (the reference's "{:15}{:22.15e}" can)."""
    if long:
        fmt = "{0:15} {1:22.15e} {2:22.15e} {3:22.15e}"
        fmt2 = fmt + " {4:22.15e}"
        fmt3 = "%22.15e"
    else:
        fmt = "{0:5} {1:12.6f} {2:12.6f} {3:12.6f}"
        fmt2 = fmt + " {4:12.6f}"
        fmt3 = "%13.5e"
    arr = _host(data)
    rank = arr.shape[0]
    shape = arr.shape[1:]
    v = arr.reshape((rank, -1)).ravel(order="F") if rank > 1 else arr.ravel()
    origin = np.zeros(3) if origin is None else np.asarray(origin, dtype=float)
    with open(filename, "w") as fh:
        fh.write((header or "DFTpy") + "\nOUTER LOOP: X, MIDDLE LOOP: Y, INNER LOOP: Z\n")
        fh.write(fmt.format(ions.nat, *origin))
        if rank > 1:
            fh.write("{0:8d}".format(rank))
        fh.write("\n")
        for i in range(3):
            fh.write(fmt.format(shape[i], *(ions.cell[i] / shape[i])) + "\n")
        ch = ions.get_charges()
        for z, c, p in zip(ions.numbers, ch, ions.positions):
            fh.write(fmt2.format(int(z), c, *p) + "\n")
        for i in range(0, v.size, 6):
            fh.write(" ".join(fmt3 % x for x in v[i:i + 6]))
            if i + 6 < v.size:
                fh.write("\n")


# ---------------------------------------------------------------------------
# snpy (formats/snpy.py:28-146, serial): .npy records in the order of the
# leading description: 1 cell, 2 numbers, 3 positions, 4 data, 5 information
# ---------------------------------------------------------------------------
def write_snpy(fname, ions=None, data=None, information=None, kind="all", desc=None, **kwargs):
    if desc is None:
        desc = np.arange(1, 4) if kind == "ions" else (np.arange(1, 6) if information is not None else np.arange(1, 5))
    with open(fname, "wb") as fh:
        np.save(fh, np.asarray(desc))
        for key in desc:
            if key == 1:
                np.save(fh, ions.cell)
            elif key == 2:
                np.save(fh, ions.numbers)
            elif key == 3:
                np.save(fh, ions.positions)
            elif key == 4:
                a = _host(data)
                np.save(fh, a[0] if a.shape[0] == 1 else a)
            elif key == 5:
                np.save(fh, np.asarray(information))


def read_snpy(fname, kind="all", full=False, grid=None, device=None, **kwargs):
    ions = info = data = None
    with open(fname, "rb") as fh:
        desc = np.load(fh)
        for key in desc:
            if key == 1:
                lattice = np.load(fh)
            elif key == 2:
                numbers = np.load(fh)
            elif key == 3:
                ions = Ions(numbers=numbers, positions=np.load(fh), cell=lattice)
                if kind == "ions":
                    return ions
            elif key == 5:
                info = np.load(fh)
            elif key == 4:
                a = np.load(fh)
                shape = a.shape[1:] if a.ndim == 4 else a.shape
                if grid is None:
                    grid = DirectGrid(lattice=lattice, nr=shape, full=full, device=device)
                data = _field(grid, a)
    if kind == "data":
        return data
    return ions, data, info


# ---------------------------------------------------------------------------
# VASP POSCAR, ions only (formats/vasp.py:4-101); Angstrom
# ---------------------------------------------------------------------------
def read_vasp(infile, names=None, **kwargs):
    with open(infile) as fh:
        fh.readline()
        scale = [float(x) for x in fh.readline().split()]
        scale = np.ones(3) * scale[0] if len(scale) == 1 else np.asarray(scale)
        lat = np.array([[float(x) for x in fh.readline().split()[:3]] for _ in range(3)]) * scale[:, None]
        w = fh.readline().split()
        if w[0].isdigit():
            typ = [int(x) for x in w]
        else:
            names = w
            typ = [int(x) for x in fh.readline().split()]
        if names is None:
            raise AttributeError("Must input the ions names")
        mode = fh.readline().strip()[0]
        if mode.lower() == "s":  # selective dynamics
            mode = fh.readline().strip()[0]
        nat = sum(typ)
        pos = np.array([[float(x) for x in fh.readline().split()[:3]] for _ in range(nat)])
    symbols = [n for n, c in zip(names, typ) for _ in range(c)]
    if mode.lower() in ("c", "k"):
        return Ions(symbols=symbols, positions=pos, cell=lat, units="ase")
    return Ions(symbols=symbols, scaled_positions=pos, cell=lat, units="ase")


def write_vasp(infile, ions=None, data=None, direct=False, fmt="%22.15f", header="DFTpy", **kwargs):
    cell = ions.cell * BOHR
    pos = ions.get_scaled_positions() if direct else ions.positions * BOHR
    names, inv, counts = np.unique(ions.symbols, return_inverse=True, return_counts=True)
    order = np.argsort(inv, kind="stable")
    with open(infile, "w") as fh:
        fh.write(header + "\n1.0\n")
        for vec in cell:
            fh.write("".join(fmt % x for x in vec) + "\n")
        fh.write("".join(f" {x:<3s}" for x in names) + "\n")
        fh.write("".join(f" {x:<3d}" for x in counts) + "\n")
        fh.write("Direct\n" if direct else "Cartesian\n")
        for p in pos[order]:
            fh.write(f"{fmt % p[0]} {fmt % p[1]} {fmt % p[2]}\n")


# ---------------------------------------------------------------------------
# dispatch (formats/io.py:9-200)
# ---------------------------------------------------------------------------
_FORMATS = {
    "xsf": (read_xsf, write_xsf, ("ions", "data", "all")),
    "qepp": (read_qepp, write_qepp, ("ions", "data", "all")),
    "cube": (read_cube, write_cube, ("ions", "data", "all")),
    "snpy": (read_snpy, write_snpy, ("ions", "data", "all")),
    "vasp": (read_vasp, write_vasp, ("ions",)),
}
_ALIASES = {"pp": "qepp", "poscar": "vasp", "contcar": "vasp"}


def guess_format(infile, **kwargs):
    """guess_format (formats/io.py:35-46): the base name (POSCAR, CONTCAR)
    or the extension."""
    base = os.path.basename(infile).lower()
    ext = os.path.splitext(infile)[1].lower()[1:]
    for key in (base, ext):
        key = _ALIASES.get(key, key)
        if key in _FORMATS:
            return key
    return ext if ext else base


def _driver(infile, format):
    fmt = _ALIASES.get(format, format) if format else guess_format(infile)
    if fmt not in _FORMATS:
        raise AttributeError("%s format not support yet" % fmt)
    return fmt, _FORMATS[fmt]


def read(infile, format=None, kind="ions", **kwargs):
    """io.read (formats/io.py:70-131, serial)."""
    fmt, (rd, _, kinds) = _driver(infile, format)
    if kind != "ions" and "data" not in kinds:
        if kind == "data":
            raise AttributeError("%s format has no volumetric data" % fmt)
        return rd(infile, **kwargs), None, None
    return rd(infile, kind=kind, **kwargs)


def write(outfile, ions=None, data=None, information=None, format=None, **kwargs):
    """io.write (formats/io.py:133-175, serial); ions and data may be given
    in either order as the reference allows."""
    if isinstance(data, Ions) and (ions is None or isinstance(ions, torch.Tensor)):
        ions, data = data, ions
    fmt, (_, wr, _) = _driver(outfile, format)
    return wr(outfile, ions=ions, data=data, information=information, **kwargs)


def read_data(infile, format=None, **kwargs):
    return read(infile, format=format, kind="data", **kwargs)


def read_all(infile, format=None, **kwargs):
    return read(infile, format=format, kind="all", **kwargs)


def read_density(infile, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return read(infile, format=format, kind="data", data_type="density", **kwargs)


def read_potential(infile, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return read(infile, format=format, kind="data", data_type="potential", **kwargs)


def write_all(outfile, ions=None, data=None, information=None, format=None, **kwargs):
    return write(outfile, ions=ions, data=data, information=information, format=format, **kwargs)


def write_density(outfile, data=None, ions=None, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return write(outfile, data=data, ions=ions, format=format, data_type="density", **kwargs)


def write_potential(outfile, data=None, ions=None, format=None, **kwargs):
    kwargs.pop("data_type", None)
    return write(outfile, data=data, ions=ions, format=format, data_type="potential", **kwargs)
