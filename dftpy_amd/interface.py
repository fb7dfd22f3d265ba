"""INI front end and job driver on the GPU path: the reference's
``dftpy config.ini`` (cui/main.py:19-57), ReadConf / OptionFormat /
ConfSpecialFormat (config/config.py:58-149), ConfigParser,
OptimizeDensityConf and evaluator2results (interface.py:21-317).

The option table below carries the sections, keys, value types and
defaults of the reference's configuration schema (configentries.json) for
the options this path runs: the density optimization of one cell with the
local pseudopotential, LDA, Hartree and the TF / vW / TFvW / WT / GGA / LWT /
LMGP KEDFs, energies, forces and stress.  Keys of the schema that select something the
GPU path does not run (spin, atomic initial densities, multistep grids,
LibXC strings, other KEDF options) are accepted at their defaults and
rejected with AttributeError otherwise.
"""
import argparse
import configparser
import os
import re
import time

import numpy as np
import torch

from . import formats as io
from .field import DirectField
from .functional import KEDF, XC, Hartree, TotalFunctional
from .grid import DirectGrid
from .ions import BOHR
from .math_utils import ecut2nr
from .optimization import Optimization
from .pseudo import LocalPseudo

HARTREE_EV = 27.21138602436724  # ase.units.Ha (CODATA 2014)
_E_2014 = 1.6021766208e-19  # elementary charge, CODATA 2014 (ase.units)
# Ha/bohr^3 -> GPa (constants.py:75: Ha/Bohr3 -> eV/A^3 -> GPa)
HA_BOHR3_GPA = HARTREE_EV / BOHR ** 3 * _E_2014 * 1e30 / 1e9


# --------------------------------------------------------------------------
# option table: section -> key -> (type, default)
# --------------------------------------------------------------------------
def _bool(s):
    return s.strip().lower()[0] not in ("n", "f", "0")


def _float(s):
    # arithmetic expressions ("1e-6 * 2") as the reference's numexpr/eval path
    return float(eval(s, {"__builtins__": {}}, {}))


def _path(s):
    return re.subn(r"\$\{(.*?)\}", lambda m: os.environ.get(m.group(1), ""), s)[0]


def _intlist(s):
    return [int(x) for x in s.split()]


_TYPES = {
    "bool": _bool, "int": int, "float": _float, "str": str, "lstr": str.lower, "ustr": str.upper,
    "cstr": str.capitalize, "path": _path, "intlist": _intlist,
    "floatlist": lambda s: [_float(x) for x in s.split()], "strlist": str.split,
    "cstrlist": lambda s: s.title().split(), "lstrlist": lambda s: s.lower().split(),
}

SCHEMA = {
    "JOB": {"task": ("cstr", "Optdensity"), "calctype": ("cstrlist", ["Energy"])},
    "PATH": {"pppath": ("path", "./"), "cellpath": ("path", "./")},
    "MATH": {"linearii": ("bool", True), "linearie": ("bool", True), "twostep": ("bool", False),
             "multistep": ("int", 1), "reuse": ("bool", True)},
    "PP": {},
    "CELL": {"cellfile": ("str", "POSCAR"), "elename": ("cstrlist", None), "zval": ("floatlist", None),
             "format": ("str", None)},
    "GRID": {"ecut": ("float", 600.0), "spacing": ("float", None), "gfull": ("bool", False), "nr": ("intlist", None),
             "maxprime": ("int", 13), "scale": ("float", 0.99), "cplx": ("bool", False)},
    "DENSITY": {"nspin": ("int", 1), "magmom": ("float", 0.0), "densityini": ("lstr", "heg"),
                "densityfile": ("path", None), "densityoutput": ("path", None)},
    "EXC": {"xc": ("str", None), "libxc": ("lstrlist", ["lda_x", "lda_c_pz"]), "x_str": ("str", None),
            "c_str": ("str", None)},
    "KEDF": {"kedf": ("ustr", "WT"), "x": ("float", 1.0), "y": ("float", 1.0), "alpha": ("float", None),
             "beta": ("float", None), "sigma": ("float", None), "rho0": ("float", None), "k_str": ("str", None),
             "params": ("floatlist", None), "nsp": ("int", None), "interp": ("str", "hermite"),
             "kerneltype": ("str", None), "symmetrization": ("str", None), "lumpfactor": ("floatlist", None),
             "neta": ("int", 50000), "etamax": ("float", 50.0), "order": ("int", 3), "ratio": ("float", 1.15),
             "maxpoints": ("int", 1000), "delta": ("float", None), "kdd": ("int", 3), "kfmin": ("float", None),
             "kfmax": ("float", None), "ldw": ("float", None)},
    "OPT": {"method": ("ustr", "CG-HS"), "algorithm": ("ustr", "EMM"), "vector": ("str", "Orthogonalization"),
            "c1": ("float", 1e-4), "c2": ("float", 0.2), "maxls": ("int", 10), "econv": ("float", 1e-6),
            "maxfun": ("int", 50), "maxiter": ("int", 100), "xtol": ("float", 1e-12), "h0": ("float", 1.0)},
    "OUTPUT": {"time": ("bool", True), "stress": ("bool", True), "electrostatic_potential": ("str", None)},
}


def DefaultOption():
    """The default of every option (config/config.py:38-55)."""
    return {sec: {k: (list(v[1]) if isinstance(v[1], list) else v[1]) for k, v in keys.items()}
            for sec, keys in SCHEMA.items()}


def ConfSpecialFormat(conf):
    """config/config.py:58-95: the grid spacing (Angstrom) and the cutoff
    (eV) determine each other, Ecut = pi^2 / (2 h^2)."""
    g = conf["GRID"]
    if g["spacing"]:
        g["ecut"] = np.pi ** 2 / (2 * g["spacing"] ** 2) * HARTREE_EV * BOHR ** 2
    else:
        g["spacing"] = np.sqrt(np.pi ** 2 / g["ecut"] * 0.5 * HARTREE_EV) * BOHR
    if conf["MATH"]["twostep"] and conf["MATH"]["multistep"] == 1:
        conf["MATH"]["multistep"] = 2
    return conf


def ReadConf(infile):
    """ReadConf + OptionFormat (config/config.py:114-149): defaults, then
    the INI values converted by type (text after '#' or '!' dropped); keys
    outside the table are reported and ignored; [PP] keys are element
    names (capitalised) mapped to pseudopotential files."""
    cp = configparser.ConfigParser()
    if not cp.read(infile):
        raise FileNotFoundError(infile)
    conf = DefaultOption()
    for section in cp.sections():
        if section not in conf:
            print('!WARN : "%s" not in the dictionary' % section)
            continue
        for key in cp.options(section):
            raw = cp.get(section, key)
            if section == "PP":
                conf["PP"][key.capitalize()] = raw
            elif key not in SCHEMA[section]:
                print('!WARN : "%s.%s" not in the dictionary' % (section, key))
            else:
                value = re.split("#|!", raw)[0]
                conf[section][key] = _TYPES[SCHEMA[section][key][0]](value) if value.strip() else None
    return ConfSpecialFormat(conf)


def PrintConf(conf):
    import json
    s = json.dumps(conf, indent=4, sort_keys=True, default=str)
    print(s)
    return s


def _check_supported(conf):
    bad = []
    if conf["DENSITY"]["nspin"] != 1:
        bad.append("DENSITY.nspin > 1")
    if conf["DENSITY"]["densityini"] not in ("heg", "read"):
        bad.append("DENSITY.densityini = %s" % conf["DENSITY"]["densityini"])
    if conf["MATH"]["multistep"] > 1:
        bad.append("MATH.multistep / twostep")
    if conf["GRID"]["cplx"]:
        bad.append("GRID.cplx")
    if conf["OPT"]["algorithm"] != "EMM":
        bad.append("OPT.algorithm = %s" % conf["OPT"]["algorithm"])
    xc = conf["EXC"]["xc"]
    if xc is not None and xc.upper() != "LDA":
        bad.append("EXC.xc = %s" % xc)
    if xc is None and conf["EXC"]["libxc"] != ["lda_x", "lda_c_pz"]:
        bad.append("EXC.libxc = %s" % " ".join(conf["EXC"]["libxc"]))
    if not any(t in conf["JOB"]["task"] for t in ("Optdensity", "Calculation")):
        bad.append("JOB.task = %s" % conf["JOB"]["task"])
    if bad:
        raise AttributeError("not on the GPU path: " + ", ".join(bad))


def ConfigParser(config, ions=None, rhoini=None, pseudo=None, grid=None):
    """interface.py:21-142: ions (and a density when the cell file holds
    one), the grid, the functionals and the initial density."""
    if isinstance(config, str):
        config = ReadConf(config)
        PrintConf(config)
    _check_supported(config)
    field = None
    if ions is None:
        ions, field, _ = io.read(os.path.join(config["PATH"]["cellpath"], config["CELL"]["cellfile"]),
                                 format=config["CELL"]["format"], names=config["CELL"]["elename"], kind="all")
    lattice = ions.cell
    if config["GRID"]["nr"] is not None:
        nr = np.asarray(config["GRID"]["nr"])
    elif field is not None:
        nr = np.asarray(field.grid.nrR)
        rhoini = field
    else:
        nr = ecut2nr(lattice=lattice, spacing=config["GRID"]["spacing"] / BOHR,
                     max_prime=int(config["GRID"]["maxprime"]))
    print("The final grid size is ", nr)
    if grid is None:
        grid = DirectGrid(lattice=lattice, nr=nr, full=config["GRID"]["gfull"])
    pplist = {k.capitalize(): os.path.join(config["PATH"]["pppath"], v) for k, v in config["PP"].items()}
    if pseudo is None:
        pseudo = LocalPseudo(grid=grid, ions=ions, PP_list=pplist, PME=config["MATH"]["linearie"])
    kopt = {k: v for k, v in config["KEDF"].items() if k != "kedf" and v is not None}
    ke = KEDF(config["KEDF"]["kedf"], **kopt)
    evaluator = TotalFunctional(KineticEnergyFunctional=ke, XCFunctional=XC("LDA"), HARTREE=Hartree(),
                                PSEUDO=pseudo)
    ntot = ions.get_ncharges()
    if rhoini is not None:
        rho = DirectField(grid=grid, data=rhoini)
    elif config["DENSITY"]["densityini"] == "read":
        rho = DirectField(grid=grid, data=io.read_density(config["DENSITY"]["densityfile"]))
    else:
        rho = DirectField(grid=grid)
        with torch._C.DisableTorchFunctionSubclass():
            rho.fill_(ntot / grid.volume)
    # normalise the charge (e.g. when the cell changed)
    with torch._C.DisableTorchFunctionSubclass():
        rho.mul_(ntot / (float(rho.sum()) * grid.dV))
    return config, {"ions": ions, "field": rho, "rho": rho, "E_v_Evaluator": evaluator, "nr2": nr}


def OptimizeDensityConf(config, ions=None, rho=None, E_v_Evaluator=None, **kwargs):
    """interface.py:145-230: the density optimization with econv per atom,
    the density output, then evaluator2results for JOB.calctype."""
    if "Optdensity" in config["JOB"]["task"]:
        opts = dict(config["OPT"])
        opts["econv"] *= ions.nat
        method = opts.pop("method")
        opt = Optimization(EnergyEvaluator=E_v_Evaluator, optimization_options=opts, optimization_method=method)
        rho = opt.optimize_rho(guess_rho=rho)
    print("-" * 80)
    calc = set()
    ct = config["JOB"]["calctype"]
    if "Both" in ct:
        calc.update({"E", "V"})
    if "Potential" in ct:
        calc.add("V")
    if "Energy" in ct:
        calc.add("E")
    if "Force" in ct:
        calc.add("F")
    if "Stress" in ct:
        calc.add("S")
    if config["DENSITY"]["densityoutput"]:
        print("Write Density...")
        io.write(config["DENSITY"]["densityoutput"], ions=ions, data=rho)
    if config["OUTPUT"]["electrostatic_potential"]:
        # properties.get_electrostatic_potential: Hartree plus local PP potential
        print("Write electrostatic potential...")
        v = E_v_Evaluator.HARTREE.compute(rho, calcType=("V",)).potential
        with torch._C.DisableTorchFunctionSubclass():
            v = DirectField(grid=rho.grid, data=v + E_v_Evaluator.PSEUDO.vreal)
        io.write(config["OUTPUT"]["electrostatic_potential"], ions=ions, data=v, data_type="potential")
    return evaluator2results(E_v_Evaluator, rho=rho, calcType=calc, ions=ions,
                             print_stress=config["OUTPUT"]["stress"])


def evaluator2results(evaluator, rho=None, calcType=("E", "F"), ions=None, print_stress=False, split=True, **kwargs):
    """interface.py:245-317: the energy table, force summary and stress
    printout, and the results dictionary."""
    calcType = set(calcType)
    forces, stress = {}, {}
    ep = {}
    if "E" in calcType or "V" in calcType:
        print("Calculate Energy/Potential...")
        ep = evaluator.get_energy_potential(rho, calcType=tuple(calcType & {"E", "V"}) or ("E",), split=split)
    if "E" in calcType:
        print(format("Energy information", "-^80"))
        ke = 0.0
        for key in sorted(ep):
            if key == "TOTAL":
                continue
            val = ep[key].energy
            print("{:>10s} energy (eV): {:22.15E}".format(key, val * HARTREE_EV))
            if key.startswith("KEDF"):
                ke += val
        etot = ep["TOTAL"].energy
        print("{:>10s} energy (eV): {:22.15E}".format("TOTAL", etot * HARTREE_EV))
        print("-" * 80)
        f = "  {:<30s} : {:30.15f}"
        print(f.format("kedfs energy (a.u.)", ke))
        print(f.format("kedfs energy (eV)", ke * HARTREE_EV))
        print(f.format("total energy (a.u.)", etot))
        print(f.format("total energy (eV)", etot * HARTREE_EV))
        if ions is not None:
            print(f.format("total energy (eV/atom)", etot * HARTREE_EV / ions.nat))
    if "F" in calcType:
        print("Calculate Force...")
        forces = evaluator.get_forces(rho, split=split)
        fa = np.abs(forces["TOTAL"])
        fs = " " * 8 + "{0:>22s} : {1:<22.5f}"
        print(fs.format("Max force (a.u.)", fa.max()))
        print(fs.format("Min force (a.u.)", fa.min()))
        print(fs.format("Ave force (a.u.)", fa.mean()))
        print("-" * 80)
    if "S" in calcType:
        print("Calculate Stress...")
        stress = evaluator.get_stress(rho, split=split)
        ss = " " * 16 + "{0[0]:12.5f} {0[1]:12.5f} {0[2]:12.5f}"
        if print_stress:
            for key in sorted(stress):
                if key == "TOTAL":
                    continue
                print("{:>10s} stress (a.u.): ".format(key))
                for i in range(3):
                    print(ss.format(stress[key][i]))
        print("{:>10s} stress (a.u.): ".format("TOTAL"))
        for i in range(3):
            print(ss.format(stress["TOTAL"][i]))
        print("{:>10s} stress (GPa): ".format("TOTAL"))
        for i in range(3):
            print(ss.format(stress["TOTAL"][i] * HA_BOHR3_GPA))
        print("-" * 80)
    results = {"density": rho, "energypotential": ep, "forces": forces, "stress": stress, "evaluator": evaluator}
    if hasattr(evaluator, "PSEUDO"):
        results["pseudo"] = evaluator.PSEUDO
    return results


def RunJob(args):
    """cui/main.py:19-57 (density optimization jobs)."""
    print("DFTpy (MI355X) Begin on : {}".format(time.strftime("%Y-%m-%d %H:%M:%S", time.localtime())))
    confs = list(args.confs) or [args.input]
    out = []
    for fname in confs:
        config, others = ConfigParser(fname)
        print("#" * 80)
        t0 = time.perf_counter()
        out.append(OptimizeDensityConf(config, **others))
        print("{:>30s} : {:10.3f} s".format("TOTAL", time.perf_counter() - t0))
        print("-" * 80)
    print("#" * 80)
    print("DFTpy (MI355X) Finished on : {}".format(time.strftime("%Y-%m-%d %H:%M:%S", time.localtime())))
    return out


def main(argv=None):
    parser = argparse.ArgumentParser(description="Process task")
    parser.add_argument("confs", nargs="*")
    parser.add_argument("-i", "--ini", "--input", dest="input", type=str, default="config.ini",
                        help="Input file (default: config.ini)")
    return RunJob(parser.parse_args(argv))
