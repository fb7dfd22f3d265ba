"""The C-ABI host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The C-ABI unit (dftpy_amd/csrc/ofdft_api.hip: argument checks, plan
creation and destruction, batch validation, the host side of every entry
point) is recompiled for the host only with -fsanitize=address,undefined,
embedding the device code object of the normal build (its .hip_fatbin
section), and linked with the other units of the normal build and the C
driver tests/csrc/abi_host_checks.c.  The driver needs no GPU: it runs
every entry point through its argument checks and sweeps plan creation over
5000 grid shapes (factorisation, LDS bounds, lattice inversion and, with no
device, the clean-up of a half-built plan).  Any heap or stack overflow,
use after free, double free or undefined behaviour aborts the driver."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "dftpy_amd", "_lib", "obj")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# each sanitizer flag goes to the host compilation only
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, "%s\n%s\n%s" % (" ".join(cmd), r.stdout[-3000:], r.stderr[-3000:])
    return r


def test_c_abi_host_code_under_asan_and_ubsan(tmp_path):
    api_obj = os.path.join(OBJ, "ofdft_api.o")
    if not (os.path.exists(HIPCC) and shutil.which("objcopy") and os.path.exists(api_obj)):
        pytest.skip("needs hipcc, objcopy and the built objects (python -m dftpy_amd.build_lib)")
    others = sorted(o for o in glob.glob(os.path.join(OBJ, "*.o")) if os.path.basename(o) != "ofdft_api.o")
    fatbin = str(tmp_path / "api.hipfb")
    _run(["objcopy", "--dump-section", ".hip_fatbin=" + fatbin, api_obj, str(tmp_path / "unused.o")])
    inc = ["-I", os.path.join(ROOT, "include")]
    api = str(tmp_path / "api_san.o")
    _run([HIPCC, "--offload-arch=gfx950", "--offload-host-only", "-Xclang", "-fcuda-include-gpubinary",
          "-Xclang", fatbin, "-O1", "-g", "-std=c++17", "-fPIC"] + SAN + inc +
         ["-c", os.path.join(ROOT, "dftpy_amd", "csrc", "ofdft_api.hip"), "-o", api])
    drv_o = str(tmp_path / "drv.o")
    _run([HIPCC, "-x", "c", "-g"] + SAN + inc + ["-c", os.path.join(ROOT, "tests", "csrc", "abi_host_checks.c"),
                                               "-o", drv_o])
    drv = str(tmp_path / "abi_host_checks")
    _run([HIPCC, "--offload-arch=gfx950"] + SAN + [drv_o, api] + others + ["-o", drv])
    env = dict(os.environ)
    # leaks inside the HIP runtime's own start-up are not this library's
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    r = _run([drv], env=env, timeout=600)
    print(r.stdout)
    assert r.stdout.strip().splitlines()[-1].startswith("ok ")
