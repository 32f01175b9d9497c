"""Build the Fortran host modules (rhy_capi.f90, rhyme_host.f90) with flang
for NDIM = 1, 2, 3, and the host drivers used by the tests
(tests/fhost/rhy_fhost_driver.f90, the main loop; tests/fhost/rhy_fhost_units.f90,
the unit procedures and forced errors), linked against librhyme_hip.so.

Outputs under rhyme_amd/fhost/_build/<N>d/: the .mod files, librhyme_fhost.a
and the executables rhy_fhost_driver.exe and rhy_fhost_units.exe.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
DRIVER = os.path.join(ROOT, "tests", "fhost", "rhy_fhost_driver.f90")
UNITS = os.path.join(ROOT, "tests", "fhost", "rhy_fhost_units.f90")


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("flang failed")


def build_one(nd):
    out = os.path.join(HERE, "_build", "%dd" % nd)
    os.makedirs(out, exist_ok=True)
    fl = ["-cpp", "-DNDIM=%d" % nd, "-O2", "-fPIC", "-module-dir", out, "-I", out]
    objs = []
    for src in ("rhy_capi.f90", "rhyme_host.f90"):
        o = os.path.join(out, src.replace(".f90", ".o"))
        run([FLANG] + fl + ["-c", os.path.join(HERE, src), "-o", o])
        objs.append(o)
    lib = os.path.join(out, "librhyme_fhost.a")
    if os.path.exists(lib):
        os.remove(lib)
    run(["ar", "rcs", lib] + objs)
    exe = os.path.join(out, "rhy_fhost_driver.exe")
    for src, name in ((DRIVER, "rhy_fhost_driver.exe"), (UNITS, "rhy_fhost_units.exe")):
        if os.path.exists(src):
            run([FLANG] + fl + [src, lib, "-L", PKG, "-lrhyme_hip",
                                "-Wl,-rpath,$ORIGIN/../../..", "-Wl,-rpath,/opt/rocm/lib",
                                "-o", os.path.join(out, name)])
    return exe


def build(dims=(1, 2, 3)):
    with ThreadPoolExecutor(3) as ex:
        return list(ex.map(build_one, dims))


if __name__ == "__main__":
    for e in build():
        print(e)
