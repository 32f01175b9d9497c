"""Build the gfx950 shared library librhyme_hip.so in-tree with hipcc.

All device code is compiled for gfx950 only, with FP contraction off (the
reference's arithmetic order is part of the FP64 parity contract) and packed
FP32 instructions disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librhyme_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["rhy_api.hip", "rhy_batch.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building librhyme_hip.so")


def build(verbose=False, extra=None, out=None):
    """Compile every translation unit in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor
    lib = out or LIB
    tmp = lib + ".tmp"
    objdir = os.path.join(HERE, "_build", "obj" + ("" if out is None else
                                                   "_" + os.path.basename(out)))
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def one(src):
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        _run([HIPCC] + compile_flags + (extra or []) + ["-c", os.path.join(CSRC, src),
                                                        "-o", obj], verbose)
        return obj
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(one, SOURCES))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, lib)
    return lib


# Libraries built from the same sources with the documented build options
# switched, for the tests that run the parity checks against each of them
# (selected per call through _lib.using).
VARIANT_DIR = os.path.join(HERE, "variants")
VARIANTS = {
    "powlib": ["-DRHY_FP64_POW_FAST=0"],       # device libm pow in the FP64 path
    "nocontact": ["-DRHY_CONTACT_EXACT=0"],    # Newton iteration on contacts too
    "noldswalk": ["-DRHY_LDS_WALK=0"],         # y/z walks re-read their cells
    "fastriemann": ["-DRHY_FP64_FAST_RIEMANN=1"],  # mixed-form FP64 Riemann policy
}


def variant_path(name):
    return os.path.join(VARIANT_DIR, "librhyme_hip_%s.so" % name)


def build_all(verbose=False):
    """The default library and every variant, built concurrently."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(VARIANT_DIR, exist_ok=True)
    jobs = [(None, None)] + [(variant_path(n), f) for n, f in VARIANTS.items()]
    with ThreadPoolExecutor(len(jobs)) as ex:
        futs = [ex.submit(build, verbose, f, o) for o, f in jobs]
        return [f.result() for f in futs]


if __name__ == "__main__":
    print(build_all(verbose=True))
