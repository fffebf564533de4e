"""Builds libhz.so (HIP kernels + C-ABI host library) in-tree for gfx950.

    python build.py            # compile every source, link libhz.so next to this file
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhz.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "hz_setup.cpp",
    "hz_abi.cpp",
    "hz_krylov.cpp",
    "hz_ref_shim.cpp",
    "hz_tune.cpp",
    "hz_io.cpp",
    "k_fft.hip",
    "k_stencil.hip",
    "k_blocktri.hip",
    "k_vec.hip",
    "hz_noalloc.hip",
    "k_hodlr.hip",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # no packed FP32 instructions in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return out, warn


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            if w:
                print(w)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_ref_driver()
    return LIB


DRIVER_SRC = os.path.join(HERE, "..", "tests", "ref_shim_driver.cpp")
DRIVER = os.path.join(OBJ, "ref_shim_driver")


def build_ref_driver():
    """The test caller of the reference-signature layer (tests/ref_shim_driver.cpp),
    a host-only C++ program linked against libhz.so."""
    if not os.path.exists(DRIVER_SRC):
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(HERE, "..", "include"), DRIVER_SRC, "-o", DRIVER,
           "-L" + HERE, "-lhz", "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
