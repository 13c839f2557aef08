"""Build recipe of libfftsvd_hip.so (gfx950) with hipcc.

python -m fftsvd_amd.build          # or fftsvd_amd.build.build()

Objects go to fftsvd_amd/_build/, the shared library to
fftsvd_amd/_lib/libfftsvd_hip.so (in-tree, so it travels with the tree).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libfftsvd_hip.so")
PBE = os.path.join(LIBDIR, "fftsvd_pbe")          # drop-in driver built against the C interface
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["geometry.cpp", "tree.cpp", "host_linalg.cpp", "op.cpp", "solver.cpp", "capi.cpp", "mesh.cpp", "frontend.cpp",
           "fill.hip", "apply.hip", "solve.hip", "noalloc.hip"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-x", "hip",
          # no packed FP32 instructions (v_pk_*_f32) anywhere in device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Xarch_host", "-ffp-contract=off", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I", CSRC]


def _compile(src, outdir=BUILD, extra=()):
    obj = os.path.join(outdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + list(extra) + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    # -Bsymbolic-functions: calls inside the library bind to its own
    # definitions (error() in particular, which glibc also exports)
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-Wl,-Bsymbolic-functions", "-o", tmp] + objs + \
        ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    # the reference-style driver, plain C against include/FFTSVDpbeAPI.h
    cmd = ["gcc", "-O2", "-std=c99", "-Wall", "-I", INCLUDE, os.path.join(HERE, "apps", "fftsvd_pbe.c"), "-L", LIBDIR,
           "-lfftsvd_hip", "-Wl,-rpath,$ORIGIN", "-o", PBE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    if verbose:
        print("built", LIB, "and", PBE)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
