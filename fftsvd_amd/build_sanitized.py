"""Host-sanitized build of the library and the drop-in driver.

python -m fftsvd_amd.build_sanitized

The host translation units (.cpp) are compiled with AddressSanitizer and
UndefinedBehaviorSanitizer on the host pass only; device code is never
instrumented, and the .hip files (kernels and their launch wrappers) are
the objects of the regular build.  Outputs go to fftsvd_amd/_lib/asan/:
libfftsvd_hip.so (load it with FFTSVD_LIBRARY and the sanitizer runtime
preloaded) and fftsvd_pbe, the drop-in driver as an executable that
carries the runtime itself.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

from fftsvd_amd import build as B

ASAN_BUILD = os.path.join(B.BUILD, "asan")
ASAN_LIBDIR = os.path.join(B.LIBDIR, "asan")
LIB = os.path.join(ASAN_LIBDIR, "libfftsvd_hip.so")
PBE = os.path.join(ASAN_LIBDIR, "fftsvd_pbe")
LLVM_BIN = os.path.join(os.path.dirname(os.path.realpath(B.HIPCC)), "..", "llvm", "bin")
CLANG = os.path.join(LLVM_BIN, "clang")
CLANGXX = os.path.join(LLVM_BIN, "clang++")
ROCM_LIB = os.path.join(os.path.dirname(os.path.realpath(B.HIPCC)), "..", "lib")
HOST_ONLY = "-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined " \
            "-Xarch_host -fno-omit-frame-pointer -Xarch_host -g".split()
SAN = "-fsanitize=address,undefined -shared-libsan".split()


def runtime():
    """Path of the shared AddressSanitizer runtime (with UBSan inside)."""
    return subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                          text=True).stdout.strip()


def build(verbose=False):
    os.makedirs(ASAN_BUILD, exist_ok=True)
    os.makedirs(ASAN_LIBDIR, exist_ok=True)
    cpp = [s for s in B.SOURCES if s.endswith(".cpp")]
    hip = [os.path.join(B.BUILD, os.path.splitext(s)[0] + ".o") for s in B.SOURCES if s.endswith(".hip")]
    for o in hip:
        if not os.path.exists(o):
            raise RuntimeError("run fftsvd_amd.build first: %s is missing" % o)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: B._compile(s, ASAN_BUILD, HOST_ONLY), cpp))
    rpath = "-Wl,-rpath," + os.path.dirname(runtime())
    # the link is a host link (the device code objects are already bundled)
    cmd = [CLANGXX, "-shared", "-fPIC", "-Wl,-Bsymbolic-functions", rpath, "-o", LIB] + SAN + objs + hip + \
        ["-L", ROCM_LIB, "-lamdhip64", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitized link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    cmd = [CLANG, "-O1", "-g", "-std=c99", "-Wall", "-fno-omit-frame-pointer", "-I", B.INCLUDE,
           os.path.join(B.HERE, "apps", "fftsvd_pbe.c"), "-L", ASAN_LIBDIR, "-lfftsvd_hip", "-Wl,-rpath,$ORIGIN",
           rpath, "-o", PBE] + SAN
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitized driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    if verbose:
        print("built", LIB, "and", PBE)
    return LIB, PBE


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
