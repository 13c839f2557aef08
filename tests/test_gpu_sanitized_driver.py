"""The drop-in driver built with AddressSanitizer and UndefinedBehaviorSanitizer
on every host translation unit (fftsvd_amd/build_sanitized.py) runs the
whole reference flow on the GPU -- parameters, SIZ/CRG assignment, the SES
mesher, readSRF, operator construction (octree, lists, near field, bases,
apply plans with the pipelined translation schedule), GMRES, potentials
and the FRC output -- for the salt, Stern-layer and qualocation
formulations without a sanitizer report, and prints the same solvation
energy as the regular driver."""
import os
import subprocess

import pytest

from fftsvd_amd import build
from test_gpu_pbe_driver import write_inputs

pytestmark = pytest.mark.gpu

ASAN_PBE = os.path.join(build.LIBDIR, "asan", "fftsvd_pbe")
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def teardown_only(err):
    """The one report accepted: the sanitizer runtime's own device-allocator
    CHECK when, at process exit (__cxa_finalize of the HIP runtime), a GPU
    allocation still held in its quarantine is recycled after the HSA
    runtime has been unloaded.  It comes after the driver's last output and
    involves no frame of this library."""
    rep = err[err.find("AddressSanitizer"):] if "AddressSanitizer" in err else ""
    return (rep.startswith("AddressSanitizer: CHECK failed: sanitizer_allocator_device.h") and
            "__cxa_finalize" in rep and "ERROR: AddressSanitizer" not in err and "runtime error:" not in err and
            "libfftsvd_hip" not in rep)


def energy(stdout):
    return float([l for l in stdout.splitlines() if l.startswith("solvation energy")][0].split(":")[1].split()[0])


@pytest.mark.parametrize("salt,ionrad,extra", [(0.145, 0.0, ""), (0.145, 2.0, ""), (0.0, 0.0, "usequal=true\n")],
                         ids=["salt", "stern", "qualocation"])
def test_sanitized_driver(tmp_path, salt, ionrad, extra):
    assert os.path.exists(ASAN_PBE), "host-sanitized driver not built (python -m fftsvd_amd.build_sanitized)"
    runs = {}
    for name, exe in (("plain", build.PBE), ("sanitized", ASAN_PBE)):
        d = tmp_path / name
        d.mkdir()
        write_inputs(d, salt, ionrad=ionrad, extra=extra)
        print("case %s, %s driver" % (tmp_path.name, name), flush=True)
        env = dict(os.environ, **SAN_ENV) if name == "sanitized" else None
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=240, env=env)
        out = r.stdout + r.stderr
        print(out[:4000], "\n...\n" if len(out) > 8000 else "", out[max(4000, len(out) - 4000):], flush=True)
        runs[name] = energy(r.stdout)
        if name == "sanitized" and r.returncode != 0:
            assert teardown_only(r.stderr), "sanitizer report during the run"
            continue
        assert r.returncode == 0, name
        assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert abs(runs["sanitized"] - runs["plain"]) <= 1e-9 * abs(runs["plain"])
