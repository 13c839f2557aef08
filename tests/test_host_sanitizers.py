"""The host-side tests (parameter, PDB/CRD/XYZR, SIZ/CRG readers, the SES
mesher and generateFlatSRF, readSRF, argument checks of the C interface)
rerun against the host-sanitized library (fftsvd_amd/build_sanitized.py:
AddressSanitizer and UndefinedBehaviorSanitizer on every host translation
unit).  Any sanitizer report aborts the run."""
import os
import subprocess
import sys

import pytest

from fftsvd_amd import build_sanitized as S

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_TESTS = ["test_frontend_host.py", "test_refapi_host.py", "test_capi_symbols.py"]
SAN_ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


@pytest.mark.skipif(bool(os.environ.get("LD_PRELOAD")), reason="a preload is already set: the sanitizer "
                    "runtime must come first, and an existing preload is left alone")
def test_host_tests_under_asan_ubsan():
    assert os.path.exists(S.LIB), "host-sanitized library not built (python -m fftsvd_amd.build_sanitized)"
    env = dict(os.environ, LD_PRELOAD=S.runtime(), FFTSVD_LIBRARY=S.LIB, **SAN_ENV)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider"] +
                       [os.path.join(HERE, t) for t in HOST_TESTS], env=env, cwd=os.path.dirname(HERE),
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    print(out[-3000:])
    assert r.returncode == 0
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert " passed" in out
