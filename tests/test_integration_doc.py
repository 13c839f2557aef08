"""The complete C examples in INTEGRATION.md (the blocks that include a
header; the others are fragments) compile against include/ and link
against the built library, so they cannot drift from the header (CPU only:
nothing is run)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "fftsvd_amd", "_lib")


def _c_blocks():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    return [b for b in re.findall(r"```c\n(.*?)```", text, flags=re.S) if "#include" in b]


def test_integration_c_example_compiles_and_links(tmp_path):
    blocks = _c_blocks()
    assert blocks, "INTEGRATION.md has no C example"
    if not os.path.exists(os.path.join(LIBDIR, "libfftsvd_hip.so")):
        pytest.skip("library not built")
    for k, code in enumerate(blocks):
        src = tmp_path / ("example%d.c" % k)
        src.write_text(code + "\nint main(void) { return 0; }\n")
        exe = tmp_path / ("example%d" % k)
        r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-Wno-unused-function", "-I",
                            os.path.join(ROOT, "include"), str(src), "-L", LIBDIR, "-lfftsvd_hip", "-o", str(exe)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
