"""The C-ABI library loads on a CPU-only host and exports every function
declared in include/*.h (no compute call is made here)."""
import ctypes
import os
import re

import pytest

import fftsvd_amd as fa


def declared_functions():
    names = []
    for h in fa.HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9\s\*]*?\b([A-Za-z_][A-Za-z_0-9]*)\s*\(", text,
                             flags=re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "switch", "return", "sizeof"):
                names.append(name)
    return sorted(set(names))


def test_headers_declare_solve_path():
    names = declared_functions()
    for required in ("readSRF", "generateSurfaceOperator", "generateSurfaceOperatorPreconditioner", "generateRHS",
                     "solveSurfaceOperator", "collectPotentials", "fftsvd_problem_create", "fftsvd_problem_solve",
                     "fftsvd_tree_apply", "fftsvd_tree_apply_transpose"):
        assert required in names


def test_library_exports_every_declared_function():
    if not os.path.exists(fa.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(fa.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_declare_extended_paths():
    names = declared_functions()
    for required in ("generateQualocationOperator", "solveQualocationOperator", "collectQualocationPotentials",
                     "solveBEMproblem", "fftsvd_qual_create", "fftsvd_problem_solve_batch",
                     "fftsvd_problem_solve_adjoint", "fftsvd_problem_grad_charges", "fftsvd_tree_cube_points"):
        assert required in names


def test_argument_errors_before_any_device_call():
    """Invalid sizes are rejected with a message before the library touches
    a device (so this runs on a CPU-only host)."""
    if not os.path.exists(fa.LIB_PATH):
        pytest.skip("library not built")
    import numpy as np
    v = np.zeros((4, 9))
    q = np.zeros((1, 3))
    c = np.zeros(1)
    with pytest.raises(fa.FFTSVDError, match="grid points"):
        fa.Problem(v, q, c, gridpoints=5)
    with pytest.raises(fa.FFTSVDError, match="grid points"):
        fa.Qual(v, q, c, gridpoints=2)
    with pytest.raises(fa.FFTSVDError, match="sizes"):
        fa.Qual(np.zeros((0, 9)), q, c)


def test_batch_entry_points_reject_bad_sizes():
    """The batched entry points check their sizes (and null handles) before
    any device call."""
    if not os.path.exists(fa.LIB_PATH):
        pytest.skip("library not built")
    import ctypes as C
    L = fa.lib()
    assert L.fftsvd_problems_solve(0, None, 0, None, None, 1e-6, 10, None, None) < 0
    assert "nprob" in L.fftsvd_last_error().decode()
    hs = (C.c_void_p * 1)(None)
    assert L.fftsvd_problems_solve(1, hs, 0, None, None, 1e-6, 0, None, None) < 0
    assert "maxit" in L.fftsvd_last_error().decode()
