"""Many right-hand sides through the column-batched operator.

fftsvd_problem_matvec_batch applies the operator to ncol columns with one
launch per apply stage (each block of the single-vector launch has one copy
per column; the copies of a block read its operator rows together), and
fftsvd_problem_solve_batch runs one GMRES per column in lockstep over it
(and, in mixed mode, the FP64 refinement, itself batched).  Every column is
computed by the single-vector kernels' arithmetic, so the gates are
bit-for-bit equality with the single-vector calls:
- the batched matvec against fftsvd_problem_matvec, both precisions, on a
  depth-4 tree, a depth-7 tree (far-field cubes of 4 to 9 grid points per
  side) and a Stern layout with a salt cavity (five regions, four trees);
- the batched solve against fftsvd_problem_solve column by column:
  solution, iteration count, refinements, matvec counts and final residual,
  both modes, with a zero column (the zero-solution test) and with an
  iteration cap that stops columns unconverged;
- more than 64 columns (two lockstep groups).
The single-vector path itself is pinned against the oracle by the parity
tests; the reference's hessian() (FFTSVDpbe.c:52-99) solves the columns
one after another, which is what the single-vector calls do.
"""
import numpy as np
import pytest

import fftsvd_amd as fa
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu


def _ellipsoid(sub):
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=sub, jitter=0.04, seed=7)
    return meshes.panels_from_mesh(V, F)


def _sphere(R, sub, seed):
    V, F = meshes.icosphere(R, sub)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    Q[:, 0] *= np.sign(np.linalg.det(Q))
    return meshes.panels_from_mesh(V @ Q.T, F)


def _charges(n, ncol, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.8, 0.8, size=(n, 3)), rng.uniform(-0.8, 0.8, size=(n, ncol))


_cache = {}


def _problem(name):
    if name not in _cache:
        qpos, Q = _charges(8, 1, 3)
        if name == "d4":
            g = fa.Problem(_ellipsoid(3), qpos, Q[:, 0], 4.0, 80.0, 0.124)
        elif name == "d7":
            g = fa.Problem(_ellipsoid(4), qpos, Q[:, 0], 4.0, 80.0, 0.124, maxpanels=2)
        else:
            qpos = np.array([[0.0, 0.0, 2.2], [0.4, -0.3, 2.0]])
            g = fa.Problem(_sphere(3.0, 3, 1), qpos, [0.7, -0.4], 2.0, 80.0, 0.5, stern_verts9=_sphere(3.8, 3, 2),
                           cavities=[_sphere(1.6, 2, 3)], salt_cavities=[_sphere(1.0, 2, 4)], salt_cavity_parent=[0])
        _cache[name] = g
    return _cache[name]


@pytest.mark.parametrize("name", ["d4", "d7", "stern_cavities"])
@pytest.mark.parametrize("prec", [fa.F64, fa.MIXED])
def test_matvec_batch_equals_single(name, prec):
    g = _problem(name)
    X = np.random.default_rng(11).normal(size=(g.nvec, 5))
    print("batched matvec", name, prec, X.shape, flush=True)
    Y = g.matvec_batch(X, prec)
    for c in range(X.shape[1]):
        np.testing.assert_array_equal(Y[:, c], g.matvec(X[:, c], prec))


def _compare_solves(g, R, prec, tol, maxit):
    X, its, info = g.solve_batch(R, tol, maxit=maxit, precision=prec)
    for c in range(R.shape[1]):
        x, it, _, inf = g.solve(R[:, c], tol, maxit=maxit, precision=prec)
        np.testing.assert_array_equal(X[:, c], x)
        assert its[c] == it
        for k in ("iterations", "refinements", "matvecs64", "matvecs32", "final_residual"):
            assert info[c][k] == inf[k], (c, k, info[c][k], inf[k])
    return X, its, info


@pytest.mark.parametrize("prec", [fa.F64, fa.MIXED])
def test_solve_batch_equals_single(prec):
    qpos, Q = _charges(8, 6, 5)
    gq = fa.Problem(_ellipsoid(3), qpos, Q[:, 0], 4.0, 80.0, 0.124)
    R = gq.rhs_batch(Q)
    R[:, 2] = 0.0   # zero right-hand side: zero solution without iterating
    print("batched solve", prec, R.shape, flush=True)
    X, its, info = _compare_solves(gq, R, prec, 1e-6, 200)
    assert its[2] == 0 and not X[:, 2].any()


def test_solve_batch_iteration_cap():
    g = _problem("d7")
    R = np.random.default_rng(2).normal(size=(g.nvec, 4))
    R[:, 1] *= 1e-3
    print("batched solve, capped", R.shape, flush=True)
    _, its, info = _compare_solves(g, R, fa.F64, 1e-9, 6)
    assert all(i == 6 for i in its) and all(d["final_residual"] >= 1e-9 for d in info)


def test_solve_batch_two_groups():
    g = _problem("d4")
    R = np.random.default_rng(9).normal(size=(g.nvec, 66))
    print("batched solve, 66 columns", flush=True)
    X, its, info = g.solve_batch(R, 1e-6, precision=fa.MIXED)
    for c in (0, 37, 64, 65):
        x, it, _, inf = g.solve(R[:, c], 1e-6, precision=fa.MIXED)
        np.testing.assert_array_equal(X[:, c], x)
        assert its[c] == it and info[c]["matvecs32"] == inf["matvecs32"]
