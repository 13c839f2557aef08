"""Adjoint of the surface operator and charge gradients of the solvation
energy on the GPU.

The reference has no adjoint path: FFTSVDpbe.c's hessian() (:52-99) forms
derivatives column by column from forward solves.  Here A^T is the exact
transpose of the compressed region operator (k_split / tree apply /
k_join transposed; Tree_multiply_transpose, Tree.c:769-832, per layer),
pinned by the dot-product test, and dE/dq = 1/2 (phi + B^T lambda) with
A^T lambda = C^T q is pinned by central finite differences of the energy,
which is exactly quadratic in the charges, on the same operator."""
import numpy as np
import pytest

import fftsvd_amd as fa
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu


def blob_panels():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    return meshes.panels_from_mesh(V, F)


@pytest.fixture(scope="module")
def salt_blob():
    P = blob_panels()
    rng = np.random.default_rng(31)
    qpos = rng.uniform(-0.8, 0.8, size=(6, 3))
    q = rng.uniform(-0.8, 0.8, size=6)
    return P, qpos, q, fa.Problem(P, qpos, q, 4.0, 80.0, 0.124)


def _dot_test(g, prec, tol):
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, g.nvec)
    y = rng.uniform(-1, 1, g.nvec)
    Ax = g.matvec(x, prec)
    ATy = g.matvec_transpose(y, prec)
    assert abs(np.dot(y, Ax) - np.dot(ATy, x)) <= tol * np.linalg.norm(y) * np.linalg.norm(Ax)


def test_transpose_dot_product(salt_blob):
    P, qpos, q, g = salt_blob
    _dot_test(g, fa.F64, 1e-12)
    _dot_test(g, fa.MIXED, 1e-6)


def test_transpose_dot_product_stern():
    Vd, Fd = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    Vs, Fs = meshes.ellipsoid((4.2, 3.4, 2.8), subdivisions=3, jitter=0.02, seed=8)
    g = fa.Problem(meshes.panels_from_mesh(Vd, Fd), [[0.1, 0.0, 0.2]], [1.0], 4.0, 80.0, 0.124,
                   stern_verts9=meshes.panels_from_mesh(Vs, Fs))
    _dot_test(g, fa.F64, 1e-12)


def test_adjoint_solve_residual(salt_blob):
    P, qpos, q, g = salt_blob
    b = np.random.default_rng(3).normal(size=g.nvec)
    lam, it = g.solve_adjoint(b, 1e-10, precision=fa.F64)
    r = b - g.matvec_transpose(lam, fa.F64)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)
    lam32, _ = g.solve_adjoint(b, 1e-6, precision=fa.MIXED)
    assert np.linalg.norm(lam32 - lam) <= 1e-5 * np.linalg.norm(lam)


def test_charge_gradient_matches_finite_differences(salt_blob):
    """E(q) = 1/2 q^T C A^-1 B q is quadratic, so the central difference
    (E(q + h e_j) - E(q - h e_j)) / 2h is its exact derivative; the energies
    of the perturbed charge vectors come from the batched solve on the same
    operator."""
    P, qpos, q, g = salt_blob
    nq, h = len(q), 0.05
    Q = np.repeat(q[:, None], 2 * nq, axis=1)
    for j in range(nq):
        Q[j, 2 * j] += h
        Q[j, 2 * j + 1] -= h
    X, _, _ = g.solve_batch(g.rhs_batch(Q), 1e-11, precision=fa.F64)
    E = 0.5 * np.sum(Q * g.potentials_batch(X), axis=0) * 332.112 / 4.0
    fd = (E[0::2] - E[1::2]) / (2 * h)
    x, _, _, _ = g.solve(g.rhs(), 1e-11, precision=fa.F64)
    grad = g.grad_charges(x, 1e-11, precision=fa.F64)
    np.testing.assert_allclose(grad, fd, rtol=1e-7, atol=1e-9 * np.abs(fd).max())
    # the operator is not symmetric: the potentials alone are not the gradient
    phi = g.reaction_potentials(x) * 332.112 / 4.0
    assert np.abs(phi - grad).max() > 1e-6 * np.abs(grad).max()
    # mixed precision at the production tolerance
    x32, _, _, _ = g.solve(g.rhs(), 1e-6, precision=fa.MIXED)
    grad32 = g.grad_charges(x32, 1e-6, precision=fa.MIXED)
    assert np.linalg.norm(grad32 - fd) <= 1e-5 * np.linalg.norm(fd)
