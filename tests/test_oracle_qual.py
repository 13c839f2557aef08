"""Pins the oracle's qualocation operator (QualocationOperator.c;
Solv_ecf_multiply_qual, GMRES.c:159-181; GMRES_QualocationOperator,
GMRES.c:541-640; Preconditioner_fill_diagonal_solv_ecf_qual,
Preconditioner.c:49-60; Charge_makerhs_ecf_qual, Charge.c:129-142):
the analytic Born sphere, agreement with the first-kind formulation on an
off-centre charge set, the compressed apply against the dense double-layer
operator (which fails if the bases are sampled from the wrong layer), the
preconditioner against its closed form (self double layer 2 pi), Gauss's
law for the right-hand side and GMRES against a dense solve."""
import numpy as np
import pytest

import oracle
from fftsvd_amd import meshes

EPS_IN, EPS_OUT = 4.0, 80.0
C0 = -EPS_OUT / ((EPS_OUT - EPS_IN) * EPS_IN)


@pytest.fixture(scope="module")
def blob():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    P = meshes.panels_from_mesh(V, F)
    rng = np.random.default_rng(1)
    qpos = rng.uniform(-0.8, 0.8, size=(6, 3))
    q = rng.uniform(-0.8, 0.8, size=6)
    return P, qpos, q, oracle.Qual(P, qpos, q, EPS_IN, EPS_OUT)


def _areas(P):
    v = P.reshape(-1, 3, 3)
    return 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)


def test_qual_born_energy():
    """332.06 q^2/(2a) (1/eps_out - 1/eps_in) = -19.716 kcal/mol at a = 2; the
    second-kind formulation converges in a few iterations."""
    a = 2.0
    V, F = meshes.icosphere(a, 3)
    P = meshes.panels_from_mesh(V, F)
    o = oracle.Qual(P, [[0, 0, 0]], [1.0], EPS_IN, EPS_OUT)
    sol, it, _ = o.gmres(o.rhs(), 1e-6)
    E = o.energy(sol, 332.06)
    Ea = 332.06 / (2 * a) * (1 / EPS_OUT - 1 / EPS_IN)
    assert abs(E - Ea) < 0.01 * abs(Ea)
    assert it < 15


def test_qual_matches_first_kind(blob):
    """Same mesh and charges, kappa = 0: the qualocation and the first-kind
    (SurfaceOperator) energies agree to the flat-panel discretisation error."""
    P, qpos, q, o = blob
    sol, it, _ = o.gmres(o.rhs(), 1e-8)
    Eq = o.energy(sol)
    pr = oracle.Problem(P, qpos, q, EPS_IN, EPS_OUT, 0.0)
    s1, _, _ = pr.gmres(pr.rhs(), 1e-8)
    E1 = pr.energy(s1)
    assert abs(Eq - E1) < 0.01 * abs(E1)


def test_qual_matvec_matches_dense(blob):
    """Compressed transposed double layer (bases sampled from the double layer
    only) against the dense operator, to the FFT-SVD tolerance."""
    P, qpos, q, o = blob
    n = o.n
    S, D = o.tree().dense(n)
    a = _areas(P)
    dg = np.diag(D).copy()
    x = np.random.default_rng(2).uniform(-1, 1, n)
    xa = a * x
    dt = D.T @ xa / (4 * np.pi * EPS_IN)
    ref = dt - dg * xa / (4 * np.pi * EPS_IN) + (C0 + dg / (4 * np.pi * EPS_IN)) * xa
    y = o.matvec(x)
    # FFT-SVD with 4^3 grids: ~4e-4 of the double-layer apply (the same as a
    # tree sampled from both layers)
    assert np.linalg.norm(y - ref) <= 1e-3 * np.linalg.norm(dt)
    # the same apply with the single layer in place of the double is far off
    bad = (S.T @ xa - np.diag(S) * xa) / (4 * np.pi * EPS_IN) + (C0 + dg / (4 * np.pi * EPS_IN)) * xa
    assert np.linalg.norm(bad - ref) > 0.1 * np.linalg.norm(dt)


def test_qual_precond_closed_form(blob):
    """Self double layer of a flat panel at its centroid is 2 pi
    (FlatIntegration.c:424-425), so p_i = (c0 + 2 pi/(4 pi eps_in)) area_i."""
    P, qpos, q, o = blob
    p = o.precond_diag()
    np.testing.assert_allclose(p, (C0 + 2 * np.pi / (4 * np.pi * EPS_IN)) * _areas(P), rtol=1e-12)


def test_qual_rhs_gauss_law():
    """rhs_i = -dlp(charge, panel i) q/(4 pi eps_in); the double layer of a
    closed surface seen from inside sums to a full solid angle, so
    |sum rhs| = q/eps_in for any charge inside (to the moment expansion
    the far panels use, FlatIntegration.c:321-337: a few 1e-7)."""
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    P = meshes.panels_from_mesh(V, F)
    for x, qv in (([0.1, -0.2, 0.3], 1.0), ([-0.9, 0.4, 0.2], -0.7)):
        o = oracle.Qual(P, [x], [qv], EPS_IN, EPS_OUT)
        assert abs(abs(o.rhs().sum()) - abs(qv) / EPS_IN) <= 1e-5 * abs(qv) / EPS_IN


def test_qual_gmres_matches_dense_solve():
    V, F = meshes.icosphere(1.5, 2)
    P = meshes.panels_from_mesh(V, F)
    o = oracle.Qual(P, [[0.2, -0.1, 0.3]], [1.0], 2.0, 80.0)
    n = o.n
    p = o.precond_diag()
    A = np.column_stack([o.matvec(e) for e in np.eye(n)])
    b = o.rhs()
    x_dense = np.linalg.solve(A, b)
    x, it, hist = o.gmres(b, tol=1e-12, maxit=200)
    assert hist[-1] < 1e-12
    np.testing.assert_allclose(x, x_dense, rtol=0, atol=1e-9 * np.abs(x_dense).max())
    # the preconditioned residual the solver reports is ||P^-1 (b - A x)|| / ||P^-1 b||
    assert np.linalg.norm((b - A @ x) / p) <= 1e-11 * np.linalg.norm(b / p)
    # zero right-hand side: zero solution, no iterations
    z, it0, _ = o.gmres(np.zeros(n), 1e-6)
    assert it0 == 0 and not z.any()
