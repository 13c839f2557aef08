"""kappa gradients of the solvation energy on the GPU, and the torch
autograd wrapper over charges and kappa.

dE/dkappa = -1/2 lambda^T (dA/dkappa) x (A^T lambda = C^T q) applies the
derivative of every Yukawa region's compressed operator at fixed tree,
ranks and bases: near-field blocks of the reference scheme differentiated
term by term (the kernel pinned on the CPU against central differences of
the scheme, and here the GPU kernel against the oracle), and the
derivatives of the equivalent densities, translation spectra and
target-side projections.  The gradient is pinned by central differences
of the energy with the operator rebuilt at kappa +- h (Richardson over
h = 2e-3 and 1e-3, truncation O(h^4)) to 1e-5 relative, at an SVD
tolerance of 1e-8 so that re-sampled bases move the energy by less than
that: a salt dielectric, and the Stern layout with a salt cavity (two
Yukawa regions).  At the default tolerance 1e-5 the rebuilt operators'
bases and ranks shift with kappa and the central difference itself moves
by a few 1e-5; the gradient stays within 1e-4 of it."""
import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu

QPOS = np.array([[0.6, -0.3, 0.2], [-0.7, 0.4, -0.1], [0.1, 0.5, 0.4], [-0.2, -0.6, -0.3], [0.9, 0.1, -0.2]])
QVAL = np.array([0.7, -0.4, 0.55, -0.9, 0.3])


def blob():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    return meshes.panels_from_mesh(V, F)


def sphere(R, seed, sub=3):
    V, F = meshes.icosphere(R, sub)
    Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    Q[:, 0] *= np.sign(np.linalg.det(Q))
    return meshes.panels_from_mesh(V @ Q.T, F)


TIGHT = dict(svdeps=1e-8)


def energy_at(kappa, make, tol=1e-12):
    g = make(kappa)
    x, _, _, _ = g.solve(g.rhs(), tol, precision=fa.F64)
    return g.energy(x)


def fd_kappa(kappa, make):
    d = {}
    for h in (2e-3, 1e-3):
        d[h] = (energy_at(kappa + h, make) - energy_at(kappa - h, make)) / (2 * h)
    return (4 * d[1e-3] - d[2e-3]) / 3


def test_dkappa_integrals_match_oracle():
    P = blob()
    rng = np.random.default_rng(3)
    Pt = P[rng.choice(len(P), 50, replace=False)]
    cent = P.reshape(-1, 3, 3).mean(1)
    pts = np.vstack([cent[rng.choice(len(P), 80, replace=False)], Pt.reshape(-1, 3, 3).mean(1)[:6]])
    gs, gd = fa.integrate(Pt, pts, fa.YUKAWA_DKAPPA, 0.124)
    os_, od = oracle.yukawa_dk(Pt, pts, 0.124)
    assert np.abs(gs - os_).max() <= 1e-12 * np.abs(os_).max()
    assert np.abs(gd - od).max() <= 1e-12 * np.abs(od).max()


@pytest.mark.parametrize("svd,gate", [(1e-8, 1e-5), (1e-5, 1e-4)], ids=["svd1e-8", "svd1e-5"])
def test_kappa_gradient_salt(svd, gate):
    P = blob()
    kappa = 0.124
    make = lambda k: fa.Problem(P, QPOS, QVAL, 4.0, 80.0, k, svdeps=svd)
    g = make(kappa)
    x, _, _, _ = g.solve(g.rhs(), 1e-12, precision=fa.F64)
    grad = g.grad_kappa(x, 1e-12, precision=fa.F64)
    fd = fd_kappa(kappa, make)
    print("svdeps %g: kappa gradient %.10f, central differences %.10f" % (svd, grad, fd), flush=True)
    assert abs(grad - fd) <= gate * abs(fd)
    # mixed precision at the production tolerance
    x32, _, _, _ = g.solve(g.rhs(), 1e-6, precision=fa.MIXED)
    grad32 = g.grad_kappa(x32, 1e-6, precision=fa.MIXED)
    assert abs(grad32 - grad) <= 1e-5 * abs(grad)


def test_kappa_gradient_stern_with_salt_cavity():
    """Two Yukawa regions (outside the ion-exclusion surface and inside
    the salt cavity) both contribute."""
    R = [1.0, 1.6, 3.0, 3.8]
    Pl, Pc, Pd, Ps = (sphere(r, 40 + i) for i, r in enumerate(R))
    qpos, q = [[0.0, 0.0, 2.2], [0.3, -2.0, 0.5]], [1.0, -0.6]
    kw = dict(stern_verts9=Ps, cavities=[Pc], salt_cavities=[Pl], salt_cavity_parent=[0])
    make = lambda k: fa.Problem(Pd, qpos, q, 2.0, 80.0, k, **kw, **TIGHT)
    kappa = 0.5
    g = make(kappa)
    x, _, _, _ = g.solve(g.rhs(), 1e-12, precision=fa.F64)
    grad = g.grad_kappa(x, 1e-12, precision=fa.F64)
    fd = fd_kappa(kappa, make)
    print("kappa gradient %.10f, central differences %.10f" % (grad, fd), flush=True)
    assert abs(grad - fd) <= 1e-5 * abs(fd)


def test_kappa_gradient_is_zero_without_salt():
    P = blob()
    g = fa.Problem(P, QPOS, QVAL, 4.0, 80.0, 0.0)
    x, _, _, _ = g.solve(g.rhs(), 1e-10, precision=fa.F64)
    assert g.grad_kappa(x, 1e-10) == 0.0


def test_torch_autograd_wrapper():
    """torch.autograd through the C-ABI: d/dq and d/dkappa of a loss built
    on the energy, against central differences (the energy is quadratic in
    the charges; kappa as above)."""
    import torch
    from torch_energy import Surface, solvation_energy

    P = blob()
    surf = Surface(P, QPOS, tol=1e-12, **TIGHT)
    q = torch.tensor(QVAL, dtype=torch.float64, requires_grad=True)
    kappa = torch.tensor(0.124, dtype=torch.float64, requires_grad=True)
    E = solvation_energy(q, kappa, surf)
    loss = 3.0 * E + E * E
    loss.backward()
    e = float(E.detach())
    dloss = 3.0 + 2.0 * e
    # charges: central differences of the quadratic energy
    h = 0.05
    fdq = []
    for j in range(len(QVAL)):
        qp, qm = QVAL.copy(), QVAL.copy()
        qp[j] += h
        qm[j] -= h
        ep = float(solvation_energy(torch.tensor(qp), torch.tensor(0.124), surf))
        em = float(solvation_energy(torch.tensor(qm), torch.tensor(0.124), surf))
        fdq.append((ep - em) / (2 * h))
    np.testing.assert_allclose(q.grad.numpy(), dloss * np.array(fdq), rtol=1e-7, atol=0)
    fdk = fd_kappa(0.124, lambda k: fa.Problem(P, QPOS, QVAL, 4.0, 80.0, k, **TIGHT))
    assert abs(float(kappa.grad) - dloss * fdk) <= 1e-5 * abs(dloss * fdk)
