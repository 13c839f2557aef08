"""Pins the oracle's cavity layouts (generateSurfaceOperator,
FFTSVDpbeAPI.c:1064-1730; SurfaceOperator_multiply, SurfaceOperator.c:114-158)
against the closed-form reaction energy of a charge among concentric
spheres (tests/spheres.py):

- the series itself reproduces the Born, Kirkwood (salt), Stern-sphere and
  off-centre Kirkwood closed forms;
- a dielectric shell around a solvent-filled cavity, without and with salt
  (the cavity region then takes the exterior Yukawa kernel), and the Stern
  layout with a dielectric cavity holding a salt cavity, each within the
  flat-panel discretisation error of the series, while the series without
  the cavity (or without salt in the salt cavity) lies well outside that
  error -- so a cavity dropped, given the wrong permittivity ratio or the
  wrong kernel fails;
- a cavity whose parent is 999 is dropped (readSRF, FFTSVDpbeAPI.c:1006-1029).
"""
import numpy as np
import pytest

import oracle
import spheres
from fftsvd_amd import meshes

EI, EO = 2.0, 80.0


def sphere(R, sub=3):
    V, F = meshes.icosphere(R, sub)
    return meshes.panels_from_mesh(V, F)


def test_series_closed_forms():
    a, b, k = 2.0, 3.0, 0.124
    born = 332.06 / (2 * a) * (1 / 80 - 1 / 4)
    assert spheres.reaction_energy([a], [4.0, 80.0], [0, 0], 0, 0.0) == pytest.approx(born, rel=1e-13)
    kirk = 332.06 * 0.5 * (1 / (80 * a * (1 + k * a)) - 1 / (4 * a))
    assert spheres.reaction_energy([a], [4.0, 80.0], [0, k], 0, 0.0) == pytest.approx(kirk, rel=1e-13)
    stern = 0.5 * 332.06 * (1 / (80 * a) - k / (80 * (1 + k * b)) - 1 / (4 * a))
    assert spheres.reaction_energy([a, b], [4.0, 80.0, 80.0], [0, 0, k], 0, 0.0) == pytest.approx(stern, rel=1e-12)
    # off-centre charge in one sphere: Kirkwood's Legendre series
    ei, eo, s = 4.0, 80.0, 1.2
    ser = sum((l + 1) * (ei - eo) / (ei * ((l + 1) * eo + l * ei)) * (s / a) ** (2 * l) for l in range(400)) / a
    assert spheres.reaction_potential([a], [ei, eo], [0, 0], 0, s) == pytest.approx(ser, rel=1e-12)
    # a cavity of the surrounding medium's own permittivity changes nothing
    e1 = spheres.reaction_potential([a], [ei, eo], [0, 0], 0, s)
    e2 = spheres.reaction_potential([0.5, a], [ei, ei, eo], [0, 0, 0], 1, s)
    assert e2 == pytest.approx(e1, rel=1e-12)


@pytest.mark.parametrize("kappa", [0.0, 0.124])
def test_dielectric_cavity_energy(kappa):
    """Dielectric shell (eps 2) between a solvent cavity of radius 1.2 and
    the surface of radius 3; charge at 1.9 from the centre."""
    Rc, Rd, s = 1.2, 3.0, 1.9
    Pd, Pc = sphere(Rd), sphere(Rc)
    pr = oracle.Problem(Pd, [[0, 0, s]], [1.0], EI, EO, kappa, cavities=[Pc])
    assert pr.n == len(Pd) + len(Pc)
    # regions in pre-order: exterior over the dielectric panels, the
    # dielectric over its panels and the cavity's, the cavity over its own
    assert [pr.tree(w).n for w in (0, 1, 2)] == [len(Pd), len(Pd) + len(Pc), len(Pc)]
    sol, it, _ = pr.gmres(pr.rhs(), 1e-6)
    E = pr.energy(sol, 332.06)
    Ea = spheres.reaction_energy([Rc, Rd], [EO, EI, EO], [kappa, 0, kappa], 1, s)
    E_nocav = spheres.reaction_energy([Rd], [EI, EO], [0, kappa], 0, s)
    E_lapcav = spheres.reaction_energy([Rc, Rd], [EO, EI, EO], [0, 0, kappa], 1, s)
    assert abs(E - Ea) < 0.005 * abs(Ea)
    assert abs(E_nocav - Ea) > 0.2 * abs(Ea)
    if kappa > 0:
        # the salt inside the cavity is seen: a Laplace cavity is 0.6 % off,
        # the discretisation error 0.15 %
        assert abs(E - E_lapcav) > abs(E - Ea) * 2


def test_stern_dielectric_and_salt_cavity_energy():
    """Stern layout, concentric: salt cavity (radius 1) inside a dielectric
    cavity (1.6) inside the dielectric (3, eps 2, charge at 2.2) inside the
    ion-exclusion surface (3.8); kappa 0.5 outside and in the salt cavity."""
    R, s, k = [1.0, 1.6, 3.0, 3.8], 2.2, 0.5
    Pl, Pc, Pd, Ps = (sphere(r) for r in R)
    pr = oracle.Problem(Pd, [[0, 0, s]], [1.0], EI, EO, k, stern_verts9=Ps, cavities=[Pc], salt_cavities=[Pl],
                        salt_cavity_parent=[0])
    n = [len(Pl), len(Pc), len(Pd), len(Ps)]
    assert pr.n == sum(n)
    # pre-order: exterior (salt panels), Stern layer (salt + dielectric),
    # dielectric (dielectric + cavity), cavity (cavity + salt cavity), salt cavity
    assert [pr.tree(w).n for w in range(5)] == [n[3], n[3] + n[2], n[2] + n[1], n[1] + n[0], n[0]]
    r = pr.rhs()
    # unknown blocks in pre-order: salt, dielectric, cavity, salt cavity;
    # the charge term sits on the dielectric phi and the cavity dphi rows
    o = np.cumsum([0, 2 * n[3], 2 * n[2], 2 * n[1], 2 * n[0]])
    assert np.all(r[o[0]:o[1]] == 0.0) and np.all(r[o[3]:o[4]] == 0.0)
    assert np.all(r[o[1]:o[1] + n[2]] > 0.0) and np.all(r[o[1] + n[2]:o[2]] == 0.0)
    assert np.all(r[o[2]:o[2] + n[1]] == 0.0) and np.all(r[o[2] + n[1]:o[3]] > 0.0)
    sol, it, _ = pr.gmres(r, 1e-6)
    E = pr.energy(sol, 332.06)
    Ea = spheres.reaction_energy(R, [EO, EO, EI, EO, EO], [k, 0, 0, 0, k], 2, s)
    E_nosalt = spheres.reaction_energy(R, [EO, EO, EI, EO, EO], [0, 0, 0, 0, k], 2, s)
    assert abs(E - Ea) < 0.003 * abs(Ea)
    assert abs(E_nosalt - Ea) > 0.02 * abs(Ea)


def test_cavity_marked_999_is_dropped():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=2, jitter=0.04, seed=7)
    Pd = meshes.panels_from_mesh(V, F)
    Pc = sphere(0.6, 1)
    qpos, q = [[0.9, 0.2, -0.1], [-0.7, 0.3, 0.4]], [0.5, -0.8]
    a = oracle.Problem(Pd, qpos, q, 4.0, 80.0, 0.124, cavities=[Pc], cavity_parent=[999])
    b = oracle.Problem(Pd, qpos, q, 4.0, 80.0, 0.124)
    assert a.n == b.n == len(Pd)
    x = np.random.default_rng(3).normal(size=2 * a.n)
    np.testing.assert_array_equal(a.matvec(x), b.matvec(x))
    np.testing.assert_array_equal(a.rhs(), b.rhs())
