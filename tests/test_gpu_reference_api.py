"""The reference's call sequence through the drop-in C interface
(FFTSVDpbe.c:158-238: readSRF -> generateSurfaceOperator ->
generateSurfaceOperatorPreconditioner -> generateRHS -> solveSurfaceOperator
-> collectPotentials), checked against the extended API on the same panels
and charges, and against the FP64 oracle's energy."""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
import refapi
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu

KAPPA = 0.124


def setup_case(tmp_path, kappa):
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    meshes.write_vertface(str(tmp_path / "blob"), V, F)
    V = np.loadtxt(tmp_path / "blob.vert")   # the vertices as written (9 decimals)
    refapi.write_srf(tmp_path / "op.srf", ["blob"])
    rng = np.random.default_rng(1)
    qpos, q = rng.uniform(-0.8, 0.8, size=(6, 3)), rng.uniform(-0.8, 0.8, size=6)
    pdb = (refapi.PDBentry * len(q))()
    for i in range(len(q)):
        pdb[i].x, pdb[i].y, pdb[i].z = qpos[i]
        pdb[i].charge = q[i]
        pdb[i].radius = 1.5
        pdb[i].chain = b"A"
    refapi.set_global("innerdielectric", 4.0)
    refapi.set_global("outerdielectric", 80.0)
    refapi.set_global("ionicstrength", (kappa * 3.047) ** 2)
    refapi.set_global("ionexclusionradius", 0.0)
    refapi.set_global("errortolerance", 1e-6)
    refapi.set_global("addcoulombic", 0, C.c_uint)
    return V, F, qpos, q, pdb


@pytest.mark.parametrize("kappa", [0.0, KAPPA])
def test_reference_sequence_matches_extended_api(tmp_path, kappa):
    V, F, qpos, q, pdb = setup_case(tmp_path, kappa)
    s = refapi.read_srf(tmp_path / "op.srf")
    op, rhs, sol = refapi.solve_path(s, pdb)
    n = s.total.value
    # the caller-visible structure FFTSVDpbe.c:54-57,75 walks
    root = op.contents
    assert root.numchildren == 1
    child = root.children[0].contents
    assert child.charges.contents.numcharges == len(q)
    np.testing.assert_allclose([child.charges.contents.charges[i] for i in range(len(q))], q, rtol=0, atol=0)
    assert child.mynumpanels == n
    pot = np.array([pdb[i].potential for i in range(len(q))])

    P = meshes.panels_from_mesh(V, F)
    assert len(P) == n
    prob = fa.Problem(P, qpos, q, 4.0, 80.0, kappa)
    b = prob.rhs()
    np.testing.assert_allclose(np.ctypeslib.as_array(rhs, shape=(2 * n,)), b, rtol=0, atol=1e-12 * np.abs(b).max())
    x, it, hist, info = prob.solve(b, 1e-6, precision=fa.F64)
    phi = prob.reaction_potentials(x)
    ref = 561.0 * phi / 4.0
    np.testing.assert_allclose(pot, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # energy against the FP64 oracle on the same panels
    o = oracle.Problem(P, qpos, q, 4.0, 80.0, kappa)
    xo, ito, _ = o.gmres(o.rhs(), 1e-6)
    E_lib = 0.5 * np.dot(q, pot) * 332.112 / 561.0
    E_orc = o.energy(xo)
    assert abs(E_lib - E_orc) <= 1e-6 * abs(E_orc)
    refapi.free(rhs)
    refapi.free(sol)


def test_reference_sequence_mixed_precision(tmp_path):
    V, F, qpos, q, pdb = setup_case(tmp_path, KAPPA)
    s = refapi.read_srf(tmp_path / "op.srf")
    fa.lib().fftsvd_set_precision.argtypes = [C.c_int]
    fa.lib().fftsvd_set_precision(1)
    try:
        op, rhs, sol = refapi.solve_path(s, pdb)
    finally:
        fa.lib().fftsvd_set_precision(0)
    pot32 = np.array([pdb[i].potential for i in range(len(q))])
    op64, rhs64, sol64 = refapi.solve_path(s, pdb)
    pot64 = np.array([pdb[i].potential for i in range(len(q))])
    E32, E64 = np.dot(q, pot32), np.dot(q, pot64)
    assert abs(E32 - E64) <= 1e-6 * abs(E64)


def test_addcoulombic_adds_direct_term(tmp_path):
    V, F, qpos, q, pdb = setup_case(tmp_path, KAPPA)
    s = refapi.read_srf(tmp_path / "op.srf")
    refapi.solve_path(s, pdb)
    reac = np.array([pdb[i].potential for i in range(len(q))])
    refapi.set_global("addcoulombic", 1, C.c_uint)
    try:
        refapi.solve_path(s, pdb)
    finally:
        refapi.set_global("addcoulombic", 0, C.c_uint)
    tot = np.array([pdb[i].potential for i in range(len(q))])
    d = np.linalg.norm(qpos[:, None, :] - qpos[None, :, :], axis=2)
    coul = np.array([sum(561.0 * q[a] / (4.0 * d[c, a]) for a in range(len(q)) if a != c) for c in range(len(q))])
    np.testing.assert_allclose(tot - reac, coul, rtol=1e-9, atol=1e-9)
