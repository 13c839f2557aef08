"""Qualocation formulation on the GPU (generateQualocationOperator,
FFTSVDpbeAPI.c:1741-1839; QualocationOperator.c; GMRES.c:159-181,541-640)
and the GAMESS entry solveBEMproblem (FFTSVDpbeAPI.c:1974-2026).

Gates: identical trees and ranks to the oracle (bases sampled from the
double layer only); FP64 matvec of an identical operator <= 1e-10 relative;
RHS, preconditioner and potentials <= 1e-12; GMRES iterations within one
and energies within 1e-8 of the oracle; mixed energy within 1e-6 of FP64;
the analytic Born sphere; the reference call sequence with usequalocation
equal to the extended API; three grid points per side (solveBEMproblem's
setting) through both the transposed and the forward apply."""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
import refapi
from fftsvd_amd import meshes
from test_gpu_parity import _import_gpu_operator, _key

pytestmark = pytest.mark.gpu

EPS_IN, EPS_OUT = 4.0, 80.0


def blob_mesh():
    return meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)


@pytest.fixture(scope="module")
def qual_pair():
    V, F = blob_mesh()
    P = meshes.panels_from_mesh(V, F)
    rng = np.random.default_rng(1)
    qpos, q = rng.uniform(-0.8, 0.8, size=(6, 3)), rng.uniform(-0.8, 0.8, size=6)
    g = fa.Qual(P, qpos, q, EPS_IN, EPS_OUT)
    o = oracle.Qual(P, qpos, q, EPS_IN, EPS_OUT)
    return P, g, o


def _same_trees(gt, ot):
    assert gt.depth == ot.depth
    gc = {_key(c): c for c in gt.cubes()}
    oc = {_key(c): c for c in ot.cubes()}
    assert set(gc) == set(oc)
    for k, c in oc.items():
        d = gc[k]
        assert sorted(d["pan"].tolist()) == sorted(c["pan"].tolist())
        assert d["rowrank"] == c["rowrank"] and d["colrank"] == c["colrank"]


def test_qual_tree_matches_oracle(qual_pair):
    P, g, o = qual_pair
    assert g.n == o.n == len(P) and g.nvec == len(P)
    _same_trees(g.tree(0), o.tree(0))


def test_qual_matvec_identical_operator_fp64(qual_pair):
    P, g, o = qual_pair
    _import_gpu_operator(g, o, 0)
    x = np.random.default_rng(9).uniform(-1, 1, g.n)
    yo = o.matvec(x)
    yg = g.matvec(x, fa.F64)
    assert np.linalg.norm(yg - yo) <= 1e-10 * np.linalg.norm(yo)
    y32 = g.matvec(x, fa.MIXED)
    assert np.linalg.norm(y32 - yg) <= 1e-5 * np.linalg.norm(yg)


def test_qual_rhs_precond_potentials(qual_pair):
    P, g, o = qual_pair
    ro, rg = o.rhs(), g.rhs()
    assert np.abs(rg - ro).max() <= 1e-12 * np.abs(ro).max()
    b = np.random.default_rng(4).normal(size=g.n)
    po = b / o.precond_diag()
    pg = g.precond_solve(b)
    assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()
    sol = np.random.default_rng(6).normal(size=g.n)
    fo, fg = o.potentials(sol), g.potentials(sol)
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()


def test_qual_gmres_and_energy_match_oracle(qual_pair):
    P, g, o = qual_pair
    _import_gpu_operator(g, o, 0)
    rhs = o.rhs()
    so, ito, _ = o.gmres(rhs, 1e-6)
    sg, itg, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    assert abs(itg - ito) <= 1
    Eo, Eg = o.energy(so), g.energy(sg)
    assert abs(Eg - Eo) <= 1e-8 * abs(Eo)
    s32, _, _, info = g.solve(rhs, 1e-6, precision=fa.MIXED)
    assert abs(g.energy(s32) - Eg) <= 1e-6 * abs(Eg)
    # zero right-hand side: zero solution without iterating
    z, it0, _, _ = g.solve(np.zeros(g.n), 1e-6, precision=fa.F64)
    assert it0 == 0 and not z.any()


def test_qual_born_sphere_gpu():
    a = 2.0
    V, F = meshes.icosphere(a, 3)
    P = meshes.panels_from_mesh(V, F)
    g = fa.Qual(P, [[0, 0, 0]], [1.0], EPS_IN, EPS_OUT)
    rhs = g.rhs()
    s64, it, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    s32, _, _, _ = g.solve(rhs, 1e-6, precision=fa.MIXED)
    E64, E32 = g.energy(s64, 332.06), g.energy(s32, 332.06)
    Ea = 332.06 / (2 * a) * (1 / EPS_OUT - 1 / EPS_IN)
    assert abs(E64 - Ea) < 0.01 * abs(Ea)
    assert abs(E32 - E64) <= 1e-6 * abs(E64)
    assert it < 15


def test_qual_grid3_matches_oracle():
    """solveBEMproblem's tree: 27 panels per leaf, 3 grid points per side
    at the finest levels (cube kernels of 27 grid points)."""
    V, F = blob_mesh()
    P = meshes.panels_from_mesh(V, F)
    g = fa.Qual(P, np.zeros((0, 3)), np.zeros(0), EPS_IN, EPS_OUT, maxpanels=27, gridpoints=3)
    o = oracle.Qual(P, np.zeros((0, 3)), np.zeros(0), EPS_IN, EPS_OUT, maxpanels=27, gridpoints=3)
    assert o.tree(0).gpl()[-1] == 3
    _same_trees(g.tree(0), o.tree(0))
    _import_gpu_operator(g, o, 0)
    x = np.random.default_rng(3).uniform(-1, 1, g.n)
    yo, yg = o.matvec(x), g.matvec(x, fa.F64)
    assert np.linalg.norm(yg - yo) <= 1e-10 * np.linalg.norm(yo)


def test_surface_operator_grid3_matches_oracle():
    """The forward tree apply with 3 grid points per side (source and target
    cubes of 27 points take the block path)."""
    V, F = blob_mesh()
    P = meshes.panels_from_mesh(V, F)
    qpos, q = np.array([[0.1, 0.2, -0.3]]), np.array([1.0])
    g = fa.Problem(P, qpos, q, EPS_IN, EPS_OUT, 0.124, maxpanels=27, gridpoints=3)
    o = oracle.Problem(P, qpos, q, EPS_IN, EPS_OUT, 0.124, maxpanels=27, gridpoints=3)
    rng = np.random.default_rng(8)
    for which in (0, 1):
        _same_trees(g.tree(which), o.tree(which))
        _import_gpu_operator(g, o, which)
        xs, xd = rng.uniform(-1, 1, len(P)), rng.uniform(-1, 1, len(P))
        yo = o.tree(which).apply_both(xs, xd)
        for prec, tol in ((fa.F64, 1e-10), (fa.MIXED, 1e-5)):
            yg = g.tree(which).apply_both(xs, xd, prec)
            assert np.linalg.norm(yg - yo) <= tol * np.linalg.norm(yo)


def test_solve_bem_problem():
    """solveBEMproblem on panels given as 9 doubles: the same solution as the
    qualocation operator built with its settings, and the oracle's to the
    GMRES tolerance."""
    V, F = blob_mesh()
    P = meshes.panels_from_mesh(V, F)
    rhs = np.random.default_rng(11).normal(size=len(P))
    sol = refapi.solve_bem_problem(P, EPS_IN, EPS_OUT, rhs)
    g = fa.Qual(P, np.zeros((0, 3)), np.zeros(0), EPS_IN, EPS_OUT, maxpanels=27, gridpoints=3, want_f32=False)
    sg, _, _, _ = g.solve(rhs, 1e-5, precision=fa.F64)
    np.testing.assert_allclose(sol, sg, rtol=0, atol=1e-12 * np.abs(sg).max())
    o = oracle.Qual(P, np.zeros((0, 3)), np.zeros(0), EPS_IN, EPS_OUT, maxpanels=27, gridpoints=3)
    so, _, _ = o.gmres(rhs, 1e-5, early_exit=False)
    assert np.linalg.norm(sol - so) <= 1e-4 * np.linalg.norm(so)


def test_qual_reference_sequence(tmp_path):
    """readSRF with usequalocation, then the reference's qualocation call
    sequence: potentials equal the extended API's on the same panels."""
    V, F = blob_mesh()
    meshes.write_vertface(str(tmp_path / "blob"), V, F)
    Vr = np.loadtxt(tmp_path / "blob.vert")
    refapi.write_srf(tmp_path / "op.srf", ["blob"])
    rng = np.random.default_rng(1)
    qpos, q = rng.uniform(-0.8, 0.8, size=(6, 3)), rng.uniform(-0.8, 0.8, size=6)
    pdb = (refapi.PDBentry * len(q))()
    for i in range(len(q)):
        pdb[i].x, pdb[i].y, pdb[i].z = qpos[i]
        pdb[i].charge = q[i]
        pdb[i].radius = 1.5
        pdb[i].chain = b"A"
    refapi.set_global("innerdielectric", EPS_IN)
    refapi.set_global("outerdielectric", EPS_OUT)
    refapi.set_global("ionicstrength", 0.0)
    refapi.set_global("errortolerance", 1e-6)
    refapi.set_global("addcoulombic", 0, C.c_uint)
    refapi.set_global("usequalocation", 1, C.c_uint)
    try:
        s = refapi.read_srf(tmp_path / "op.srf")
        rhs, sol = refapi.qual_path(s, pdb)
    finally:
        refapi.set_global("usequalocation", 0, C.c_uint)
    n = s.total.value
    assert n == len(F)
    pot = np.array([pdb[i].potential for i in range(len(q))])
    P = meshes.panels_from_mesh(Vr, F)
    g = fa.Qual(P, qpos, q, EPS_IN, EPS_OUT)
    sg, _, _, _ = g.solve(g.rhs(), 1e-6, precision=fa.F64)
    phi = g.potentials(sg)
    # collectQualocationPotentials scales by 561/eps_in (kT/e)
    np.testing.assert_allclose(pot, 561.0 * phi / EPS_IN, rtol=1e-6, atol=1e-9 * np.abs(pot).max())
    refapi.free(rhs)
    refapi.free(sol)
