"""Stern-layer formulation on the GPU (generateSurfaceOperator's Stern
branch, FFTSVDpbeAPI.c:1064-1377): a salt (ion-exclusion) surface around
the dielectric gives three regions -- exterior Yukawa over the salt panels,
the Laplace layer between the surfaces, the dielectric interior.

Gates as for the two-region operator: identical integer decisions and
near field in all three trees, FP64 matvec of an identical operator
<= 1e-10 relative, RHS / preconditioner / potentials <= 1e-12, GMRES
iterations within one and energies within 1e-8 relative of the FP64
oracle, mixed energy within 1e-6 of FP64; the analytic Stern sphere within
the flat-panel discretisation error; and the reference call sequence with
a salt surface in the .srf equal to the extended API."""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
import refapi
from fftsvd_amd import meshes
from test_gpu_parity import _import_gpu_operator, _key

pytestmark = pytest.mark.gpu

KAPPA = 0.124


def stern_meshes():
    Vd, Fd = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    Vs, Fs = meshes.ellipsoid((4.2, 3.4, 2.8), subdivisions=3, jitter=0.02, seed=8)
    return (Vd, Fd), (Vs, Fs)


@pytest.fixture(scope="module")
def stern_pair():
    (Vd, Fd), (Vs, Fs) = stern_meshes()
    Pd, Ps = meshes.panels_from_mesh(Vd, Fd), meshes.panels_from_mesh(Vs, Fs)
    rng = np.random.default_rng(1)
    qpos, q = rng.uniform(-0.8, 0.8, size=(6, 3)), rng.uniform(-0.8, 0.8, size=6)
    g = fa.Problem(Pd, qpos, q, 4.0, 80.0, KAPPA, stern_verts9=Ps)
    o = oracle.Problem(Pd, qpos, q, 4.0, 80.0, KAPPA, stern_verts9=Ps)
    return Pd, Ps, g, o


def test_stern_trees_match_oracle(stern_pair):
    Pd, Ps, g, o = stern_pair
    assert g.n == o.n == len(Pd) + len(Ps)
    # region trees in pre-order: salt (Yukawa), layer (Laplace, salt +
    # dielectric panels), interior (Laplace, dielectric panels)
    sizes = []
    for which in (0, 1, 2):
        gt, ot = g.tree(which), o.tree(which)
        assert gt.depth == ot.depth
        gc = {_key(c): c for c in gt.cubes()}
        oc = {_key(c): c for c in ot.cubes()}
        assert set(gc) == set(oc)
        for k, c in oc.items():
            d = gc[k]
            assert sorted(d["pan"].tolist()) == sorted(c["pan"].tolist())
            assert d["rowrank"] == c["rowrank"] and d["colrank"] == c["colrank"]
        sizes.append(gt.n)
    assert sizes == [len(Ps), len(Ps) + len(Pd), len(Pd)]


def test_stern_matvec_identical_operator_fp64(stern_pair):
    Pd, Ps, g, o = stern_pair
    for which in (0, 1, 2):
        _import_gpu_operator(g, o, which)
    x = np.random.default_rng(9).uniform(-1, 1, 2 * g.n)
    yo = o.matvec(x)
    yg = g.matvec(x, fa.F64)
    assert np.linalg.norm(yg - yo) <= 1e-10 * np.linalg.norm(yo)
    y32 = g.matvec(x, fa.MIXED)
    assert np.linalg.norm(y32 - yg) <= 1e-5 * np.linalg.norm(yg)


def test_stern_rhs_precond_potentials(stern_pair):
    Pd, Ps, g, o = stern_pair
    ro, rg = o.rhs(), g.rhs()
    assert np.abs(rg - ro).max() <= 1e-14 * np.abs(ro).max()
    assert np.all(rg[:2 * len(Ps)] == 0.0)   # no charge outside the dielectric
    b = np.random.default_rng(4).normal(size=2 * g.n)
    po, pg = o.precond_solve(b), g.precond_solve(b)
    assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()
    sol = np.random.default_rng(6).normal(size=2 * g.n)
    fo, fg = o.reaction_potentials(sol), g.reaction_potentials(sol)
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()


def test_stern_gmres_and_energy_match_oracle(stern_pair):
    Pd, Ps, g, o = stern_pair
    for which in (0, 1, 2):
        _import_gpu_operator(g, o, which)
    rhs = o.rhs()
    so, ito, ho = o.gmres(rhs, 1e-6)
    sg, itg, hg, info = g.solve(rhs, 1e-6, precision=fa.F64)
    assert abs(itg - ito) <= 1
    Eo, Eg = o.energy(so), g.energy(sg)
    assert abs(Eg - Eo) <= 1e-8 * abs(Eo)
    s32, it32, _, _ = g.solve(rhs, 1e-6, precision=fa.MIXED)
    assert abs(g.energy(s32) - Eg) <= 1e-6 * abs(Eg)


def test_stern_sphere_analytic_gpu():
    """Dielectric sphere a inside an ion-exclusion sphere b, central
    charge: reaction potential q/(eps_out a) - q k/(eps_out (1+kb)) -
    q/(eps_in a); FP64 within 1 %, mixed within 1e-6 of FP64."""
    a, b = 2.0, 3.0
    Vd, Fd = meshes.icosphere(a, 3)
    Vs, Fs = meshes.icosphere(b, 3)
    Pd, Ps = meshes.panels_from_mesh(Vd, Fd), meshes.panels_from_mesh(Vs, Fs)
    g = fa.Problem(Pd, [[0, 0, 0]], [1.0], 4.0, 80.0, KAPPA, stern_verts9=Ps)
    rhs = g.rhs()
    s64, _, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    s32, _, _, _ = g.solve(rhs, 1e-6, precision=fa.MIXED)
    E64, E32 = g.energy(s64, 332.06), g.energy(s32, 332.06)
    Ea = 0.5 * 332.06 * (1 / (80 * a) - KAPPA / (80 * (1 + KAPPA * b)) - 1 / (4 * a))
    assert abs(E64 - Ea) < 0.01 * abs(Ea)
    assert abs(E32 - E64) <= 1e-6 * abs(E64)


def test_stern_reference_sequence(tmp_path):
    """readSRF with a salt surface, ionexclusionradius > 0 and salt: the
    reference call sequence builds the Stern operator; its potentials equal
    the extended API's on the same panels."""
    (Vd, Fd), (Vs, Fs) = stern_meshes()
    meshes.write_vertface(str(tmp_path / "blob"), Vd, Fd)
    meshes.write_vertface(str(tmp_path / "salt"), Vs, Fs)
    Vd = np.loadtxt(tmp_path / "blob.vert")   # the vertices as written
    Vs = np.loadtxt(tmp_path / "salt.vert")
    refapi.write_srf(tmp_path / "op.srf", ["blob"], salt_names=["salt"])
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
    refapi.set_global("ionicstrength", (KAPPA * 3.047) ** 2)
    refapi.set_global("ionexclusionradius", 2.0)
    refapi.set_global("errortolerance", 1e-6)
    refapi.set_global("addcoulombic", 0, C.c_uint)
    try:
        s = refapi.read_srf(tmp_path / "op.srf")
        assert s.num[0].value == 1 and s.num[1].value == 1
        op, rhs, sol = refapi.solve_path(s, pdb)
    finally:
        refapi.set_global("ionexclusionradius", 0.0)
    n = s.total.value
    root = op.contents
    assert root.numchildren == 1 and root.mynumpanels == 0
    layer = root.children[0].contents
    assert layer.numchildren == 1 and layer.kernel == 4 and layer.mynumpanels == len(Fs)   # POISSON_KERNEL
    inner = layer.children[0].contents
    assert inner.charges.contents.numcharges == len(q) and inner.mynumpanels == len(Fd)
    pot = np.array([pdb[i].potential for i in range(len(q))])

    Pd, Ps = meshes.panels_from_mesh(Vd, Fd), meshes.panels_from_mesh(Vs, Fs)
    assert len(Pd) + len(Ps) == n
    prob = fa.Problem(Pd, qpos, q, 4.0, 80.0, KAPPA, stern_verts9=Ps)
    s64, _, _, _ = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    phi = prob.reaction_potentials(s64)
    # collectPotentials scales by 561/eps_in (kT/e)
    np.testing.assert_allclose(pot, 561.0 * phi / 4.0, rtol=1e-6, atol=1e-9 * np.abs(pot).max())
    refapi.free(rhs)
    refapi.free(sol)
