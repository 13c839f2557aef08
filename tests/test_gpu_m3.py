"""M3 trees (accelerateM3; FFTSVDpbeAPI.c:1256-1266,1503-1513): per region
with charges a Laplace tree from the region's panels to the charges, and
the reaction potentials through it (SurfaceOperator_calculateReactionPotentials,
M3 branch, SurfaceOperator.c:347-363).

Gates: the M3 tree (panels and charges partitioned together) and its ranks
equal the oracle's; with the GPU-built blocks loaded into the oracle's
tree, the potentials agree to 1e-10; against the direct sums the M3
potentials and energy agree to the FFT-SVD tolerance; the reference call
sequence with accelerateM3 gives the extended API's M3 potentials."""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
import refapi
from fftsvd_amd import meshes
from test_gpu_parity import _key

pytestmark = pytest.mark.gpu

KAPPA = 0.124


def blob():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    rng = np.random.default_rng(1)
    return V, F, rng.uniform(-0.8, 0.8, size=(40, 3)), rng.uniform(-0.8, 0.8, size=40)


@pytest.fixture(scope="module")
def m3_pair():
    V, F, qpos, q = blob()
    P = meshes.panels_from_mesh(V, F)
    refapi.set_global("accelerateM3", 1, C.c_uint)
    try:
        g = fa.Problem(P, qpos, q, 4.0, 80.0, KAPPA)
    finally:
        refapi.set_global("accelerateM3", 0, C.c_uint)
    o = oracle.Problem(P, qpos, q, 4.0, 80.0, KAPPA)
    return P, qpos, q, g, o


def _points(g, which, cube_id, nq):
    pt = np.zeros(max(nq, 1), dtype=np.int32)
    fa._check(fa.lib().fftsvd_tree_cube_points(g.h, which, cube_id, pt.ctypes.data_as(C.POINTER(C.c_int))))
    return pt[:nq]


def test_m3_tree_matches_oracle(m3_pair):
    P, qpos, q, g, o = m3_pair
    gt, ot = g.tree(-2), o.m3_tree(1)   # region 1 (the dielectric) holds the charges
    assert gt.depth == ot.depth
    gc = {_key(c): c for c in gt.cubes()}
    oc = {_key(c): c for c in ot.cubes()}
    assert set(gc) == set(oc)
    for k, c in oc.items():
        d = gc[k]
        assert sorted(d["pan"].tolist()) == sorted(c["pan"].tolist())
        assert sorted(_points(g, -2, d["id"], len(c["pt"])).tolist()) == sorted(c["pt"].tolist())
        assert d["rowrank"] == c["rowrank"] and d["colrank"] == c["colrank"]


def _import_m3(g, which, ot):
    """Load the GPU-built blocks of an M3 tree into the oracle's (panel and
    point orders mapped per cube)."""
    gt = g.tree(which)
    gpl = ot.gpl()
    gc = {_key(c): c for c in gt.cubes()}
    for c in ot.cubes():
        d = gc[_key(c)]
        gm = gt.cube_mats(d, int(gpl[c["level"]]) ** 3)
        ppos = {int(p): k for k, p in enumerate(d["pan"])}
        pperm = np.array([ppos[int(p)] for p in c["pan"]], dtype=int)
        gpts = _points(g, which, d["id"], len(c["pt"]))
        qpos_ = {int(p): k for k, p in enumerate(gpts)}
        qperm = np.array([qpos_[int(p)] for p in c["pt"]], dtype=int)
        leafD = c["leaf"] and len(c["pt"])
        ot.set_cube_mats(c["id"], d["rowrank"], d["colrank"], gm["Ds"] if leafD else None,
                         gm["Dd"] if leafD else None, gm["VT"][:, pperm] if d["rowrank"] else None,
                         gm["U"][qperm, :] if d["colrank"] else None, gm["PVs"] if d["rowrank"] else None,
                         gm["PVd"] if d["rowrank"] else None, gm["UTI"] if d["colrank"] else None)


def test_m3_potentials_identical_operator(m3_pair):
    P, qpos, q, g, o = m3_pair
    _import_m3(g, -2, o.m3_tree(1))
    sol = np.random.default_rng(6).normal(size=2 * len(P))
    po, pg = o.m3_potentials(sol), g.reaction_potentials(sol)
    assert np.abs(pg - po).max() <= 1e-10 * np.abs(po).max()


def test_m3_against_direct(m3_pair):
    P, qpos, q, g, o = m3_pair
    sol, _, _, _ = g.solve(g.rhs(), 1e-8, precision=fa.F64)
    pm = g.reaction_potentials(sol)
    direct = fa.Problem(P, qpos, q, 4.0, 80.0, KAPPA)
    pd = direct.reaction_potentials(sol)
    assert np.linalg.norm(pm - pd) <= 2e-3 * np.linalg.norm(pd)
    Em, Ed = 0.5 * np.dot(q, pm), 0.5 * np.dot(q, pd)
    assert abs(Em - Ed) <= 2e-3 * abs(Ed)
    # the batched potentials and the gradient stay on the direct sums
    np.testing.assert_array_equal(g.potentials_batch(sol[:, None])[:, 0], pd)


def test_m3_reference_sequence(tmp_path):
    V, F, qpos, q = blob()
    meshes.write_vertface(str(tmp_path / "blob"), V, F)
    Vr = np.loadtxt(tmp_path / "blob.vert")
    refapi.write_srf(tmp_path / "op.srf", ["blob"])
    pdb = (refapi.PDBentry * len(q))()
    for i in range(len(q)):
        pdb[i].x, pdb[i].y, pdb[i].z = qpos[i]
        pdb[i].charge = q[i]
        pdb[i].radius = 1.5
        pdb[i].chain = b"A"
    refapi.set_global("innerdielectric", 4.0)
    refapi.set_global("outerdielectric", 80.0)
    refapi.set_global("ionicstrength", (KAPPA * 3.047) ** 2)
    refapi.set_global("ionexclusionradius", 0.0)
    refapi.set_global("errortolerance", 1e-6)
    refapi.set_global("addcoulombic", 0, C.c_uint)
    refapi.set_global("accelerateM3", 1, C.c_uint)
    try:
        s = refapi.read_srf(tmp_path / "op.srf")
        op, rhs, sol = refapi.solve_path(s, pdb)
        assert bool(op.contents.children[0].contents.M3)
        P = meshes.panels_from_mesh(Vr, F)
        prob = fa.Problem(P, qpos, q, 4.0, 80.0, KAPPA)
    finally:
        refapi.set_global("accelerateM3", 0, C.c_uint)
    pot = np.array([pdb[i].potential for i in range(len(q))])
    x, _, _, _ = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    np.testing.assert_allclose(pot, 561.0 * prob.reaction_potentials(x) / 4.0, rtol=1e-6,
                               atol=1e-9 * np.abs(pot).max())
    refapi.free(rhs)
    refapi.free(sol)


def test_qual_m1m3():
    """Qualocation with accelerateM3: the M1M3 tree carries the right-hand
    side (transposed double layer) and the potentials (single layer)
    (QualocationOperator.c:33-70); identical operator against the oracle,
    and against the direct sums to the tree's accuracy (the reference
    itself calls the trees inaccurate here, SurfaceOperator.c:364)."""
    V, F, qpos, q = blob()
    P = meshes.panels_from_mesh(V, F)
    refapi.set_global("accelerateM3", 1, C.c_uint)
    try:
        g = fa.Qual(P, qpos, q, 4.0, 80.0)
    finally:
        refapi.set_global("accelerateM3", 0, C.c_uint)
    o = oracle.Qual(P, qpos, q, 4.0, 80.0)
    ot = o.use_m1m3()
    _import_m3(g, -1, ot)
    ro, rg = o.rhs(), g.rhs()
    assert np.abs(rg - ro).max() <= 1e-10 * np.abs(ro).max()
    sol = np.random.default_rng(3).normal(size=len(P))
    po, pg = o.potentials(sol), g.potentials(sol)
    assert np.abs(pg - po).max() <= 1e-10 * np.abs(po).max()
    direct = fa.Qual(P, qpos, q, 4.0, 80.0)
    rd, pd = direct.rhs(), direct.potentials(sol)
    assert np.linalg.norm(rg - rd) <= 3e-2 * np.linalg.norm(rd)
    assert np.linalg.norm(pg - pd) <= 2e-3 * np.linalg.norm(pd)
