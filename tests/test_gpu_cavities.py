"""Cavities on the GPU (generateSurfaceOperator, FFTSVDpbeAPI.c:1064-1730;
SurfaceOperator_multiply, SurfaceOperator.c:114-158).

Salt without Stern: a dielectric with an off-centre solvent cavity gives
three regions -- exterior Yukawa over the dielectric panels, the dielectric
(Laplace, the charges) over its panels and the cavity's, the cavity
(Yukawa) over its own.  Stern: salt surface, dielectric, dielectric cavity
and salt cavity give five regions.  Against the FP64 oracle on the same
panels: identical integer decisions and ranks in every tree, FP64 matvec of
an identical operator <= 1e-10 relative, RHS / preconditioner / potentials
<= 1e-12, GMRES iterations within one and energies within 1e-8, mixed
energy within 1e-6 of FP64.  Concentric spheres against the closed-form
series within the flat-panel error; a cavity whose parent is 999 dropped
as the reference's readSRF does; and the reference call sequence with
cavities in the .srf equal to the extended API."""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
import refapi
import spheres
from fftsvd_amd import meshes
from test_gpu_parity import _import_gpu_operator, _key

pytestmark = pytest.mark.gpu

KAPPA = 0.124
QPOS = [[-1.5, 0.3, 0.2], [-0.8, -0.5, 0.1], [0.2, 0.6, -0.3], [-0.3, -0.2, 0.5], [0.2, -0.7, -0.2]]
QVAL = [0.7, -0.4, 0.55, -0.9, 0.3]


def sphere(R, sub=3, center=(0.0, 0.0, 0.0), seed=None):
    """Icosphere panels; with a seed, turned by a random rotation about the
    centre, so that concentric spheres share no symmetry plane with each
    other or with the octree (identical twin blocks put a basis column's
    norm at the rank cut in both builds)."""
    V, F = meshes.icosphere(R, sub)
    if seed is not None:
        Q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
        Q[:, 0] *= np.sign(np.linalg.det(Q))   # a rotation keeps the panels' orientation
        V = V @ Q.T
    return meshes.panels_from_mesh(V + np.asarray(center), F)


def blob_and_cavity():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.03, seed=11)
    return meshes.panels_from_mesh(V, F), sphere(0.6, 2, (1.2, 0.0, 0.0))


def check_trees(g, o, sizes):
    for which, n in enumerate(sizes):
        gt, ot = g.tree(which), o.tree(which)
        assert ot.n == n
        assert gt.depth == ot.depth
        gc = {_key(c): c for c in gt.cubes()}
        oc = {_key(c): c for c in ot.cubes()}
        assert set(gc) == set(oc)
        for k, c in oc.items():
            d = gc[k]
            assert sorted(d["pan"].tolist()) == sorted(c["pan"].tolist())
            assert d["rowrank"] == c["rowrank"] and d["colrank"] == c["colrank"]


def check_against_oracle(g, o, nreg):
    ro, rg = o.rhs(), g.rhs()
    assert np.abs(rg - ro).max() <= 1e-14 * np.abs(ro).max()
    b = np.random.default_rng(4).normal(size=2 * g.n)
    po, pg = o.precond_solve(b), g.precond_solve(b)
    assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()
    sol = np.random.default_rng(6).normal(size=2 * g.n)
    fo, fg = o.reaction_potentials(sol), g.reaction_potentials(sol)
    assert np.abs(fg - fo).max() <= 1e-12 * np.abs(fo).max()
    for which in range(nreg):
        _import_gpu_operator(g, o, which)
    x = np.random.default_rng(9).uniform(-1, 1, 2 * g.n)
    yo = o.matvec(x)
    yg = g.matvec(x, fa.F64)
    assert np.linalg.norm(yg - yo) <= 1e-10 * np.linalg.norm(yo)
    y32 = g.matvec(x, fa.MIXED)
    assert np.linalg.norm(y32 - yg) <= 1e-5 * np.linalg.norm(yg)
    so, ito, _ = o.gmres(ro, 1e-6)
    sg, itg, _, _ = g.solve(ro, 1e-6, precision=fa.F64)
    assert abs(itg - ito) <= 1
    Eo, Eg = o.energy(so), g.energy(sg)
    assert abs(Eg - Eo) <= 1e-8 * abs(Eo)
    s32, _, _, _ = g.solve(ro, 1e-6, precision=fa.MIXED)
    assert abs(g.energy(s32) - Eg) <= 1e-6 * abs(Eg)


def test_salt_dielectric_cavity_matches_oracle():
    Pd, Pc = blob_and_cavity()
    print("case: salt, dielectric %d panels + cavity %d panels" % (len(Pd), len(Pc)), flush=True)
    g = fa.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA, cavities=[Pc])
    o = oracle.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA, cavities=[Pc])
    assert g.n == o.n == len(Pd) + len(Pc)
    check_trees(g, o, [len(Pd), len(Pd) + len(Pc), len(Pc)])
    check_against_oracle(g, o, 3)


def test_stern_cavities_match_oracle_and_series():
    R, s, k = [1.0, 1.6, 3.0, 3.8], 2.2, 0.5
    Pl, Pc, Pd, Ps = (sphere(r, seed=20 + i) for i, r in enumerate(R))
    print("case: Stern, concentric spheres", R, flush=True)
    kw = dict(stern_verts9=Ps, cavities=[Pc], salt_cavities=[Pl], salt_cavity_parent=[0])
    g = fa.Problem(Pd, [[0, 0, s]], [1.0], 2.0, 80.0, k, **kw)
    o = oracle.Problem(Pd, [[0, 0, s]], [1.0], 2.0, 80.0, k, **kw)
    n = [len(x) for x in (Pl, Pc, Pd, Ps)]
    assert g.n == o.n == sum(n)
    check_trees(g, o, [n[3], n[3] + n[2], n[2] + n[1], n[1] + n[0], n[0]])
    # the series before the oracle's operator is replaced by the GPU's
    sg, _, _, _ = g.solve(g.rhs(), 1e-6, precision=fa.F64)
    Ea = spheres.reaction_energy(R, [80.0, 80.0, 2.0, 80.0, 80.0], [k, 0, 0, 0, k], 2, s)
    assert abs(g.energy(sg, 332.06) - Ea) < 0.003 * abs(Ea)
    check_against_oracle(g, o, 5)


def test_no_salt_cavity_series():
    Rc, Rd, s = 1.2, 3.0, 1.9
    g = fa.Problem(sphere(Rd, seed=30), [[0, 0, s]], [1.0], 2.0, 80.0, 0.0, cavities=[sphere(Rc, seed=31)])
    rhs = g.rhs()
    s64, _, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    s32, _, _, _ = g.solve(rhs, 1e-6, precision=fa.MIXED)
    E64, E32 = g.energy(s64, 332.06), g.energy(s32, 332.06)
    Ea = spheres.reaction_energy([Rc, Rd], [80.0, 2.0, 80.0], [0, 0, 0], 1, s)
    assert abs(E64 - Ea) < 0.005 * abs(Ea)
    assert abs(E32 - E64) <= 1e-6 * abs(E64)


def test_cavity_marked_999_is_dropped():
    Pd, Pc = blob_and_cavity()
    a = fa.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA, cavities=[Pc], cavity_parent=[999])
    b = fa.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA)
    assert a.n == b.n == len(Pd)
    x = np.random.default_rng(3).normal(size=2 * a.n)
    np.testing.assert_array_equal(a.rhs(), b.rhs())
    ya, yb = a.matvec(x, fa.F64), b.matvec(x, fa.F64)
    assert np.abs(ya - yb).max() <= 1e-14 * np.abs(yb).max()
    with pytest.raises(fa.FFTSVDError, match="cavity parent"):
        fa.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA, cavities=[Pc], cavity_parent=[3])


def test_reference_sequence_with_cavities(tmp_path):
    """readSRF with a dielectric cavity and a 999-marked one, salt without
    a Stern layer: the reference call sequence equals the extended API."""
    Vd, Fd = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.03, seed=11)
    Vc, Fc = meshes.icosphere(0.6, 2, (1.2, 0.0, 0.0))
    Vb, Fb = meshes.icosphere(0.3, 1, (-2.0, 0.0, 0.0))
    meshes.write_vertface(str(tmp_path / "blob"), Vd, Fd)
    # cavity faces as the mesher writes them: counter-clockwise seen from inside
    meshes.write_vertface(str(tmp_path / "cav"), Vc, Fc[:, [0, 2, 1]])
    meshes.write_vertface(str(tmp_path / "bad"), Vb, Fb[:, [0, 2, 1]])
    Vd, Vc = np.loadtxt(tmp_path / "blob.vert"), np.loadtxt(tmp_path / "cav.vert")
    refapi.write_srf(tmp_path / "op.srf", ["blob"], cavity_names=["cav", "bad"], cavity_parents=[0, 999])
    pdb = (refapi.PDBentry * len(QVAL))()
    for i in range(len(QVAL)):
        pdb[i].x, pdb[i].y, pdb[i].z = QPOS[i]
        pdb[i].charge = QVAL[i]
        pdb[i].radius = 1.5
        pdb[i].chain = b"A"
    refapi.set_global("innerdielectric", 4.0)
    refapi.set_global("outerdielectric", 80.0)
    refapi.set_global("ionicstrength", (KAPPA * 3.047) ** 2)
    refapi.set_global("ionexclusionradius", 0.0)
    refapi.set_global("errortolerance", 1e-6)
    refapi.set_global("addcoulombic", 0, C.c_uint)
    s = refapi.read_srf(tmp_path / "op.srf")
    assert s.num[1].value == 1 and s.num[2].value == 2
    op, rhs, sol = refapi.solve_path(s, pdb)
    root = op.contents
    assert root.numchildren == 1 and root.mynumpanels == 0
    diel = root.children[0].contents
    assert diel.numchildren == 1 and diel.mynumpanels == len(Fd)
    cav = diel.children[0].contents
    assert cav.mynumpanels == len(Fc) and cav.kernel == 5 and cav.epsilon == 80.0   # HELMHOLTZ_KERNEL
    pot = np.array([pdb[i].potential for i in range(len(QVAL))])
    Pd, Pc = meshes.panels_from_mesh(Vd, Fd), meshes.panels_from_mesh(Vc, Fc)
    assert s.total.value == len(Pd) + len(Pc)
    prob = fa.Problem(Pd, QPOS, QVAL, 4.0, 80.0, KAPPA, cavities=[Pc])
    s64, _, _, _ = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    phi = prob.reaction_potentials(s64)
    # collectPotentials scales by 561/eps_in (kT/e)
    np.testing.assert_allclose(pot, 561.0 * phi / 4.0, rtol=1e-6, atol=1e-9 * np.abs(pot).max())
    refapi.free(rhs)
    refapi.free(sol)


def test_driver_meshes_and_solves_a_cavity(tmp_path):
    """The drop-in driver on a shell of atoms around a void: generateFlatSRF
    writes the void as a dielectric cavity (parent 0), the solve runs with
    the cavity region, and the FRC potentials equal the extended API's on
    the same two surfaces (and differ from the outer surface alone)."""
    import subprocess

    from fftsvd_amd import build
    from test_frontend_host import hollow_shell, pdb_line
    from test_gpu_pbe_driver import CHARGES, NAMES, RADII

    xyz = np.round(hollow_shell(), 3)
    lines = [pdb_line("ATOM", i + 1, NAMES[i % 4], "ALA", "A", i // 4 + 1, *p) for i, p in enumerate(xyz)]
    for f in ("fort.13", "fort.15"):
        (tmp_path / f).write_text("".join(lines) + "END\n")
    (tmp_path / "fort.11").write_text("atom__res_radius_\n" +
                                      "".join("%-6s%-3s%8.3f\n" % (k, "", v) for k, v in RADII.items()))
    (tmp_path / "fort.12").write_text("atom__resnumbc_charge_\n" +
                                      "".join("%-6s%-3s%-4s%1s%8.3f\n" % (k, "", "", "", v) for k, v in CHARGES.items()))
    salt = 0.145
    (tmp_path / "fort.10").write_text("indi=4.0\nexdi=80.0\nsalt=%g\nionrad=0\nprbrad=1.4\nde=1e-6\ndvert=2.0\n"
                                      "svert=1.0\nenergy(g)\nout(frc)\n" % salt)
    r = subprocess.run([build.PBE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    srf = (tmp_path / "operator.srf").read_text().split("\n")
    assert srf[3:5] == ["operator.srf.d0", "operator.srf.c0"] and srf[7] == "0"
    frc = (tmp_path / "fort.16").read_text().splitlines()
    pot = np.array([float(l[35:50]) for l in frc[12:] if not l.startswith("total energy")])
    q = np.array([CHARGES[NAMES[i % 4]] for i in range(len(xyz))])

    def panels(tag, flip):
        V = np.loadtxt(tmp_path / ("operator.srf.%s.vert" % tag))
        F = np.loadtxt(tmp_path / ("operator.srf.%s.face" % tag), dtype=np.int64)[:, :3] - 1
        return meshes.panels_from_mesh(V, F, flip=flip)

    Pd, Pc = panels("d0", True), panels("c0", False)
    kappa = np.sqrt(salt) / 3.047
    prob = fa.Problem(Pd, xyz, q, 4.0, 80.0, kappa, cavities=[Pc])
    x, _, _, _ = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    ref = 561.0 * prob.reaction_potentials(x) / 4.0
    np.testing.assert_allclose(pot, ref, rtol=0, atol=2e-6 * np.abs(ref).max() + 1e-6)
    outer = fa.Problem(Pd, xyz, q, 4.0, 80.0, kappa)
    x0, _, _, _ = outer.solve(outer.rhs(), 1e-6, precision=fa.F64)
    ref0 = 561.0 * outer.reaction_potentials(x0) / 4.0
    assert np.abs(ref - ref0).max() > 1e-3 * np.abs(ref).max()
