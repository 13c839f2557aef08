"""Pins the oracle's compressed operator, region operator and solver:
tree partition and list completeness (Tree.c:12-178, Cube.c:352-713);
compressed apply vs the dense all-pairs operator to the SVD tolerance
(Tree.c:597-683); exact transpose by the dot-product test (Tree.c:769-832);
2x2 preconditioner vs its closed form (SurfaceOperator.c:170-197); GMRES
vs a dense solve (GMRES.c:446-539); analytic Born, Kirkwood and
Stern-layer spheres end to end; the .vert/.face reader (VertFace.c:26-202)."""
import numpy as np
import pytest

import oracle
from fftsvd_amd import meshes


@pytest.fixture(scope="module")
def blob():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    P = meshes.panels_from_mesh(V, F)
    rng = np.random.default_rng(1)
    qpos = rng.uniform(-0.8, 0.8, size=(6, 3))
    q = rng.uniform(-0.8, 0.8, size=6)
    return P, qpos, q


@pytest.fixture(scope="module")
def prob_salt(blob):
    P, qpos, q = blob
    return oracle.Problem(P, qpos, q, 4.0, 80.0, 0.124)


def test_tree_partition_and_lists(prob_salt):
    pr = prob_salt
    n = pr.n
    for which in (0, 1):
        t = pr.tree(which)
        cubes = t.cubes()
        leaves = [c for c in cubes if c["leaf"]]
        allp = np.concatenate([c["pan"] for c in leaves])
        assert sorted(allp.tolist()) == list(range(n))
        assert all(c["level"] == t.depth - 1 for c in leaves)
        within = sum(len(c["pan"]) <= 32 for c in leaves)
        assert within / len(leaves) > 0.9
        # near (D columns) + far (ancestors' interaction lists) cover each (point, panel) once
        byid = {c["id"]: c for c in cubes}
        leaf_of = {}
        for c in leaves:
            for p in c["pt"]:
                leaf_of[int(p)] = c
        for p in range(0, n, 37):
            cov = np.zeros(n, dtype=int)
            leaf = leaf_of[p]
            cov[leaf["pan"]] += 1
            for l in leaf["loc"]:
                cov[byid[l]["pan"]] += 1
            c = leaf
            while True:
                for s in c["itr"]:
                    cov[byid[s]["pan"]] += 1
                if c["parent"] < 0:
                    break
                c = byid[c["parent"]]
            assert (cov == 1).all()


def test_compressed_matches_dense(prob_salt):
    pr = prob_salt
    rng = np.random.default_rng(2)
    for which in (0, 1):
        t = pr.tree(which)
        S, D = t.dense(pr.n)
        xs = rng.uniform(-1, 1, pr.n)
        xd = rng.uniform(-1, 1, pr.n)
        y = t.apply_both(xs, xd)
        ref = S @ xs + D @ xd
        # FFT-SVD with 4^3 grids and 49 equivalent sources: ~3e-4 overall (far field ~1e-3)
        assert np.linalg.norm(y - ref) <= 1e-3 * np.linalg.norm(ref)


def test_transpose_dot_product(prob_salt):
    pr = prob_salt
    rng = np.random.default_rng(3)
    t = pr.tree(0)
    for layer in (0, 1):
        x = rng.uniform(-1, 1, pr.n)
        y = rng.uniform(-1, 1, pr.n)
        z = np.zeros(pr.n)
        Ax = t.apply_both(x, z) if layer == 0 else t.apply_both(z, x)
        ATy = t.apply_transpose(layer, y)
        assert abs(np.dot(Ax, y) - np.dot(x, ATy)) <= 1e-12 * np.linalg.norm(Ax) * np.linalg.norm(y)


def test_preconditioner_closed_form(prob_salt, blob):
    pr = prob_salt
    t = pr.tree(1)
    ds = np.zeros(pr.n)
    dd = np.zeros(pr.n)
    for c in t.cubes():
        if c["leaf"]:
            m = t.cube_mats(c)
            for k, p in enumerate(c["pan"]):
                ds[p] = m["Ds"][k, k]
                dd[p] = m["Dd"][k, k]
    b = np.random.default_rng(4).normal(size=2 * pr.n)
    x = pr.precond_solve(b)
    n = pr.n
    rel = 4.0 / 80.0
    for i in range(0, n, 17):
        M = np.array([[dd[i], -ds[i]], [dd[i], rel * ds[i]]])
        xi = np.linalg.solve(M, [b[i], b[n + i]])
        np.testing.assert_allclose([x[i], x[n + i]], xi, rtol=1e-13)


def test_gmres_matches_dense_solve():
    V, F = meshes.icosphere(1.5, 2)
    P = meshes.panels_from_mesh(V, F)
    pr = oracle.Problem(P, [[0.2, -0.1, 0.3]], [1.0], 2.0, 80.0, 0.3)
    n2 = 2 * pr.n
    A = np.column_stack([pr.matvec(e) for e in np.eye(n2)])
    b = pr.rhs()
    x_dense = np.linalg.solve(A, b)
    x, it, hist = pr.gmres(b, tol=1e-12, maxit=400)
    assert hist[-1] < 1e-12
    np.testing.assert_allclose(x, x_dense, rtol=0, atol=1e-9 * np.abs(x_dense).max())


def test_born_energy():
    """Analytic Born: 332.06 q^2/(2a) (1/eps_out - 1/eps_in) = -19.716 kcal/mol;
    flat-panel discretisation error at 1280 panels is a few tenths of a percent."""
    V, F = meshes.icosphere(2.0, 3)
    P = meshes.panels_from_mesh(V, F)
    pr = oracle.Problem(P, [[0, 0, 0]], [1.0], 4.0, 80.0, 0.0)
    sol, it, _ = pr.gmres(pr.rhs(), 1e-6)
    E = pr.energy(sol, 332.06)
    Ea = 332.06 / (2 * 2.0) * (1 / 80 - 1 / 4)
    assert abs(E - Ea) < 0.01 * abs(Ea)
    assert it < 40


def test_kirkwood_energy_salt():
    """Debye-Hueckel sphere, central charge: 332.06 q^2/2 [1/(eps_out a (1+ka)) - 1/(eps_in a)]."""
    V, F = meshes.icosphere(2.0, 3)
    P = meshes.panels_from_mesh(V, F)
    k = 0.124
    pr = oracle.Problem(P, [[0, 0, 0]], [1.0], 4.0, 80.0, k)
    sol, it, _ = pr.gmres(pr.rhs(), 1e-6)
    E = pr.energy(sol, 332.06)
    Ea = 332.06 * 0.5 * (1 / (80 * 2.0 * (1 + k * 2.0)) - 1 / (4 * 2.0))
    assert abs(E - Ea) < 0.01 * abs(Ea)


def test_stern_sphere_energy():
    """Stern layer (FFTSVDpbeAPI.c:1064-1377): dielectric sphere a inside an
    ion-exclusion sphere b, central charge.  Potential A/r + B between the
    spheres, C exp(-kr)/r outside, continuity of phi and eps dphi/dr give the
    reaction potential q/(eps_out a) - q k/(eps_out (1+kb)) - q/(eps_in a).
    The energy is pinned to the flat-panel discretisation error, and the
    Stern shift (against the same mesh without the layer) to 2 %."""
    a, b, k = 2.0, 3.0, 0.124
    Vd, Fd = meshes.icosphere(a, 3)
    Vs, Fs = meshes.icosphere(b, 3)
    Pd, Ps = meshes.panels_from_mesh(Vd, Fd), meshes.panels_from_mesh(Vs, Fs)
    pr = oracle.Problem(Pd, [[0, 0, 0]], [1.0], 4.0, 80.0, k, stern_verts9=Ps)
    assert pr.n == len(Pd) + len(Ps)
    sol, it, _ = pr.gmres(pr.rhs(), 1e-6)
    E = pr.energy(sol, 332.06)
    Ea = 0.5 * 332.06 * (1 / (80 * a) - k / (80 * (1 + k * b)) - 1 / (4 * a))
    assert abs(E - Ea) < 0.01 * abs(Ea)
    assert it < 40
    # the unknowns of the salt surface come first: the RHS has no charge
    # terms on the salt phi rows, only on the dielectric's
    r = pr.rhs()
    assert np.all(r[:len(Ps)] == 0.0) and np.all(r[2 * len(Ps):2 * len(Ps) + len(Pd)] > 0.0)
    p0 = oracle.Problem(Pd, [[0, 0, 0]], [1.0], 4.0, 80.0, k)
    s0, _, _ = p0.gmres(p0.rhs(), 1e-6)
    shift = E - p0.energy(s0, 332.06)
    shift_a = Ea - 0.5 * 332.06 * (1 / (80 * a * (1 + k * a)) - 1 / (4 * a))
    assert abs(shift - shift_a) < 0.02 * abs(shift_a)


def test_vertface_reader(tmp_path):
    V, F = meshes.icosphere(1.0, 1)
    F2 = np.vstack([F, [[0, 0, 1]]])  # a zero-area face is dropped
    g = np.ones(len(F2), dtype=int)
    g[3] = 2
    base = str(tmp_path / "s")
    meshes.write_vertface(base, V, F2, g)
    P = oracle.read_vertface(base + ".vert", base + ".face", flip=True)
    assert P.shape == (len(F), 9)
    ref = meshes.panels_from_mesh(np.round(V, 9), F, flip=True)
    np.testing.assert_allclose(P, ref, atol=1e-9)
    Pn = oracle.read_vertface(base + ".vert", base + ".face", flip=False)
    np.testing.assert_allclose(Pn[:, 3:6], P[:, 6:9])
    Pa = oracle.read_vertface(base + ".vert", base + ".face", flip=True, accessible=True)
    assert len(Pa) == len(F) - 1
