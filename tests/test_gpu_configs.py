"""The benchmark workloads at their stated sizes.

Born ion: icosphere of radius 2 A subdivided 4 times (5120 flat panels,
2562 vertices), +1 e at the centre, eps 4/80, kappa 0, at most 32 panels
per leaf, GMRES tol 1e-6.  Gates: the FP64 energy within 1e-8 of the
oracle's (operators built independently from the same panels), GMRES
iterations within one, the analytic Born energy 332.06 q^2/(2a)
(1/eps_out - 1/eps_in) within the flat-panel discretisation error (0.5 %),
and the mixed energy within 1e-6 of FP64.

Synthetic globule: the 2000-atom seeded walk meshed to ~100k panels (the
throughput workload).  The oracle's FP64 energy on that mesh is a
committed fixture (tests/golden/make_globule100k.py: five minutes of single-core
CPU); the mesh checksum in the fixture guards that the mesher produced the
same panels.  Gates: FP64 GPU energy within 1e-8 of the oracle's, GMRES
iterations within one, mixed within 1e-6 of FP64; on the same operator, 8
of the batched workload's seeded charge vectors solved in lockstep to 1e-8:
every column's mixed energy within 1e-6 of FP64, and a column equal to its
single solve bit for bit.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_born_icosphere_5120():
    V, F = meshes.icosphere(2.0, 4)
    P = meshes.panels_from_mesh(V, F)
    assert len(P) == 5120 and len(V) == 2562
    qpos, q = [[0.0, 0.0, 0.0]], [1.0]
    g = fa.Problem(P, qpos, q, 4.0, 80.0, 0.0)
    o = oracle.Problem(P, qpos, q, 4.0, 80.0, 0.0)
    assert g.tree(0).depth == o.tree(0).depth
    rhs = g.rhs()
    s64, it64, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    s32, _, _, info = g.solve(rhs, 1e-6, precision=fa.MIXED)
    so, ito, _ = o.gmres(o.rhs(), 1e-6)
    E64, E32, Eo = g.energy(s64, 332.06), g.energy(s32, 332.06), o.energy(so, 332.06)
    Eborn = 332.06 * 1.0 / (2 * 2.0) * (1 / 80.0 - 1 / 4.0)
    print("born 5120: gpu %.8f (its %d) oracle %.8f (its %d) mixed %.8f analytic %.6f" %
          (E64, it64, Eo, ito, E32, Eborn))
    assert abs(it64 - ito) <= 1
    assert abs(E64 - Eo) <= 1e-8 * abs(Eo)
    assert abs(E64 - Eborn) <= 5e-3 * abs(Eborn)
    assert abs(E32 - E64) <= 1e-6 * abs(E64)
    g.close()


def _panel_digest(P):
    return hashlib.sha256(np.ascontiguousarray(P, dtype=np.float64).tobytes()).hexdigest()


def test_globule_100k_against_oracle_fixture():
    path = os.path.join(GOLDEN, "globule100k_oracle.json")
    with open(path) as f:
        ref = json.load(f)
    xyz, radii, q = meshes.synthetic_globule(natoms=ref["atoms"], seed=ref["seed"])
    V, F = meshes.ses_mesh(xyz, radii, probe=1.4, spacing=ref["spacing"])
    P = meshes.panels_from_mesh(V, F)
    assert len(P) == ref["panels"]
    assert _panel_digest(P) == ref["panel_sha256"], "the mesher produced other panels than the fixture's"
    g = fa.Problem(P, xyz, q, 4.0, 80.0, ref["kappa"])
    assert g.tree(0).depth == ref["depth"]
    rhs = g.rhs()
    s64, it64, _, _ = g.solve(rhs, 1e-6, precision=fa.F64)
    s32, it32, _, info = g.solve(rhs, 1e-6, precision=fa.MIXED)
    E64, E32 = g.energy(s64), g.energy(s32)
    print("globule %d panels: gpu fp64 %.10f (its %d) mixed %.10f (its %d) oracle %.10f (its %d)" %
          (len(P), E64, it64, E32, it32, ref["energy_fp64"], ref["iterations"]))
    assert abs(it64 - ref["iterations"]) <= 1
    assert abs(E64 - ref["energy_fp64"]) <= 1e-8 * abs(ref["energy_fp64"])
    assert abs(E32 - E64) <= 1e-6 * abs(E64)
    assert info["final_residual"] < 1e-6
    # the batched charge-vector workload on this mesh (8 of its 64 seeded
    # columns, GMRES to 1e-8): every column's mixed energy within 1e-6 of
    # FP64, and the lockstep solve equal to the single solve bit for bit
    Q = np.column_stack([np.random.default_rng(ref["seed"] + 1 + c).uniform(-0.8, 0.8, len(q)) for c in range(8)])
    R = g.rhs_batch(Q)
    X64, _, _ = g.solve_batch(R, 1e-8, precision=fa.F64)
    X32, its, _ = g.solve_batch(R, 1e-8, precision=fa.MIXED)
    Eb64 = 0.5 * np.sum(Q * g.potentials_batch(X64), axis=0)
    Eb32 = 0.5 * np.sum(Q * g.potentials_batch(X32), axis=0)
    print("batched columns: mixed vs fp64 max rel %.2e" % np.max(np.abs(Eb32 - Eb64) / np.abs(Eb64)), flush=True)
    assert np.all(np.abs(Eb32 - Eb64) <= 1e-6 * np.abs(Eb64))
    x5, it5, _, _ = g.solve(R[:, 5], 1e-8, precision=fa.MIXED)
    np.testing.assert_array_equal(X32[:, 5], x5)
    assert its[5] == it5
    g.close()
