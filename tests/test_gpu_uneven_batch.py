"""Uneven batch: several seeded random-walk molecules of different sizes,
each meshed by the library's SES mesher and solved on its own operator in
one process (the uneven-batch workload at test sizes: the recipe of the
globule benchmark with log-uniform atom counts).

Gates per molecule: the FP64 energy within 1e-7 of the oracle's (both
build their operator from the same mesh), the mixed energy within 1e-6 of
FP64, and every operator released before the next is built.  The
concurrent solve of a batch (fftsvd_problems_solve, larger molecules: 80
to 600 atoms on the globule benchmark's grid) gives every molecule exactly
its own single solve: solution bit for bit, iteration counts and
residuals equal, in both precisions."""
import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu


def test_uneven_batch_energies():
    rng = np.random.default_rng(20261202)
    counts = np.exp(rng.uniform(np.log(12), np.log(60), size=4)).astype(int)
    sizes = []
    for k, n in enumerate(counts):
        xyz, radii, q = meshes.synthetic_globule(natoms=int(n), seed=100 + k, confine=3.0 + 0.6 * n ** (1 / 3))
        V, F = meshes.ses_mesh(xyz, radii, probe=1.4, spacing=1.0)
        P = meshes.panels_from_mesh(V, F)
        sizes.append(len(P))
        g = fa.Problem(P, xyz, q, 4.0, 80.0, 0.124)
        s64, it64, _, _ = g.solve(g.rhs(), 1e-8, precision=fa.F64)
        s32, _, _, _ = g.solve(g.rhs(), 1e-8, precision=fa.MIXED)
        E64, E32 = g.energy(s64), g.energy(s32)
        o = oracle.Problem(P, xyz, q, 4.0, 80.0, 0.124)
        so, ito, _ = o.gmres(o.rhs(), 1e-8)
        Eo = o.energy(so)
        assert abs(E64 - Eo) <= 1e-7 * abs(Eo), (k, n, len(P))
        assert abs(E32 - E64) <= 1e-6 * abs(E64), (k, n, len(P))
        assert abs(it64 - ito) <= 1
        g.close()
    assert max(sizes) > 2 * min(sizes)   # uneven


def test_uneven_batch_concurrent_equals_single():
    rng = np.random.default_rng(7)
    counts = np.exp(rng.uniform(np.log(80), np.log(600), size=6)).astype(int)
    probs, rhs = [], []
    for k, n in enumerate(counts):
        xyz, radii, q = meshes.synthetic_globule(natoms=int(n), seed=300 + k,
                                                 confine=26.0 * (n / 2000.0) ** (1 / 3))
        V, F = meshes.ses_mesh(xyz, radii, probe=1.4, spacing=1.266)
        g = fa.Problem(meshes.panels_from_mesh(V, F), xyz, q, 4.0, 80.0, 0.124)
        probs.append(g)
        rhs.append(g.rhs())
    print("concurrent batch, panels per molecule:", [p.nvec // 2 for p in probs], flush=True)
    assert max(p.nvec for p in probs) > 3 * min(p.nvec for p in probs)
    for prec in (fa.F64, fa.MIXED):
        X, its, info = fa.solve_many(probs, rhs, 1e-6, precision=prec)
        for k, (g, r) in enumerate(zip(probs, rhs)):
            x, it, _, inf = g.solve(r, 1e-6, precision=prec)
            np.testing.assert_array_equal(X[k], x)
            assert its[k] == it
            for key in ("iterations", "refinements", "matvecs64", "matvecs32", "final_residual"):
                assert info[k][key] == inf[key], (k, key)
    for g in probs:
        g.close()
