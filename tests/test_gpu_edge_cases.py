"""Edge cases of the solve path on the GPU: zero charges (zero right-hand
side: zero solution without iterating, zero potentials), kappa below 1e-6
(the Yukawa integrals fall back to Laplace, FlatIntegration.c:1035-1049,
as in the oracle), an iteration cap below convergence (GMRES
non-convergence is silent, GMRES.c:446-539) and a charge exactly at a
panel centroid (the -4 pi self term of the direct potentials,
SurfaceOperator.c:382-383)."""
import numpy as np
import pytest

import fftsvd_amd as fa
import oracle
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu


def panels():
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.04, seed=7)
    return meshes.panels_from_mesh(V, F)


def test_zero_charges():
    P = panels()
    g = fa.Problem(P, [[0.1, 0.2, 0.3], [-0.2, 0.0, 0.1]], [0.0, 0.0], 4.0, 80.0, 0.124)
    r = g.rhs()
    assert not r.any()
    x, it, hist, info = g.solve(r, 1e-6, precision=fa.MIXED)
    assert it == 0 and not x.any()
    assert not g.reaction_potentials(x).any()


def test_tiny_kappa_is_laplace():
    P = panels()
    qpos, q = [[0.1, 0.2, 0.3]], [1.0]
    g = fa.Problem(P, qpos, q, 4.0, 80.0, 1e-8)
    o = oracle.Problem(P, qpos, q, 4.0, 80.0, 1e-8)
    gl = fa.Problem(P, qpos, q, 4.0, 80.0, 0.0)
    x, _, _, _ = g.solve(g.rhs(), 1e-8, precision=fa.F64)
    xo, _, _ = o.gmres(o.rhs(), 1e-8)
    xl, _, _, _ = gl.solve(gl.rhs(), 1e-8, precision=fa.F64)
    E, Eo, El = g.energy(x), o.energy(xo), gl.energy(xl)
    assert abs(E - Eo) <= 1e-7 * abs(Eo)
    assert abs(E - El) <= 1e-7 * abs(El)


def test_iteration_cap_is_silent():
    P = panels()
    g = fa.Problem(P, [[0.1, 0.2, 0.3]], [1.0], 4.0, 80.0, 0.124)
    x, it, hist, info = g.solve(g.rhs(), 1e-12, maxit=5, precision=fa.F64)
    assert it == 5 and np.isfinite(x).all() and info["final_residual"] > 1e-12


def test_charge_at_centroid():
    P = panels()
    c = oracle.panel_geom(P[17])["centroid"]
    qpos, q = np.array([c, [0.1, 0.0, -0.2]]), np.array([0.5, -0.3])
    g = fa.Problem(P, qpos, q, 4.0, 80.0, 0.124)
    o = oracle.Problem(P, qpos, q, 4.0, 80.0, 0.124)
    sol = np.random.default_rng(2).normal(size=2 * len(P))
    pg, po = g.reaction_potentials(sol), o.reaction_potentials(sol)
    assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()


def test_translation_autotuner():
    """With autotuning on, the first FP32 operator times the translation
    candidates and keeps the fastest; the mixed matvec stays within the
    FP32 rounding of the hand-tuned plan's (the candidates tile the same
    sums differently), the FP64 matvec is untouched, and turning autotuning
    off restores the defaults."""
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=4, jitter=0.04, seed=7)
    P = meshes.panels_from_mesh(V, F)
    q = [[0.2, -0.3, 0.1], [-0.5, 0.4, 0.2]], [0.7, -0.4]
    x = np.random.default_rng(8).uniform(-1, 1, 2 * len(P))
    base = fa.Problem(P, *q, 4.0, 80.0, 0.124)
    y32, y64 = base.matvec(x, fa.MIXED), base.matvec(x, fa.F64)
    assert fa.translation_params()["tuned"] is False
    try:
        fa.set_autotune(True)
        tuned = fa.Problem(P, *q, 4.0, 80.0, 0.124)
        tp = fa.translation_params()
        print("autotuned translation:", tp, flush=True)
        assert tp["tuned"] and tp["mode"] in (0, 1) and tp["budget"] in (880, 720) and tp["tile"] in (512, 256)
        assert tp["tuned_ms"] > 0
        assert np.linalg.norm(tuned.matvec(x, fa.MIXED) - y32) <= 1e-6 * np.linalg.norm(y32)
        assert np.array_equal(tuned.matvec(x, fa.F64), y64)
    finally:
        fa.set_autotune(False)
    assert fa.translation_params() == dict(mode=1, budget=880, tile=512, tuned=False, tuned_ms=0.0)
