"""Pins the oracle's panel integrals to mathematics the reference ships no
golden vectors for (the reference ships no tests):

* Laplace single/double layer (FlatIntegration.c:295-429) vs an independent
  Duffy-transformed brute-force quadrature, and the double layer vs the
  Van Oosterom-Strackee solid angle; the far-field moment branch within its
  truncation order; the 2*pi self term.
* Yukawa (calcp, FlatIntegration.c:473-949): the edge/radial decomposition
  with a 16-point angular rule equals brute force to 1e-11 (pins geometry,
  signs and integrands); with the reference's 4-point rule it stays within
  that rule's quadrature error; kappa -> 0 continuity to Laplace.
* The 4-point rule equals Golub-Welsch (QuadratureRule.c:79-125).
* The radial line's angular pieces (calcp_LogIntDoLine,
  FlatIntegration.c:489-501,615-622): the 4-point line integral equals the
  4-point rule on piece lists worked out by hand from that text, per case
  to 1e-13, for every branch of the schedule.
The split of an edge at the foot of the perpendicular (calcp_LogInt,
FlatIntegration.c:637-799) is pinned only through the panel totals above.
"""
import numpy as np
import pytest

import oracle
import quad
from fftsvd_amd import meshes


def random_tri(rng, scale=1.0):
    while True:
        t = rng.normal(size=(3, 3)) * scale
        a = np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) / 2
        e = [np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3)]
        if a > 0.15 * max(e) ** 2:
            return t


def normal_L(t):
    n = np.cross(t[2] - t[0], t[1] - t[0])
    return n / np.linalg.norm(n)


def solid_angle(x, t):
    R = t - x
    r = np.linalg.norm(R, axis=1)
    num = np.dot(R[0], np.cross(R[1], R[2]))
    den = r[0] * r[1] * r[2] + np.dot(R[0], R[1]) * r[2] + np.dot(R[0], R[2]) * r[1] + np.dot(R[1], R[2]) * r[0]
    return 2.0 * np.arctan2(num, den)


def test_gl4_is_golub_welsch():
    n = 4
    J = np.zeros((n, n))
    for i in range(1, n):
        J[i, i - 1] = J[i - 1, i] = 0.5 * (1 - (2.0 * i) ** -2) ** -0.5
    w, v = np.linalg.eigh(J)
    x, wt = oracle.gl4()
    np.testing.assert_allclose(x, w / 2 + 0.5, rtol=0, atol=1e-15)
    np.testing.assert_allclose(wt, v[0] ** 2, rtol=0, atol=1e-15)


def test_laplace_near_matches_bruteforce_and_solid_angle():
    rng = np.random.default_rng(11)
    for _ in range(12):
        t = random_tri(rng)
        c = t.mean(0)
        diag = max(np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3))
        for r in (0.03, 0.3, 1.0, 2.5):
            d = rng.normal(size=3)
            x = c + r * diag * d / np.linalg.norm(d)
            s, dl = oracle.integrate(t.reshape(1, 9), x[None])
            bs, bd = quad.layer_integrals(x, t, 0.0, n=200)
            tol = 1e-9 if r < 0.1 else 1e-11
            assert abs(s[0, 0] - bs) <= tol * abs(bs)
            # the brute-force rule is not converged for the 1/r^3 kernel very near the panel
            assert abs(dl[0, 0] - bd) <= (1e-8 if r < 0.1 else 1e-11) * max(abs(bd), 1e-3)
            om = solid_angle(x, t)
            assert abs(abs(dl[0, 0]) - abs(om)) <= 1e-12 * max(abs(om), 1e-3)
            assert np.sign(dl[0, 0]) == np.sign(np.dot(x - c, normal_L(t)))


def test_laplace_far_moments_within_truncation_order():
    rng = np.random.default_rng(5)
    for _ in range(10):
        t = random_tri(rng)
        c = t.mean(0)
        diag = max(np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3))
        for r, tol in ((3.2, 3e-4), (5.5, 3e-5), (6.5, 2e-3), (12.0, 3e-4), (40.0, 1e-5)):
            d = rng.normal(size=3)
            x = c + r * diag * d / np.linalg.norm(d)
            s, dl = oracle.integrate(t.reshape(1, 9), x[None])
            bs, bd = quad.layer_integrals(x, t, 0.0, n=160)
            assert abs(s[0, 0] - bs) <= tol * abs(bs)
            # double layer: error relative to the monopole scale area*|z|/r^3 ~ area/r^2
            area = 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
            assert abs(dl[0, 0] - bd) <= 5 * tol * area / (r * diag) ** 2


def test_laplace_self_terms():
    rng = np.random.default_rng(3)
    for _ in range(10):
        t = random_tri(rng)
        # the collocation point is the panel's own centroid, bit for bit
        # (the 2*pi self term keys on exact equality, FlatIntegration.c:424-425)
        c = oracle.panel_geom(t.reshape(9))["centroid"]
        s, dl = oracle.integrate(t.reshape(1, 9), c[None])
        assert dl[0, 0] == 2 * np.pi
        bs, _ = quad.layer_integrals(c, t, 0.0, n=96)
        assert abs(s[0, 0] - bs) <= 1e-10 * bs
        # in-plane point outside the panel: double layer exactly 0 up to rounding
        x = c + 3.0 * (t[0] - c)
        s2, d2 = oracle.integrate(t.reshape(1, 9), x[None])
        assert abs(d2[0, 0]) < 1e-12


def test_moments_are_polynomial_integrals():
    rng = np.random.default_rng(9)
    t = random_tri(rng)
    g = oracle.panel_geom(t.reshape(9))
    ax = g["ax"]
    P = (t - g["centroid"]) @ ax.T
    u, w = np.polynomial.legendre.leggauss(20)
    u = 0.5 * (u + 1)
    w = 0.5 * w
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(w, w) * U
    pts = P[0][None, None, :] + U[..., None] * (P[1] - P[0]) + (U * V)[..., None] * (P[2] - P[1])
    J = abs(np.cross(P[1] - P[0], P[2] - P[1])[2])
    X, Y = pts[..., 0], pts[..., 1]
    CS = [0.0, 1.0, 1.0, 1.5, 1.5, 3.75, 1.0, 3.0, 1.5, 7.5, 1.5, 1.5, 3.75, 1.5, 7.5, 3.75]
    k = 0
    for m in range(5):
        for n in range(5 - m):
            k += 1
            exact = np.sum(W * X ** m * Y ** n) * J
            if m + n == 1:
                exact = 0.0  # about the centroid
            assert abs(g["moments"][k - 1] - exact * CS[k]) <= 1e-12 * max(1.0, abs(exact))


@pytest.mark.parametrize("kappa", [0.124, 1.0])
def test_yukawa_decomposition_exact_with_fine_rule(kappa):
    rng = np.random.default_rng(21)
    x16, w16 = np.polynomial.legendre.leggauss(16)
    x16 = 0.5 * (x16 + 1)
    w16 = 0.5 * w16
    for _ in range(10):
        t = random_tri(rng)
        c = t.mean(0)
        diag = max(np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3))
        for r in (0.02, 0.3, 1.0, 3.0):
            d = rng.normal(size=3)
            x = c + r * diag * d / np.linalg.norm(d)
            s, dl = oracle.yukawa_rule(x, t.reshape(9), kappa, x16, w16)
            bs, bd = quad.layer_integrals(x, t, kappa, n=200)
            tol = 1e-9 if r < 0.1 else 1e-11
            assert abs(s - bs) <= tol * abs(bs)
            assert abs(dl - bd) <= 10 * tol * max(abs(bd), 1e-4)
            # the reference's 4-point rule: within its own quadrature error
            s4, d4 = oracle.integrate(t.reshape(1, 9), x[None], oracle.YUKAWA, kappa)
            assert abs(s4[0, 0] - bs) <= 5e-3 * abs(bs)
            assert abs(d4[0, 0] - bd) <= 5e-3 * max(abs(bd), 1e-2)


def test_yukawa_on_panel_and_kappa_continuity():
    rng = np.random.default_rng(4)
    for _ in range(8):
        t = random_tri(rng)
        c = oracle.panel_geom(t.reshape(9))["centroid"]
        s, dl = oracle.integrate(t.reshape(1, 9), c[None], oracle.YUKAWA, 0.5)
        assert dl[0, 0] == 2 * np.pi
        x = c + 2.5 * (t[1] - c)
        s, dl = oracle.integrate(t.reshape(1, 9), x[None], oracle.YUKAWA, 0.5)
        assert dl[0, 0] == 0.0
        # kappa -> 0: Yukawa (numerical) approaches Laplace (analytic), same sign convention
        d = rng.normal(size=3)
        diag = max(np.linalg.norm(t[i] - t[(i + 1) % 3]) for i in range(3))
        x = c + 0.7 * diag * d / np.linalg.norm(d)
        sy, dy = oracle.integrate(t.reshape(1, 9), x[None], oracle.YUKAWA, 2e-6)
        sl, dlap = oracle.integrate(t.reshape(1, 9), x[None], oracle.LAPLACE, 0.0)
        assert abs(sy[0, 0] - sl[0, 0]) <= 5e-3 * abs(sl[0, 0])
        assert abs(dy[0, 0] - dlap[0, 0]) <= 5e-3 * max(abs(dlap[0, 0]), 1e-2)
        assert np.sign(dy[0, 0]) == np.sign(dlap[0, 0])


# calcp_LogIntDoLine (FlatIntegration.c:489-501, 615-622) splits the angular
# range of a radial line at distance x, from y1 to y2, into pieces: the first
# ends at min(10 y1, y2) when x < y1, else at min(x, y2); each next piece
# ends at ten times the previous end, or at y2 once that passes y2.  Piece
# lists worked out by hand from that text, one per branch of the schedule:
LINE_SCHEDULES = [
    # (x, y1, y2, pieces)
    (0.05, 0.02, 3.0, [(0.02, 0.05), (0.05, 0.5), (0.5, 3.0)]),       # x >= y1, growth then clip
    (0.01, 0.02, 3.0, [(0.02, 0.2), (0.2, 2.0), (2.0, 3.0)]),         # x < y1, y2 > 10 y1
    (0.01, 0.5, 3.0, [(0.5, 3.0)]),                                  # x < y1, y2 <= 10 y1
    (5.0, 0.1, 3.0, [(0.1, 3.0)]),                                   # x beyond y2
    (0.3, 0.0, 0.2, [(0.0, 0.2)]),                                   # from the foot, x beyond y2
    (0.001, 0.0, 2.0, [(0.0, 0.001), (0.001, 0.01), (0.01, 0.1), (0.1, 1.0), (1.0, 2.0)]),   # four growth steps
]


def _line_by_pieces(z, x, pieces, kappa):
    """The line integrand of calcp_LogIntDoLine, HELMHOLTZ kernel, integrated
    with the 4-point Gauss-Legendre rule on each given piece."""
    g, w = np.polynomial.legendre.leggauss(4)
    g, w = 0.5 * (g + 1), 0.5 * w
    az, sz = abs(z), np.sign(z)
    s = d = 0.0
    for a, b in pieces:
        p1, p2 = np.arctan2(a, x), np.arctan2(b, x)
        th = p1 + (p2 - p1) * g
        R = np.sqrt((x / np.cos(th)) ** 2 + z * z)
        s -= np.sum(w * (p2 - p1) * (np.exp(-kappa * R) - np.exp(-kappa * az)) / kappa)
        d -= np.sum(w * (p2 - p1) * (sz * np.exp(-kappa * az) - z / R * np.exp(-kappa * R)))
    return s, d


@pytest.mark.parametrize("z", [0.3, -0.05])
@pytest.mark.parametrize("case", range(len(LINE_SCHEDULES)))
def test_yukawa_line_schedule(case, z):
    """The 4-point line integral equals the 4-point rule on the hand-derived
    pieces to rounding (1e-13), per case; shifting any piece boundary --
    growth by 5 instead of 10, the first piece ending at x instead of
    10 y1, one piece over the whole range -- moves it by far more, so a
    changed split fails.  The integrand itself is pinned by the 16-point
    brute-force check above."""
    x, y1, y2, pieces = LINE_SCHEDULES[case]
    kappa = 0.5
    s, d = oracle.yukawa_line(z, x, y1, y2, kappa)
    es, ed = _line_by_pieces(z, x, pieces, kappa)
    assert abs(s - es) <= 1e-13 * abs(es)
    assert abs(d - ed) <= 1e-13 * max(abs(ed), 1e-3)
    # the same rule on other splits of [y1, y2]: one piece, and every
    # interior boundary moved by a tenth of its value
    alts = [[(y1, y2)]]
    if len(pieces) > 1:
        b = [p[0] for p in pieces] + [y2]
        moved = [b[0]] + [v * 1.1 for v in b[1:-1]] + [b[-1]]
        alts.append(list(zip(moved[:-1], moved[1:])))
        for k in range(1, len(b) - 1):
            one = list(b)
            one[k] = 0.5 * (b[k - 1] + b[k])
            alts.append(list(zip(one[:-1], one[1:])))
    for alt in alts[1 if len(pieces) == 1 else 0:]:
        als, ald = _line_by_pieces(z, x, alt, kappa)
        assert abs(als - s) > 1e-10 * abs(s) or abs(ald - d) > 1e-10 * max(abs(d), 1e-3)


def test_yukawa_kappa_derivative_matches_finite_differences():
    """orc_yukawa_dk, the 4-point scheme differentiated term by term in kappa
    (the kernel of the kappa gradient), equals the central difference of the
    scheme's own values (h = 1e-5, error O(h^2)) near, on and far from the
    panels, for both layers."""
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=3, jitter=0.03, seed=11)
    P = meshes.panels_from_mesh(V, F)
    rng = np.random.default_rng(0)
    Pt = P[rng.choice(len(P), 40, replace=False)]
    cent = P.reshape(-1, 3, 3).mean(1)
    pts = np.vstack([cent[rng.choice(len(P), 60, replace=False)], Pt.reshape(-1, 3, 3).mean(1)[:5]])
    for k in (0.124, 0.6):
        h = 1e-5
        sp, dp = oracle.integrate(Pt, pts, oracle.YUKAWA, k + h)
        sm, dm = oracle.integrate(Pt, pts, oracle.YUKAWA, k - h)
        ds, dd = oracle.yukawa_dk(Pt, pts, k)
        assert np.abs((sp - sm) / (2 * h) - ds).max() <= 2e-8 * np.abs(ds).max()
        assert np.abs((dp - dm) / (2 * h) - dd).max() <= 2e-8 * np.abs(dd).max()
