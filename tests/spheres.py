"""Closed-form reaction energy of one point charge among concentric spheres.

Regions k = 0..K from the centre out, separated by spheres of radii
R[0] < ... < R[K-1]; region k has permittivity eps[k] and inverse Debye
length kap[k] (0: Laplace).  The charge q sits on the z axis at distance s
inside region m, which must be salt free.  Per Legendre order l the
potential in region k is a_k F_k(r) + b_k G_k(r) (plus, in region m, the
direct term q / (eps_m |r - r0|) = q/eps_m sum r_<^l / r_>^(l+1) P_l), with
F regular at the centre (r^l, or i_l(kap r)) and G decaying outwards
(r^-(l+1), or k_l(kap r)); region 0 has no G term, region K no F term.
Continuity of phi and of eps dphi/dr on every sphere gives 2K equations for
the 2K coefficients.  F_k is scaled to 1 at the region's outer sphere and
G_k to 1 at its inner one, so every basis value entering the system is at
most 1 in magnitude.

The reaction potential at the charge is sum_l a_m F_m(s) + b_m G_m(s); the
energy 1/2 q phi_reac * conversion is in the units of the solvers'
energy() (Gaussian potential q / (eps r), kcal/mol with conversion 332.06).
"""
import numpy as np
from scipy import special


def _basis(l, kap, r, ref, regular):
    """Value and radial derivative at r of F (regular) or G, scaled to 1 at ref."""
    if kap == 0.0:
        if regular:
            return (r / ref) ** l, (l / r * (r / ref) ** l if r > 0 else 0.0)
        return (ref / r) ** (l + 1), -(l + 1) / r * (ref / r) ** (l + 1)
    if regular:
        f = special.spherical_in
    else:
        f = special.spherical_kn
    v0 = f(l, kap * ref)
    v = f(l, kap * r) / v0
    d = kap * f(l, kap * r, derivative=True) / v0
    return v, d


def reaction_potential(R, eps, kap, m, s, q=1.0, lmax=200, tol=1e-15):
    R = [float(x) for x in R]
    K = len(R)
    assert len(eps) == len(kap) == K + 1 and 0 <= m <= K and kap[m] == 0.0
    lo = R[m - 1] if m > 0 else 0.0
    hi = R[m] if m < K else np.inf
    assert lo < s < hi or (m == 0 and s == 0.0)
    total = 0.0
    for l in range(lmax + 1):
        # unknown columns: a_0, (a_k, b_k) for 0 < k < K, b_K
        col = {}
        for k in range(K + 1):
            if k < K:
                col[(k, "a")] = len(col)
            if k > 0:
                col[(k, "b")] = len(col)
        A = np.zeros((2 * K, 2 * K))
        rhs = np.zeros(2 * K)
        for j in range(K):          # sphere j between regions j and j+1
            r = R[j]
            for side, k in ((1.0, j), (-1.0, j + 1)):
                if k < K:
                    v, d = _basis(l, kap[k], r, R[k], True)
                    A[2 * j, col[(k, "a")]] += side * v
                    A[2 * j + 1, col[(k, "a")]] += side * eps[k] * d
                if k > 0:
                    v, d = _basis(l, kap[k], r, R[k - 1], False)
                    A[2 * j, col[(k, "b")]] += side * v
                    A[2 * j + 1, col[(k, "b")]] += side * eps[k] * d
                if k == m:
                    # direct term of the charge on this sphere, moved to the right-hand side
                    if r < s:
                        v = q / eps[m] * r ** l / s ** (l + 1)
                        d = q / eps[m] * l * r ** (l - 1) / s ** (l + 1) if l > 0 else 0.0
                    else:
                        v = q / eps[m] * s ** l / r ** (l + 1)
                        d = -q / eps[m] * (l + 1) * s ** l / r ** (l + 2)
                    rhs[2 * j] -= side * v
                    rhs[2 * j + 1] -= side * eps[m] * d
        c = np.linalg.solve(A, rhs)
        term = 0.0
        if m < K:
            term += c[col[(m, "a")]] * _basis(l, kap[m], s, R[m], True)[0]
        if m > 0:
            term += c[col[(m, "b")]] * _basis(l, kap[m], s, R[m - 1], False)[0]
        total += term
        if s == 0.0:
            break                   # only l = 0 sees a central charge
        if l > 8 and abs(term) < tol * abs(total):
            break
    else:
        raise RuntimeError("multipole series did not converge")
    return total


def reaction_energy(R, eps, kap, m, s, q=1.0, conversion=332.06):
    return 0.5 * q * reaction_potential(R, eps, kap, m, s, q) * conversion
