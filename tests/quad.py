"""Independent brute-force quadrature of flat-triangle layer potentials.

The triangle is split into the three signed sub-triangles (p, v_i, v_i+1)
around the projection p of the observation point onto the panel plane,
each mapped from the unit square by a Duffy transform whose Jacobian
vanishes at p, then integrated with a tensor Gauss-Legendre rule.  This
shares nothing with the reference's closed forms (Hess-Smith) or its
radial line-integral scheme, so agreement pins those to the mathematics.
"""
import numpy as np

_CACHE = {}


def _gl(n):
    if n not in _CACHE:
        x, w = np.polynomial.legendre.leggauss(n)
        _CACHE[n] = (0.5 * (x + 1.0), 0.5 * w)
    return _CACHE[n]


def layer_integrals(x, tri, kappa=0.0, n=48):
    """(slp, dlp) of the kernel e^{-kappa r}/r over triangle tri (3x3) at x.
    Double layer uses the normal (v2-v0) x (v1-v0) normalised, the analytic
    frame normal of the reference panel (FlatPanel.c:41-49)."""
    tri = np.asarray(tri, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    nrm = np.cross(tri[2] - tri[0], tri[1] - tri[0])
    nrm /= np.linalg.norm(nrm)
    z = np.dot(x - tri[0], nrm)
    p = x - z * nrm
    orient = np.sign(np.dot(np.cross(tri[1] - tri[0], tri[2] - tri[0]), nrm))
    u, wu = _gl(n)
    s_tot = 0.0
    d_tot = 0.0
    U, Wv = np.meshgrid(u, u, indexing="ij")
    WW = np.outer(wu, wu)
    for i in range(3):
        a = tri[i] - p
        b = tri[(i + 1) % 3] - tri[i]
        jac = orient * np.dot(np.cross(a, b), nrm)  # signed area * 2
        if abs(jac) < 1e-300:
            continue
        Y = p[None, None, :] + U[..., None] * a[None, None, :] + (U * Wv)[..., None] * b[None, None, :]
        R = x[None, None, :] - Y
        r = np.linalg.norm(R, axis=-1)
        rn = R @ nrm
        if kappa > 0:
            g = np.exp(-kappa * r) / r
            dg = rn * (1.0 + kappa * r) * np.exp(-kappa * r) / r ** 3
        else:
            g = 1.0 / r
            dg = rn / r ** 3
        J = U * jac
        s_tot += np.sum(WW * g * J)
        d_tot += np.sum(WW * dg * J)
    return s_tot, d_tot
