"""ctypes bindings of the CPU oracle (test infrastructure only).

The oracle is a plain-C FP64 restatement of the reference FFT-SVD BEM
algorithm (see oracle.h).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module; the product package
fftsvd_amd never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "orc_panel_size": (C.c_size_t, []),
        "orc_integrate_batch": (None, [C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
        "orc_yukawa_dk_batch": (None, [C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
        "orc_panel_geom": (None, [_dp, _dp]),
        "orc_gl4": (None, [_dp, _dp]),
        "orc_yukawa_rule_v9": (None, [_dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
        "orc_yukawa_line": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp,
                                   _dp, _dp]),
        "orc_dft_r2c": (None, [C.c_int, _dp, _dp]),
        "orc_dft_c2r": (None, [C.c_int, _dp, _dp]),
        "orc_pinv": (None, [_dp, C.c_int, C.c_int, _dp]),
        "orc_colbasis_mgs": (C.c_int, [_dp, C.c_int, C.c_int, C.c_double, _dp]),
        "orc_read_vertface": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(_dp)]),
        "orc_free": (None, [C.c_void_p]),
        "orc_problem_single": (C.c_void_p, [_dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                            C.c_double, C.c_int, C.c_int, C.c_double]),
        "orc_problem_stern": (C.c_void_p, [_dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                           C.c_double, C.c_int, C.c_int, C.c_double]),
        "orc_problem_layout": (C.c_void_p, [_dp, _ip, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _dp, _dp,
                                            C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                            C.c_double]),
        "orc_problem_ntotal": (C.c_int, [C.c_void_p]),
        "orc_problem_make_dense": (None, [C.c_void_p]),
        "orc_problem_free": (None, [C.c_void_p]),
        "orc_so_matvec": (None, [C.c_void_p, _dp, _dp]),
        "orc_rhs": (None, [C.c_void_p, _dp]),
        "orc_precond_build": (None, [C.c_void_p]),
        "orc_precond_solve": (None, [C.c_void_p, _dp, _dp]),
        "orc_gmres": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_int, _dp]),
        "orc_reaction_potentials": (None, [C.c_void_p, _dp, _dp]),
        "orc_problem_tree": (C.c_void_p, [C.c_void_p, C.c_int]),
        "orc_tree_depth": (C.c_int, [C.c_void_p]),
        "orc_tree_ncubes": (C.c_int, [C.c_void_p]),
        "orc_tree_gpl": (None, [C.c_void_p, _ip]),
        "orc_tree_cube_info": (None, [C.c_void_p, C.c_int, _ip]),
        "orc_tree_cube_lists": (None, [C.c_void_p, C.c_int, _ip, _ip, _ip, _ip]),
        "orc_tree_cube_mats": (None, [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_tree_set_cube_mats": (None, [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                          _dp, _dp]),
        "orc_tree_pinv": (None, [C.c_void_p, C.c_int, _dp]),
        "orc_tree_T": (None, [C.c_void_p, C.c_int, C.c_int, _dp]),
        "orc_tree_apply_both": (None, [C.c_void_p, _dp, _dp, _dp]),
        "orc_tree_apply_transpose": (None, [C.c_void_p, C.c_int, _dp, _dp]),
        "orc_tree_dense": (None, [C.c_void_p, _dp, _dp]),
        "orc_m3_build": (None, [C.c_void_p]),
        "orc_m3_potentials": (None, [C.c_void_p, _dp, _dp]),
        "orc_problem_m3_tree": (C.c_void_p, [C.c_void_p, C.c_int]),
        "orc_qual_create": (C.c_void_p, [_dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int,
                                         C.c_int, C.c_double]),
        "orc_qual_free": (None, [C.c_void_p]),
        "orc_qual_matvec": (None, [C.c_void_p, _dp, _dp]),
        "orc_qual_precond_diag": (None, [C.c_void_p, _dp]),
        "orc_qual_rhs": (None, [C.c_void_p, _dp]),
        "orc_qual_gmres": (C.c_int, [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int, _dp]),
        "orc_qual_potentials": (None, [C.c_void_p, _dp, _dp]),
        "orc_qual_tree": (C.c_void_p, [C.c_void_p]),
        "orc_qual_m1m3_build": (None, [C.c_void_p]),
        "orc_qual_m1m3_tree": (C.c_void_p, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


LAPLACE, YUKAWA = 0, 1


def integrate(verts9, pts, kernel=LAPLACE, kappa=0.0):
    """slp, dlp of every (point, panel) pair: arrays (npt, npan)."""
    v = f64(verts9).reshape(-1, 9)
    p = f64(pts).reshape(-1, 3)
    s = np.zeros((len(p), len(v)))
    d = np.zeros((len(p), len(v)))
    lib().orc_integrate_batch(kernel, kappa, _d(v), len(v), _d(p), len(p), _d(s), _d(d))
    return s, d


def yukawa_dk(verts9, pts, kappa):
    """d/dkappa of the Yukawa (slp, dlp) of integrate(): arrays (npt, npan)."""
    v = f64(verts9).reshape(-1, 9)
    p = f64(pts).reshape(-1, 3)
    s = np.zeros((len(p), len(v)))
    d = np.zeros((len(p), len(v)))
    lib().orc_yukawa_dk_batch(kappa, _d(v), len(v), _d(p), len(p), _d(s), _d(d))
    return s, d


def yukawa_rule(x, v9, kappa, nodes, weights):
    """Yukawa (slp, dlp) by the reference's edge/radial scheme with a given angular rule."""
    s = C.c_double()
    d = C.c_double()
    xq = f64(nodes)
    wq = f64(weights)
    lib().orc_yukawa_rule_v9(_d(f64(x)), _d(f64(v9)), kappa, len(xq), _d(xq), _d(wq), C.byref(s), C.byref(d))
    return s.value, d.value


def yukawa_line(z, x, y1, y2, kappa, nodes=None, weights=None):
    """calcp_LogIntDoLine's radial line integral (slp, dlp part) with the
    4-point Gauss-Legendre rule or a given angular rule."""
    if nodes is None:
        nodes, weights = np.zeros(4), np.zeros(4)
        lib().orc_gl4(_d(nodes), _d(weights))
    s = C.c_double()
    d = C.c_double()
    xq = f64(nodes)
    wq = f64(weights)
    lib().orc_yukawa_line(z, x, y1, y2, kappa, len(xq), _d(xq), _d(wq), C.byref(s), C.byref(d))
    return s.value, d.value


def panel_geom(v9):
    out = np.zeros(39)
    lib().orc_panel_geom(_d(f64(v9)), _d(out))
    return {"centroid": out[0:3], "area": out[3], "maxd": out[4], "mind": out[5], "moments": out[6:21],
            "ax": out[21:30].reshape(3, 3), "axn": out[30:39].reshape(3, 3)}


def gl4():
    x = np.zeros(4)
    w = np.zeros(4)
    lib().orc_gl4(_d(x), _d(w))
    return x, w


def dft_r2c(grid):
    N = grid.shape[0]
    out = np.zeros((N, N, N // 2 + 1, 2))
    lib().orc_dft_r2c(N, _d(f64(grid)), _d(out))
    return out[..., 0] + 1j * out[..., 1]


def dft_c2r(spec, N):
    s = np.zeros(spec.shape + (2,))
    s[..., 0] = spec.real
    s[..., 1] = spec.imag
    out = np.zeros((N, N, N))
    lib().orc_dft_c2r(N, _d(f64(s)), _d(out))
    return out


def pinv(X):
    X = f64(X)
    m, n = X.shape
    out = np.zeros((n, m))
    lib().orc_pinv(_d(X), m, n, _d(out))
    return out


def colbasis_mgs(A, eps):
    A = f64(A)
    r, c = A.shape
    Q = np.zeros((r, c))
    rank = lib().orc_colbasis_mgs(_d(A), r, c, eps, _d(Q))
    return Q[:, :rank]


def read_vertface(vertfile, facefile, flip=True, accessible=False):
    p = _dp()
    n = lib().orc_read_vertface(vertfile.encode(), facefile.encode(), int(flip), int(accessible), C.byref(p))
    if n < 0:
        raise IOError("could not open %s / %s" % (vertfile, facefile))
    out = np.ctypeslib.as_array(p, shape=(max(n, 1) * 9,)).copy()[: n * 9].reshape(n, 9)
    lib().orc_free(p)
    return out


class Tree:
    """View of one compressed-operator tree owned by a Problem."""

    def __init__(self, handle, owner):
        self.h = handle
        self.owner = owner

    @property
    def depth(self):
        return lib().orc_tree_depth(self.h)

    @property
    def n(self):
        """Panels of the tree (those of its root cube)."""
        info = np.zeros(13, dtype=np.int32)
        lib().orc_tree_cube_info(self.h, 0, _i(info))
        return int(info[5])

    def gpl(self):
        g = np.zeros(self.depth, dtype=np.int32)
        lib().orc_tree_gpl(self.h, _i(g))
        return g

    def cubes(self):
        """List of dicts describing every cube (ids are oracle-internal)."""
        out = []
        n = lib().orc_tree_ncubes(self.h)
        for c in range(n):
            info = np.zeros(13, dtype=np.int32)
            lib().orc_tree_cube_info(self.h, c, _i(info))
            lvl, i, j, k, leaf, npan, npt, nloc, nint, rr, cr, parent, dcols = [int(v) for v in info]
            pan = np.zeros(max(npan, 1), dtype=np.int32)
            pt = np.zeros(max(npt, 1), dtype=np.int32)
            loc = np.zeros(max(nloc, 1), dtype=np.int32)
            itr = np.zeros(max(nint, 1), dtype=np.int32)
            lib().orc_tree_cube_lists(self.h, c, _i(pan), _i(pt), _i(loc), _i(itr))
            out.append(dict(id=c, level=lvl, idx=(i, j, k), leaf=leaf, pan=pan[:npan].copy(), pt=pt[:npt].copy(),
                            loc=loc[:nloc].copy(), itr=itr[:nint].copy(), rowrank=rr, colrank=cr, parent=parent,
                            dcols=dcols))
        return out

    def cube_mats(self, cube):
        gp3 = int(self.gpl()[cube["level"]]) ** 3
        drows, dcols = len(cube["pt"]), cube["dcols"]
        Ds = np.zeros(max(drows * dcols, 1))
        Dd = np.zeros(max(drows * dcols, 1))
        VT = np.zeros(max(cube["rowrank"] * len(cube["pan"]), 1))
        U = np.zeros(max(cube["colrank"] * len(cube["pt"]), 1))
        PVs = np.zeros(max(gp3 * cube["rowrank"], 1))
        PVd = np.zeros(max(gp3 * cube["rowrank"], 1))
        UTI = np.zeros(max(gp3 * cube["colrank"], 1))
        lib().orc_tree_cube_mats(self.h, cube["id"], _d(Ds), _d(Dd), _d(VT), _d(U), _d(PVs), _d(PVd), _d(UTI))
        return dict(Ds=Ds[:drows * dcols].reshape(drows, dcols), Dd=Dd[:drows * dcols].reshape(drows, dcols),
                    VT=VT[:cube["rowrank"] * len(cube["pan"])].reshape(cube["rowrank"], len(cube["pan"])),
                    U=U[:cube["colrank"] * len(cube["pt"])].reshape(len(cube["pt"]), cube["colrank"]),
                    PVs=PVs[:gp3 * cube["rowrank"]].reshape(gp3, cube["rowrank"]),
                    PVd=PVd[:gp3 * cube["rowrank"]].reshape(gp3, cube["rowrank"]),
                    UTI=UTI[:gp3 * cube["colrank"]].reshape(cube["colrank"], gp3))

    def set_cube_mats(self, cube_id, rowrank, colrank, Ds, Dd, VT, U, PVs, PVd, UTI):
        arrs = [None if a is None else f64(a) for a in (Ds, Dd, VT, U, PVs, PVd, UTI)]
        ptrs = [_d(a) if a is not None and a.size else None for a in arrs]
        lib().orc_tree_set_cube_mats(self.h, cube_id, rowrank, colrank, *ptrs)

    def pinv(self, level):
        n = int(self.gpl()[level])
        out = np.zeros((n ** 3, 49))
        lib().orc_tree_pinv(self.h, level, _d(out))
        return out

    def T(self, level, dx, dy, dz):
        n = int(self.gpl()[level])
        N = 2 * n - 1
        out = np.zeros((N, N, N // 2 + 1, 2))
        lib().orc_tree_T(self.h, level, (dx + 3) * 49 + (dy + 3) * 7 + (dz + 3), _d(out))
        return out[..., 0] + 1j * out[..., 1]

    def apply_both(self, xs, xd):
        xs = f64(xs)
        xd = f64(xd)
        y = np.zeros(len(xs))
        lib().orc_tree_apply_both(self.h, _d(xs), _d(xd), _d(y))
        return y

    def apply_transpose(self, layer, x):
        x = f64(x)
        y = np.zeros(len(x))
        lib().orc_tree_apply_transpose(self.h, layer, _d(x), _d(y))
        return y

    def dense(self, n):
        S = np.zeros((n, n))
        D = np.zeros((n, n))
        lib().orc_tree_dense(self.h, _d(S), _d(D))
        return S, D


class Problem:
    """One dielectric surface with its charges (two-region operator), or,
    with `stern_verts9` (a salt / ion-exclusion surface enclosing it, kappa
    > 0), the three-region Stern-layer operator (FFTSVDpbeAPI.c:1064-1377);
    unknowns then run salt phi, salt dphi, dielectric phi, dielectric dphi.

    `cavities` (panel arrays) are dielectric cavities inside the dielectric
    surface; `salt_cavities` (Stern layout only) sit inside the dielectric
    cavities given by `salt_cavity_parent`.  A parent of 999 drops the
    cavity, as readSRF does for surfaces the mesher marked bad
    (FFTSVDpbeAPI.c:1006-1029).  Regions and unknowns follow
    generateSurfaceOperator's layout for every branch."""

    def __init__(self, verts9, qpos, q, eps_in=4.0, eps_out=80.0, kappa=0.0, maxpanels=32, gridpoints=4,
                 svdeps=1e-5, stern_verts9=None, cavities=None, cavity_parent=None, salt_cavities=None,
                 salt_cavity_parent=None):
        self.verts9 = f64(verts9).reshape(-1, 9)
        self.qpos = f64(qpos).reshape(-1, 3)
        self.q = f64(q).reshape(-1)
        self.eps_in = eps_in
        if cavities or salt_cavities:
            salts = [] if stern_verts9 is None else [f64(stern_verts9).reshape(-1, 9)]
            cav = [f64(c).reshape(-1, 9) for c in (cavities or [])]
            scav = [f64(c).reshape(-1, 9) for c in (salt_cavities or [])]
            surf = salts + [self.verts9] + cav + scav
            allv = np.ascontiguousarray(np.vstack(surf))
            counts = np.array([len(x) for x in surf], dtype=np.int32)
            dpar = np.zeros(1, dtype=np.int32)
            cpar = np.array(cavity_parent if cavity_parent is not None else [0] * len(cav), dtype=np.int32)
            spar = np.array(salt_cavity_parent if salt_cavity_parent is not None else [0] * len(scav),
                            dtype=np.int32)
            cpar = cpar if len(cpar) else np.zeros(1, dtype=np.int32)
            spar = spar if len(spar) else np.zeros(1, dtype=np.int32)
            self.h = lib().orc_problem_layout(_d(allv), _i(counts), len(salts), 1, len(cav), len(scav), _i(dpar),
                                              _i(cpar), _i(spar), _d(self.qpos), _d(self.q), len(self.q), eps_in,
                                              eps_out, kappa, int(bool(salts)), maxpanels, gridpoints, svdeps)
            self.n = lib().orc_problem_ntotal(self.h)
        elif stern_verts9 is None:
            self.n = len(self.verts9)
            self.h = lib().orc_problem_single(_d(self.verts9), self.n, _d(self.qpos), _d(self.q), len(self.q),
                                              eps_in, eps_out, kappa, maxpanels, gridpoints, svdeps)
        else:
            self.salt9 = f64(stern_verts9).reshape(-1, 9)
            self.n = len(self.salt9) + len(self.verts9)
            self.h = lib().orc_problem_stern(_d(self.salt9), len(self.salt9), _d(self.verts9), len(self.verts9),
                                             _d(self.qpos), _d(self.q), len(self.q), eps_in, eps_out, kappa,
                                             maxpanels, gridpoints, svdeps)
        lib().orc_precond_build(self.h)

    def make_dense(self):
        """Apply every region's all-pairs integrals instead of the compressed
        operator from now on (the uncompressed BEM operator)."""
        lib().orc_problem_make_dense(self.h)
        return self

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_problem_free(self.h)
            self.h = None

    def tree(self, which):
        return Tree(lib().orc_problem_tree(self.h, which), self)

    def matvec(self, x):
        x = f64(x)
        y = np.zeros(2 * self.n)
        lib().orc_so_matvec(self.h, _d(x), _d(y))
        return y

    def rhs(self):
        r = np.zeros(2 * self.n)
        lib().orc_rhs(self.h, _d(r))
        return r

    def precond_solve(self, b):
        x = np.zeros(2 * self.n)
        lib().orc_precond_solve(self.h, _d(f64(b)), _d(x))
        return x

    def gmres(self, rhs, tol=1e-6, maxit=1000):
        sol = np.zeros(2 * self.n)
        hist = np.zeros(maxit)
        it = lib().orc_gmres(self.h, _d(f64(rhs)), _d(sol), tol, maxit, _d(hist))
        return sol, it, hist[:it]

    def reaction_potentials(self, sol):
        phi = np.zeros(len(self.q))
        lib().orc_reaction_potentials(self.h, _d(f64(sol)), _d(phi))
        return phi

    def energy(self, sol, conversion=332.112):
        """Delta G = 1/2 sum q phi * conversion / eps_in (FFTSVDsalt.c:166-172)."""
        phi = self.reaction_potentials(sol)
        return 0.5 * float(np.dot(self.q, phi)) * conversion / self.eps_in

    def m3_potentials(self, sol):
        """Reaction potentials through the panels -> charges M3 tree
        (accelerateM3), built on first use."""
        lib().orc_m3_build(self.h)
        phi = np.zeros(len(self.q))
        lib().orc_m3_potentials(self.h, _d(f64(sol)), _d(phi))
        return phi

    def m3_tree(self, which):
        lib().orc_m3_build(self.h)
        return Tree(lib().orc_problem_m3_tree(self.h, which), self)


class Qual:
    """Qualocation operator (QualocationOperator.c): one Laplace tree over
    the panels with double-layer-only bases, one unknown per panel.  The
    defaults are generateQualocationOperator's (32 panels per leaf, 4 grid
    points, SVD 1e-5); solveBEMproblem uses 27, 3 and 1e-5."""

    def __init__(self, verts9, qpos, q, eps_in=4.0, eps_out=80.0, maxpanels=32, gridpoints=4, svdeps=1e-5):
        self.verts9 = f64(verts9).reshape(-1, 9)
        self.qpos = f64(qpos).reshape(-1, 3)
        self.q = f64(q).reshape(-1)
        self.eps_in = eps_in
        self.n = len(self.verts9)
        self.h = lib().orc_qual_create(_d(self.verts9), self.n, _d(self.qpos), _d(self.q), len(self.q), eps_in,
                                       eps_out, maxpanels, gridpoints, svdeps)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_qual_free(self.h)
            self.h = None

    def tree(self, which=0):
        return Tree(lib().orc_qual_tree(self.h), self)

    def use_m1m3(self):
        """Build the M1M3 tree; rhs() and potentials() then go through it."""
        lib().orc_qual_m1m3_build(self.h)
        return Tree(lib().orc_qual_m1m3_tree(self.h), self)

    def matvec(self, x):
        y = np.zeros(self.n)
        lib().orc_qual_matvec(self.h, _d(f64(x)), _d(y))
        return y

    def precond_diag(self):
        p = np.zeros(self.n)
        lib().orc_qual_precond_diag(self.h, _d(p))
        return p

    def rhs(self):
        r = np.zeros(self.n)
        lib().orc_qual_rhs(self.h, _d(r))
        return r

    def gmres(self, rhs, tol=1e-6, maxit=1000, early_exit=True):
        sol = np.zeros(self.n)
        hist = np.zeros(maxit)
        it = lib().orc_qual_gmres(self.h, _d(f64(rhs)), _d(sol), tol, maxit, int(early_exit), _d(hist))
        return sol, it, hist[:it]

    def potentials(self, sol):
        phi = np.zeros(len(self.q))
        lib().orc_qual_potentials(self.h, _d(f64(sol)), _d(phi))
        return phi

    def energy(self, sol, conversion=332.112):
        """1/2 sum q phi * conversion / eps_in (FFTSVDsolvecfqual.c:153-167)."""
        return 0.5 * float(np.dot(self.q, self.potentials(sol))) * conversion / self.eps_in
