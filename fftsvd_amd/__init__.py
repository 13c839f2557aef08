"""MI355X-native FFT-SVD Poisson-Boltzmann boundary-element solver.

Python host mirror of the reference's solve path (FFTSVDpbeAPI.h) over the
C-ABI library ``_lib/libfftsvd_hip.so`` (hand-written gfx950 HIP kernels).
There is no CPU fallback: every call goes to the GPU through the library,
and loading fails loudly if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FFTSVD_LIBRARY selects another build of the same library (the
# host-sanitized one in _lib/asan/, for the sanitizer runs)
LIB_PATH = os.environ.get("FFTSVD_LIBRARY") or os.path.join(HERE, "_lib", "libfftsvd_hip.so")
HEADERS = [os.path.join(os.path.dirname(HERE), "include", h) for h in ("FFTSVDpbeAPI.h", "fftsvd_hip.h")]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

F64, MIXED = 0, 1
LAPLACE, YUKAWA, YUKAWA_DKAPPA = 0, 1, 2

_SIGS = {
    "fftsvd_last_error": (C.c_char_p, []),
    "fftsvd_device_info": (C.c_int, [C.c_char_p, C.c_int, _ip]),
    "fftsvd_set_precision": (C.c_int, [C.c_int]),
    "fftsvd_problem_create": (C.c_int, [_dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(_vp)]),
    "fftsvd_problem_create_multi": (C.c_int, [C.c_int, _ip, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int, C.c_double,
                                              C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                              C.POINTER(_vp)]),
    "fftsvd_problem_create_stern": (C.c_int, [C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip,
                                              _dp, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                              C.c_int, C.c_double, C.c_int, C.POINTER(_vp)]),
    "fftsvd_problem_destroy": (C.c_int, [_vp]),
    "fftsvd_problem_size": (C.c_int, [_vp, _ip, _ip]),
    "fftsvd_problem_unknowns": (C.c_int, [_vp, C.POINTER(C.c_longlong)]),
    "fftsvd_problem_rhs_batch": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "fftsvd_problem_solve_batch": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _dp]),
    "fftsvd_problem_potentials_batch": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "fftsvd_problem_matvec_transpose": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "fftsvd_problem_matvec_batch": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "fftsvd_lds_bandwidth": (C.c_int, [C.c_int, C.c_int, _dp]),
    "fftsvd_problems_solve": (C.c_int, [C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(_dp), C.POINTER(_dp), C.c_double,
                                        C.c_int, C.POINTER(C.c_int), _dp]),
    "fftsvd_tree_cube_points": (C.c_int, [_vp, C.c_int, C.c_int, _ip]),
    "fftsvd_problem_solve_adjoint": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _dp]),
    "fftsvd_problem_grad_kappa": (C.c_int, [_vp, C.c_int, _dp, C.c_double, C.c_int, _dp, _ip]),
    "fftsvd_set_autotune": (C.c_int, [C.c_int]),
    "fftsvd_translation_params": (C.c_int, [_ip, _ip, _ip, _ip, _dp]),
    "fftsvd_problem_grad_charges": (C.c_int, [_vp, C.c_int, _dp, C.c_double, C.c_int, _dp, _ip]),
    "fftsvd_qual_create": (C.c_int, [_dp, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                     C.c_double, C.c_int, C.POINTER(_vp)]),
    "fftsvd_problem_matvec": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "fftsvd_problem_precond": (C.c_int, [_vp, _dp, _dp]),
    "fftsvd_problem_rhs": (C.c_int, [_vp, _dp]),
    "fftsvd_problem_solve": (C.c_int, [_vp, C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _dp]),
    "fftsvd_problem_potentials": (C.c_int, [_vp, _dp, _dp]),
    "fftsvd_problem_stats": (C.c_int, [_vp, _dp, C.c_int]),
    "fftsvd_copy_bandwidth": (C.c_int, [C.c_longlong, C.c_int, _dp]),
    "fftsvd_problem_bench_matvec": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "fftsvd_problem_bench_matvec_batch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "fftsvd_tree_info": (C.c_int, [_vp, C.c_int, _ip, _ip, _ip]),
    "fftsvd_tree_cube_info": (C.c_int, [_vp, C.c_int, C.c_int, _ip]),
    "fftsvd_tree_cube_panels": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip]),
    "fftsvd_tree_cube_mats": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
    "fftsvd_tree_apply": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp, _dp]),
    "fftsvd_tree_apply_transpose": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "fftsvd_integrate": (C.c_int, [C.c_int, C.c_double, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
}

_lib = None


class FFTSVDError(RuntimeError):
    pass


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FFTSVDError("libfftsvd_hip.so not built (run python -m fftsvd_amd.build): %s" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise FFTSVDError(lib().fftsvd_last_error().decode())


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int()
    _check(lib().fftsvd_device_info(name, 256, C.byref(cu)))
    return name.value.decode(), cu.value


def copy_bandwidth(nbytes=1 << 30, nrep=20):
    """Device-to-device copy rate in GB/s (bytes read + written per second)."""
    g = C.c_double()
    _check(lib().fftsvd_copy_bandwidth(int(nbytes), int(nrep), C.byref(g)))
    return g.value


def lds_bandwidth(width=8, nrep=5):
    """LDS read rate of the device in GB/s with width-byte reads
    (fftsvd_lds_bandwidth)."""
    g = C.c_double()
    _check(lib().fftsvd_lds_bandwidth(int(width), int(nrep), C.byref(g)))
    return g.value


def set_autotune(on):
    """Autotune the FP32 translation of the next operator built (see
    fftsvd_set_autotune); off restores the defaults."""
    _check(lib().fftsvd_set_autotune(int(bool(on))))


def translation_params():
    """The FP32 translation parameters in use: dict(mode, budget, tile,
    tuned, tuned_ms)."""
    v = [C.c_int() for _ in range(4)]
    ms = C.c_double()
    _check(lib().fftsvd_translation_params(*[C.byref(x) for x in v], C.byref(ms)))
    return dict(mode=v[0].value, budget=v[1].value, tile=v[2].value, tuned=bool(v[3].value), tuned_ms=ms.value)


def solve_many(problems, rhs, tol=1e-6, maxit=1000, precision=F64):
    """Solves problems[k] with right-hand side rhs[k] for every k at once
    (fftsvd_problems_solve: concurrent solves, each as Problem.solve
    computes it).  Returns (solutions, iterations, info dicts)."""
    n = len(problems)
    b = [_f64(r).reshape(-1) for r in rhs]
    x = [np.zeros(p.nvec) for p in problems]
    for p, r in zip(problems, b):
        if r.size != p.nvec:
            raise ValueError("right-hand side of the wrong length")
    hs = (_vp * n)(*[p.h for p in problems])
    bp = (_dp * n)(*[_d(r) for r in b])
    xp = (_dp * n)(*[_d(v) for v in x])
    its = np.zeros(n, dtype=np.int32)
    info = np.zeros(6 * n)
    _check(lib().fftsvd_problems_solve(n, hs, precision, bp, xp, tol, maxit, _i(its), _d(info)))
    keys = ("iterations", "refinements", "final_residual", "ms", "matvecs64", "matvecs32")
    return x, its, [dict(zip(keys, info[6 * k:6 * k + 6].tolist())) for k in range(n)]


def integrate(verts9, pts, kernel=LAPLACE, kappa=0.0):
    """Panel integrals (slp, dlp) on the GPU, arrays (npt, npan)."""
    v = _f64(verts9).reshape(-1, 9)
    p = _f64(pts).reshape(-1, 3)
    s = np.zeros((len(p), len(v)))
    d = np.zeros((len(p), len(v)))
    _check(lib().fftsvd_integrate(kernel, kappa, _d(v), len(v), _d(p), len(p), _d(s), _d(d)))
    return s, d


class Tree:
    """One region's compressed operator (view)."""

    def __init__(self, prob, which):
        self.p = prob
        self.which = which
        depth, nc, npan = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fftsvd_tree_info(prob.h, which, C.byref(depth), C.byref(nc), C.byref(npan)))
        self.depth, self.ncubes, self.n = depth.value, nc.value, npan.value
        info = np.zeros(13, dtype=np.int32)   # the root cube holds every point
        _check(lib().fftsvd_tree_cube_info(prob.h, which, 0, _i(info)))
        self.npt = int(info[6])

    def cubes(self):
        out = []
        for c in range(self.ncubes):
            info = np.zeros(13, dtype=np.int32)
            _check(lib().fftsvd_tree_cube_info(self.p.h, self.which, c, _i(info)))
            lvl, i, j, k, leaf, npan, npt, nloc, nint, rr, cr, parent, dcols = [int(v) for v in info]
            pan = np.zeros(max(npan, 1), dtype=np.int32)
            dc = np.zeros(max(dcols, 1), dtype=np.int32)
            _check(lib().fftsvd_tree_cube_panels(self.p.h, self.which, c, _i(pan), _i(dc)))
            out.append(dict(id=c, level=lvl, idx=(i, j, k), leaf=leaf, npan=npan, npt=npt, nloc=nloc, nint=nint,
                            rowrank=rr, colrank=cr, parent=parent, pan=pan[:npan].copy(), dcols=dc[:dcols].copy()))
        return out

    def cube_mats(self, cube, gp3):
        np_, nq, r, rr = cube["npan"], cube["npt"], cube["rowrank"], cube["colrank"]
        m = len(cube["dcols"])
        VT = np.zeros(max(r * np_, 1))
        U = np.zeros(max(nq * rr, 1))
        PVs = np.zeros(max(gp3 * r, 1))
        PVd = np.zeros(max(gp3 * r, 1))
        UTI = np.zeros(max(gp3 * rr, 1))
        Ds = np.zeros(max(nq * m, 1))
        Dd = np.zeros(max(nq * m, 1))
        _check(lib().fftsvd_tree_cube_mats(self.p.h, self.which, cube["id"], _d(VT), _d(U), _d(PVs), _d(PVd), _d(UTI),
                                           _d(Ds), _d(Dd)))
        return dict(VT=VT[:r * np_].reshape(r, np_), U=U[:nq * rr].reshape(nq, rr),
                    PVs=PVs[:gp3 * r].reshape(gp3, r), PVd=PVd[:gp3 * r].reshape(gp3, r),
                    UTI=UTI[:gp3 * rr].reshape(rr, gp3), Ds=Ds[:nq * m].reshape(nq, m),
                    Dd=Dd[:nq * m].reshape(nq, m))

    def apply_both(self, xs, xd, precision=F64):
        y = np.zeros(self.npt)
        _check(lib().fftsvd_tree_apply(self.p.h, self.which, precision, _d(_f64(xs)), _d(_f64(xd)), _d(y)))
        return y

    def apply_transpose(self, layer, x, precision=F64):
        y = np.zeros(self.n)   # x has npt entries, y one per panel
        _check(lib().fftsvd_tree_apply_transpose(self.p.h, self.which, precision, layer, _d(_f64(x)), _d(y)))
        return y


class Problem:
    """Region operator of one dielectric surface (or several, with cavities)
    and its charges, device resident.  Mirrors the reference flow
    generateSurfaceOperator -> generateSurfaceOperatorPreconditioner ->
    generateRHS -> solveSurfaceOperator -> collectPotentials."""

    def __init__(self, verts9, qpos, q, eps_in=4.0, eps_out=80.0, kappa=0.0, maxpanels=32, gridpoints=4,
                 svdeps=1e-5, want_f32=True, cavities=None, stern_verts9=None, cavity_parent=None,
                 salt_cavities=None, salt_cavity_parent=None):
        """cavities: dielectric cavities inside the surface (panel arrays,
        normals pointing out of the cavity); salt_cavities (Stern layout
        only): cavities inside the dielectric cavities salt_cavity_parent.
        A parent of 999 drops the cavity, as the reference does for surfaces
        the mesher rejected."""
        v = _f64(verts9).reshape(-1, 9)
        self.qpos = _f64(qpos).reshape(-1, 3)
        self.q = _f64(q).reshape(-1)
        self.eps_in = eps_in
        h = _vp()
        cv = [_f64(c).reshape(-1, 9) for c in (cavities or [])]
        cp = np.array(cavity_parent if cavity_parent is not None else [0] * len(cv), dtype=np.int32)
        nc = np.array([len(c) for c in cv] or [0], dtype=np.int32)
        cp1 = cp if len(cp) else np.zeros(1, dtype=np.int32)
        if stern_verts9 is not None:
            # Stern layer: one salt (ion-exclusion) surface around the
            # dielectric (and its cavities); unknowns start with the salt's
            sv = _f64(stern_verts9).reshape(-1, 9)
            lv = [_f64(c).reshape(-1, 9) for c in (salt_cavities or [])]
            lp = np.array(salt_cavity_parent if salt_cavity_parent is not None else [0] * len(lv), dtype=np.int32)
            nl = np.array([len(c) for c in lv] or [0], dtype=np.int32)
            lp1 = lp if len(lp) else np.zeros(1, dtype=np.int32)
            allv = np.ascontiguousarray(np.vstack([sv, v] + cv + lv))
            ns = np.array([len(sv)], dtype=np.int32)
            nd = np.array([len(v)], dtype=np.int32)
            dp = np.zeros(1, dtype=np.int32)
            _check(lib().fftsvd_problem_create_stern(1, _i(ns), 1, _i(nd), _i(dp), len(cv), _i(nc), _i(cp1), len(lv),
                                                     _i(nl), _i(lp1), _d(allv), _d(self.qpos), _d(self.q), len(self.q),
                                                     eps_in, eps_out, kappa, maxpanels, gridpoints, svdeps,
                                                     int(want_f32), C.byref(h)))
        elif cv:
            if salt_cavities:
                raise ValueError("salt cavities need the Stern layout (stern_verts9)")
            allv = np.ascontiguousarray(np.vstack([v] + cv))
            nd = np.array([len(v)], dtype=np.int32)
            _check(lib().fftsvd_problem_create_multi(1, _i(nd), len(cv), _i(nc), _i(cp1), _d(allv), _d(self.qpos),
                                                     _d(self.q), len(self.q), eps_in, eps_out, kappa, maxpanels,
                                                     gridpoints, svdeps, int(want_f32), C.byref(h)))
        else:
            _check(lib().fftsvd_problem_create(_d(v), len(v), _d(self.qpos), _d(self.q), len(self.q), eps_in, eps_out,
                                               kappa, maxpanels, gridpoints, svdeps, int(want_f32), C.byref(h)))
        self.h = h
        nv = C.c_longlong()
        _check(lib().fftsvd_problem_unknowns(h, C.byref(nv)))
        self.nvec = nv.value
        self.n = self.nvec // 2

    def close(self):
        if getattr(self, "h", None):
            lib().fftsvd_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tree(self, which):
        return Tree(self, which)

    def matvec(self, x, precision=F64):
        y = np.zeros(self.nvec)
        _check(lib().fftsvd_problem_matvec(self.h, precision, _d(_f64(x)), _d(y)))
        return y

    def matvec_batch(self, X, precision=F64):
        """A applied to the columns of X (nvec x ncol) in one batched call."""
        X = np.asfortranarray(_f64(X).reshape(self.nvec, -1))
        Y = np.zeros(X.shape, order="F")
        _check(lib().fftsvd_problem_matvec_batch(self.h, precision, X.shape[1], _d(X), _d(Y)))
        return Y

    def rhs(self):
        r = np.zeros(self.nvec)
        _check(lib().fftsvd_problem_rhs(self.h, _d(r)))
        return r

    def precond_solve(self, b):
        x = np.zeros(self.nvec)
        _check(lib().fftsvd_problem_precond(self.h, _d(_f64(b)), _d(x)))
        return x

    def solve(self, rhs, tol=1e-6, maxit=1000, precision=F64):
        """Returns (solution, iterations, residual history, info dict)."""
        sol = np.zeros(self.nvec)
        it = C.c_int()
        hist = np.zeros(maxit)
        info = np.zeros(6)
        _check(lib().fftsvd_problem_solve(self.h, precision, _d(_f64(rhs)), _d(sol), tol, maxit, C.byref(it),
                                          _d(hist), _d(info)))
        keys = ("iterations", "refinements", "final_residual", "ms", "matvecs64", "matvecs32")
        return sol, it.value, hist[:max(0, min(it.value, maxit))], dict(zip(keys, info.tolist()))

    def gmres(self, rhs, tol=1e-6, maxit=1000, precision=F64):
        sol, it, hist, _ = self.solve(rhs, tol, maxit, precision)
        return sol, it, hist

    def reaction_potentials(self, sol):
        phi = np.zeros(max(len(self.q), 1))
        _check(lib().fftsvd_problem_potentials(self.h, _d(_f64(sol)), _d(phi)))
        return phi[:len(self.q)]

    def energy(self, sol, conversion=332.112):
        """Delta G = 1/2 sum q phi * conversion / eps_in (reference FFTSVDsalt.c:166-172)."""
        return 0.5 * float(np.dot(self.q, self.reaction_potentials(sol))) * conversion / self.eps_in

    # adjoint
    def matvec_transpose(self, y, precision=F64):
        x = np.zeros(self.nvec)
        _check(lib().fftsvd_problem_matvec_transpose(self.h, precision, _d(_f64(y)), _d(x)))
        return x

    def solve_adjoint(self, rhs, tol=1e-6, maxit=1000, precision=F64):
        sol = np.zeros(self.nvec)
        it = C.c_int()
        info = np.zeros(6)
        _check(lib().fftsvd_problem_solve_adjoint(self.h, precision, _d(_f64(rhs)), _d(sol), tol, maxit,
                                                  C.byref(it), _d(info)))
        return sol, it.value

    def grad_charges(self, sol, tol=1e-6, maxit=1000, precision=F64, conversion=332.112):
        """d(energy)/d(charges) at the forward solution sol, in the units of
        energy(): conversion/eps_in * 1/2 (phi + B^T lambda)."""
        g = np.zeros(max(len(self.q), 1))
        it = C.c_int()
        _check(lib().fftsvd_problem_grad_charges(self.h, precision, _d(_f64(sol)), tol, maxit, _d(g), C.byref(it)))
        return g[:len(self.q)] * conversion / self.eps_in

    def grad_kappa(self, sol, tol=1e-6, maxit=1000, precision=F64, conversion=332.112):
        """d(energy)/d(kappa) at the forward solution sol, in the units of
        energy(): conversion/eps_in * (-1/2 lambda^T dA/dkappa x)."""
        g = C.c_double()
        it = C.c_int()
        _check(lib().fftsvd_problem_grad_kappa(self.h, precision, _d(_f64(sol)), tol, maxit, C.byref(g),
                                               C.byref(it)))
        return g.value * conversion / self.eps_in

    # many charge vectors at once: arrays are (rows, ncol), one column per vector
    def rhs_batch(self, Q):
        Q = np.asfortranarray(_f64(Q).reshape(len(self.q), -1))
        ncol = Q.shape[1]
        R = np.zeros((self.nvec, ncol), order="F")
        _check(lib().fftsvd_problem_rhs_batch(self.h, ncol, _d(Q), _d(R)))
        return R

    def solve_batch(self, RHS, tol=1e-6, maxit=1000, precision=F64):
        """Returns (solutions, iterations per column, info per column)."""
        RHS = np.asfortranarray(_f64(RHS).reshape(self.nvec, -1))
        ncol = RHS.shape[1]
        X = np.zeros((self.nvec, ncol), order="F")
        its = np.zeros(ncol, dtype=np.int32)
        info = np.zeros(6 * ncol)
        _check(lib().fftsvd_problem_solve_batch(self.h, precision, ncol, _d(RHS), _d(X), tol, maxit, _i(its),
                                                _d(info)))
        keys = ("iterations", "refinements", "final_residual", "ms", "matvecs64", "matvecs32")
        return X, its, [dict(zip(keys, info[6 * c:6 * c + 6].tolist())) for c in range(ncol)]

    def potentials_batch(self, X):
        X = np.asfortranarray(_f64(X).reshape(self.nvec, -1))
        ncol = X.shape[1]
        Phi = np.zeros((max(len(self.q), 1), ncol), order="F")
        _check(lib().fftsvd_problem_potentials_batch(self.h, ncol, _d(X), _d(Phi)))
        return Phi[:len(self.q)]

    def stats(self):
        out = np.zeros(32)
        _check(lib().fftsvd_problem_stats(self.h, _d(out), 32))
        keys = ("panels", "regions", "setup_ms", "near_entries", "far_entries", "spectra_bins", "translate_macs",
                "depth", "leaves", "trees", "bytes_f64", "bytes_f32", "near_bytes_f64", "near_bytes_f32", "near_cols",
                "ut_entries", "vt_entries", "pv_entries", "uti_entries", "t_bins", "leaf_ublocks", "proj_chunks",
                "points_applied", "src_bins", "tgt_bins", "ms_tree", "ms_near", "ms_pinv_T", "ms_sampling",
                "ms_mgs", "ms_bases", "ms_plan")
        return dict(zip(keys, out.tolist()))

    def bench_matvec_batch(self, ncol, precision=F64, nrep=5):
        """ms per batched matvec of ncol columns, and per stage."""
        tot = C.c_double()
        parts = np.zeros(4)
        _check(lib().fftsvd_problem_bench_matvec_batch(self.h, precision, ncol, nrep, C.byref(tot), _d(parts)))
        return tot.value / nrep, dict(zip(("leaf", "project_source", "translate", "target"), parts.tolist()))

    def bench_matvec(self, precision=F64, nrep=20):
        tot = C.c_double()
        parts = np.zeros(4)
        _check(lib().fftsvd_problem_bench_matvec(self.h, precision, nrep, C.byref(tot), _d(parts)))
        return tot.value / nrep, dict(zip(("leaf", "project_source", "translate", "target"), parts.tolist()))


class Qual(Problem):
    """Qualocation operator (generateQualocationOperator ->
    generateQualocationOperatorPreconditioner -> generateQualocationRHS ->
    solveQualocationOperator -> collectQualocationPotentials): Laplace only,
    one unknown per panel.  The defaults are generateQualocationOperator's
    (32 panels per leaf, 4 grid points, SVD 1e-5); solveBEMproblem uses 27,
    3 and 1e-5."""

    def __init__(self, verts9, qpos, q, eps_in=4.0, eps_out=80.0, maxpanels=32, gridpoints=4, svdeps=1e-5,
                 want_f32=True):
        v = _f64(verts9).reshape(-1, 9)
        self.qpos = _f64(qpos).reshape(-1, 3)
        self.q = _f64(q).reshape(-1)
        self.eps_in = eps_in
        h = _vp()
        _check(lib().fftsvd_qual_create(_d(v), len(v), _d(self.qpos), _d(self.q), len(self.q), eps_in, eps_out,
                                        maxpanels, gridpoints, svdeps, int(want_f32), C.byref(h)))
        self.h = h
        self.n = len(v)
        nv = C.c_longlong()
        _check(lib().fftsvd_problem_unknowns(self.h, C.byref(nv)))
        self.nvec = nv.value

    def tree(self, which=0):
        return Tree(self, which)

    def potentials(self, sol):
        return self.reaction_potentials(sol)

    def energy(self, sol, conversion=332.112):
        """1/2 sum q phi * conversion / eps_in (FFTSVDsolvecfqual.c:153-167)."""
        return 0.5 * float(np.dot(self.q, self.reaction_potentials(sol))) * conversion / self.eps_in
