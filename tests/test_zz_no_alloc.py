"""The allocation-free entry point fftsvd_integrate_noalloc against
fftsvd_integrate, through the C-ABI.

Device buffers come from the HIP runtime the library itself links
(libamdhip64.so.7, loaded after the library so both resolve to one
runtime instance); the entry point only launches kernels on the stream it
is given.  Two calls in a row, the second on a smaller input, each equal
to fftsvd_integrate's output (the same integral kernel on the same panel
records, which the device builds with the host's arithmetic); panel
counts above the fixed maximum are rejected before any GPU work.
"""
import ctypes as C

import numpy as np
import pytest

import fftsvd_amd as fa
from fftsvd_amd import meshes

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _hip():
    fa.lib()   # the library's HIP runtime first, so the soname below is the same instance
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipStreamSynchronize):
        f.restype = C.c_int
    return h


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(int(nbytes), 8)) == 0
        self.nbytes = int(nbytes)

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape):
        out = np.zeros(shape)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _case(sub, npts, seed):
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=sub, jitter=0.04, seed=seed)
    P = meshes.panels_from_mesh(V, F)
    rng = np.random.default_rng(seed)
    cent = P.reshape(-1, 3, 3).mean(axis=1)
    pts = np.vstack([cent[: npts // 2], rng.normal(scale=2.5, size=(npts - npts // 2, 3))])
    return P, pts


def _noalloc(hip, P, pts, kernel, kappa, bufs):
    L = fa.lib()
    f = L.fftsvd_integrate_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    dv, dp, ds, dd = bufs
    dv.put(P)
    dp.put(pts)
    rc = f(kernel, kappa, dv.p, len(P), dp.p, len(pts), ds.p, dd.p, None)
    assert rc == 0, L.fftsvd_last_error()
    assert hip.hipStreamSynchronize(None) == 0
    return ds.get((len(pts), len(P))), dd.get((len(pts), len(P)))


def test_noalloc_matches_integrate_twice_shrinking():
    hip = _hip()
    big_P, big_pts = _case(4, 96, 3)        # 5120 panels
    small_P, small_pts = _case(2, 40, 5)    # 320 panels
    # buffers sized for the first call and reused by the second
    bufs = [DevArray(hip, big_P.nbytes), DevArray(hip, big_pts.nbytes),
            DevArray(hip, 8 * len(big_P) * len(big_pts)), DevArray(hip, 8 * len(big_P) * len(big_pts))]
    try:
        for P, pts, kernel, kappa in ((big_P, big_pts, fa.YUKAWA, 0.124), (small_P, small_pts, fa.LAPLACE, 0.0)):
            s_ref, d_ref = fa.integrate(P, pts, kernel, kappa)
            s, d = _noalloc(hip, P, pts, kernel, kappa, bufs)
            ds = np.abs(s - s_ref).max() / np.abs(s_ref).max()
            dd = np.abs(d - d_ref).max() / np.abs(d_ref).max()
            print("noalloc %d panels x %d points: max rel diff slp %.2e dlp %.2e, bitwise equal %s" %
                  (len(P), len(pts), ds, dd, bool(np.array_equal(s, s_ref) and np.array_equal(d, d_ref))))
            assert ds <= 1e-12 and dd <= 1e-12
    finally:
        for b in bufs:
            b.free()


def test_noalloc_rejects_oversize():
    L = fa.lib()
    f = L.fftsvd_integrate_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rc = f(0, 0.0, None, 16385, None, 1, None, None, None)
    assert rc != 0
    assert b"exceed the maximum of 16384" in L.fftsvd_last_error()
