"""Synthetic surface meshes in the MSMS convention used by the reference.

A mesh is (verts (V,3) float64, faces (F,3) int, 0-based) with faces
counter-clockwise seen from outside, i.e. (v1-v0) x (v2-v0) points out of
the body -- the orientation MSMS writes and that the reference's
VertFace_readface_flip (VertFace.c:82-100) turns into panels whose analytic
normal (FlatPanel.c:41-49) points outward.
"""
import numpy as np


def icosphere(radius=1.0, subdivisions=0, center=(0.0, 0.0, 0.0)):
    """Icosahedron refined `subdivisions` times, vertices projected to the sphere."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(center, dtype=np.float64)
    F = np.array(faces, dtype=np.int64)
    # orient outward
    c = V[F].mean(axis=1) - np.asarray(center)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    bad = np.einsum("ij,ij->i", n, c) < 0
    F[bad] = F[bad][:, [0, 2, 1]]
    return V, F


def ellipsoid(axes=(1.0, 1.0, 1.0), subdivisions=2, center=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    """Icosphere scaled to an ellipsoid, optionally with radial vertex jitter."""
    V, F = icosphere(1.0, subdivisions)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        V = V * (1.0 + jitter * rng.uniform(-1, 1, size=(len(V), 1)))
    V = V * np.asarray(axes, dtype=np.float64) + np.asarray(center, dtype=np.float64)
    return V, F


def panels_from_mesh(V, F, flip=True):
    """(F, 9) panel vertex array as the reference builds it from a face list:
    with flip (dielectric and salt surfaces) face "a b c" becomes panel
    (a, c, b) (VertFace_readface_flip, VertFace.c:82-100); zero-area faces
    are dropped (VertFace_fix, VertFace.c:103-190)."""
    V = np.asarray(V, dtype=np.float64)
    F = np.asarray(F)
    order = [0, 2, 1] if flip else [0, 1, 2]
    P = V[F[:, order]]
    d1 = np.linalg.norm(P[:, 0] - P[:, 1], axis=1)
    d2 = np.linalg.norm(P[:, 1] - P[:, 2], axis=1)
    d3 = np.linalg.norm(P[:, 2] - P[:, 0], axis=1)
    s = 0.5 * (d1 + d2 + d3)
    with np.errstate(invalid="ignore"):
        area = np.sqrt(s * (s - d1) * (s - d2) * (s - d3))
    keep = area > 0.0
    return np.ascontiguousarray(P[keep].reshape(-1, 9))


def write_vertface(basename, V, F, genus=None):
    """Write MSMS-style .vert / .face files (1-based faces)."""
    with open(basename + ".vert", "w") as f:
        for p in V:
            f.write("%.9f %.9f %.9f\n" % (p[0], p[1], p[2]))
    g = np.ones(len(F), dtype=np.int64) if genus is None else genus
    with open(basename + ".face", "w") as f:
        for (a, b, c), gg in zip(F, g):
            f.write("%d %d %d %d\n" % (a + 1, b + 1, c + 1, gg))


def synthetic_globule(natoms=2000, seed=20261202, step=1.5, minsep=2.4, radius_range=(1.2, 2.0),
                      charge_range=(-0.8, 0.8), confine=26.0):
    """Seeded self-avoiding random walk of `natoms` atoms (step 1.5 A, every
    non-bonded pair at least 2.4 A apart), confined to a sphere of radius `confine` so
    the chain folds into a globule; when none of 64 random directions is
    free the walk continues from a random earlier atom.  Radii uniform
    in radius_range; charges uniform in charge_range, each shifted by
    (round(sum q) - sum q)/natoms so the net charge is an integer.
    Returns (xyz (n,3), radii (n,), charges (n,))."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((natoms, 3))
    n = 1
    cur = 0
    m2 = minsep * minsep
    while n < natoms:
        # 64 candidate directions at once; the first free one is taken
        d = rng.normal(size=(64, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        cand = pos[cur] + step * d
        ok = np.ones(64, dtype=bool)
        if confine is not None:
            ok &= np.einsum("ij,ij->i", cand, cand) <= confine * confine
        d2 = ((cand[:, None, :] - pos[None, :n, :]) ** 2).sum(-1)
        d2[:, cur] = m2  # the bonded predecessor sits one step (1.5 A) away
        ok &= (d2 >= m2).all(axis=1)
        hit = np.flatnonzero(ok)
        if len(hit):
            pos[n] = cand[hit[0]]
            cur = n
            n += 1
        else:
            cur = int(rng.integers(n))
    radii = rng.uniform(radius_range[0], radius_range[1], natoms)
    q = rng.uniform(charge_range[0], charge_range[1], natoms)
    q += (np.round(q.sum()) - q.sum()) / natoms
    return pos, radii, q


def ses_mesh(xyz, radii, probe=1.4, spacing=0.6):
    """Solvent-excluded surface triangulated by the library's mesher
    (fftsvd_mesh_ses); returns (verts, faces) in the MSMS convention."""
    return _mesh("fftsvd_mesh_ses", xyz, radii, probe, spacing)


def union_mesh(xyz, radii, grow, spacing=0.6):
    """Surface of the union of the atom spheres grown by `grow` (the
    ion-exclusion surface of the Stern layer), by the library's mesher
    (fftsvd_mesh_union); (verts, faces) in the MSMS convention."""
    return _mesh("fftsvd_mesh_union", xyz, radii, grow, spacing)


def _mesh(entry, xyz, radii, probe, spacing):
    import ctypes as C

    from fftsvd_amd import lib

    L = lib()
    f = getattr(L, entry)
    f.restype = C.c_int
    f.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double, C.POINTER(C.POINTER(C.c_double)),
                  C.POINTER(C.c_int), C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]
    L.fftsvd_free.argtypes = [C.c_void_p]
    L.fftsvd_free.restype = None
    a = np.ascontiguousarray(np.column_stack([xyz, radii]), dtype=np.float64)
    pv = C.POINTER(C.c_double)()
    pf = C.POINTER(C.c_int)()
    nv, nf = C.c_int(), C.c_int()
    rc = f(a.ctypes.data_as(C.POINTER(C.c_double)), len(a), probe, spacing, C.byref(pv), C.byref(nv), C.byref(pf),
           C.byref(nf))
    if rc != 0:
        raise RuntimeError("mesher failed (%s)" % entry)
    V = np.ctypeslib.as_array(pv, shape=(nv.value * 3,)).reshape(-1, 3).copy()
    F = np.ctypeslib.as_array(pf, shape=(nf.value * 3,)).reshape(-1, 3).copy().astype(np.int64)
    L.fftsvd_free(C.cast(pv, C.c_void_p))
    L.fftsvd_free(C.cast(pf, C.c_void_p))
    return V, F


def ses_mesh_target(xyz, radii, target_faces, probe=1.4, spacing=0.6, tol=0.05, max_rounds=6):
    """SES mesh whose face count is within `tol` of `target_faces`, found by
    rescaling the grid spacing (faces scale like 1/spacing^2)."""
    best = None
    for _ in range(max_rounds):
        V, F = ses_mesh(xyz, radii, probe, spacing)
        best = (V, F, spacing)
        ratio = len(F) / float(target_faces)
        if abs(ratio - 1.0) <= tol:
            break
        spacing *= ratio ** 0.5
    return best
