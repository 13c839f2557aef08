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
