"""Host side of the reference-compatible interface (no GPU): readSRF
(FFTSVDpbeAPI.c:776-1032) reads the .srf grammar and the MSMS .vert/.face
files, flips dielectric faces (VertFace_readface_flip, VertFace.c:82-100)
and drops zero-area faces (VertFace_fix, VertFace.c:103-190)."""
import ctypes as C

import numpy as np

import refapi
from fftsvd_amd import meshes


class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class Panel(C.Structure):
    _fields_ = [("centroid", C.POINTER(Vec3)), ("normal", C.POINTER(Vec3)), ("area", C.c_double),
                ("maxedgelength", C.c_double), ("vertex", (C.c_double * 3) * 3)]


def panels_of(s, cls, i):
    arr = C.cast(s.ptrs[cls], C.POINTER(C.POINTER(C.POINTER(Panel))))[i]
    n = s.counts[cls][i]
    out = np.zeros((n, 9))
    for k in range(n):
        p = arr[k].contents
        out[k] = np.array([[p.vertex[a][d] for d in range(3)] for a in range(3)]).ravel()
    return out


def test_readsrf_flat_dielectric(tmp_path):
    V, F = meshes.ellipsoid((3.0, 2.2, 1.6), subdivisions=2, jitter=0.04, seed=3)
    # a degenerate face (repeated vertex) that must be dropped
    F = np.vstack([F, [[F[0, 0], F[0, 0], F[0, 1]]]])
    meshes.write_vertface(str(tmp_path / "blob"), V, F)
    refapi.write_srf(tmp_path / "op.srf", ["blob"])
    s = refapi.read_srf(tmp_path / "op.srf")
    assert s.num[0].value == 0 and s.num[1].value == 1 and s.num[2].value == 0 and s.num[3].value == 0
    expect = meshes.panels_from_mesh(V, F)
    assert s.total.value == len(expect) == len(F) - 1
    got = panels_of(s, 1, 0)
    np.testing.assert_allclose(got, expect, rtol=0, atol=1e-8)   # vertices are printed with %.9f
    assert s.parents[0][0] == 0


def test_readsrf_relative_paths(tmp_path):
    d = tmp_path / "geom"
    d.mkdir()
    V, F = meshes.icosphere(2.0, 1)
    meshes.write_vertface(str(d / "sph"), V, F)
    refapi.write_srf(d / "op.srf", ["sph"])
    s = refapi.read_srf(d / "op.srf")
    assert s.total.value == len(F)


def test_parameter_globals_are_exported():
    for name in ("innerdielectric", "outerdielectric", "ionicstrength", "ionexclusionradius", "errortolerance",
                 "proberadius", "dielectricvertexdensity", "saltvertexdensity"):
        refapi.set_global(name, 1.25)
        assert refapi.get_global(name) == 1.25
        refapi.set_global(name, 0.0)


def test_readsrf_cavities_and_999(tmp_path):
    """Cavity surfaces are read without the face flip (getFlatPanels,
    FFTSVDpbeAPI.c:696-699); a cavity parent of 999 takes that cavity's
    panels out of the total (FFTSVDpbeAPI.c:1006-1029)."""
    Vd, Fd = meshes.icosphere(3.0, 2)
    Vc, Fc = meshes.icosphere(1.0, 1)
    Vb, Fb = meshes.icosphere(0.5, 0, center=(1.5, 0.0, 0.0))
    for name, V, F in (("diel", Vd, Fd), ("cav", Vc, Fc[:, [0, 2, 1]]), ("bad", Vb, Fb[:, [0, 2, 1]])):
        meshes.write_vertface(str(tmp_path / name), V, F)
    refapi.write_srf(tmp_path / "op.srf", ["diel"], cavity_names=["cav", "bad"], cavity_parents=[0, 999])
    s = refapi.read_srf(tmp_path / "op.srf")
    assert s.num[1].value == 1 and s.num[2].value == 2 and s.num[3].value == 0
    assert s.total.value == len(Fd) + len(Fc)
    assert s.parents[1][0] == 0 and s.parents[1][1] == 999
    # cavity faces as written (inward CCW) become panels whose normals
    # point out of the cavity: the same panels as the flipped outward mesh
    Vw = np.loadtxt(tmp_path / "cav.vert")
    np.testing.assert_allclose(panels_of(s, 2, 0), meshes.panels_from_mesh(Vw, Fc), rtol=0, atol=1e-12)
