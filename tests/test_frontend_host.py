"""Front end of the reference interface on the host (no GPU): parameter file,
coordinate and table readers, radius/charge assignment, MODPDB/FRC writers
and flat surface generation (reference FFTSVDpbeAPI.c:36-672,1034-1049).
Inputs are written here in the formats the reference reads; expected values
follow from the format definitions (PDB fixed columns, DelPhi SIZ/CRG
columns) and the matching rules stated in the reference."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import refapi
from test_refapi_host import panels_of

L = refapi.lib()
for name in ("readParams", "checkParams", "readPDB", "readCRD", "readXYZR", "readSIZ", "readCRG", "writeMODPDB",
             "writeFRC", "assignRadiiCharges", "generateFlatSRF", "removeWhitespace"):
    getattr(L, name).restype = None
L.yesno.restype = C.c_int
L.yesno.argtypes = [C.c_char_p]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def pdb_line(rec, num, name, res, chain, resnum, x, y, z, occ=1.0, temp=0.0):
    """A PDB ATOM/HETATM record in the standard fixed columns."""
    nm = name if len(name) == 4 else " " + name.ljust(3)
    return "%-6s%5d %-4s %3s %1s%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n" % (rec, num, nm, res, chain, resnum, x, y, z, occ,
                                                                        temp)


def read_entries(fn, path):
    n = C.c_uint()
    p = C.POINTER(refapi.PDBentry)()
    getattr(L, fn)(str(path).encode(), C.byref(n), C.byref(p))
    return p, n.value


def test_yesno_and_whitespace():
    for s, v in [(b"t", 1), (b"YES", 1), (b"1", 1), (b"true", 1), (b"n", 0), (b"False", 0), (b"0", 0), (b"", 0)]:
        assert L.yesno(s) == v
    buf = C.create_string_buffer(b" a b\tc \n")
    L.removeWhitespace(buf)
    assert buf.value == b"abc"


def test_read_params(tmp_path):
    f = tmp_path / "p.prm"
    f.write_text("indi=4\nexdi = 80.0\nsalt=0.145\nionrad=0\nprbrad=1.4\nde=1e-6\ndvert=2.5\nsvert=0.5\n"
                 "energy(g)\nout(frc)\nout(modpdb)\naddcoul=yes\n")
    for g in ("outputfrc", "outputmodpdb", "addcoulombic"):
        refapi.set_global(g, 0, C.c_uint)
    L.readParams(str(f).encode())
    assert refapi.get_global("innerdielectric") == 4.0
    assert refapi.get_global("outerdielectric") == 80.0
    assert refapi.get_global("ionicstrength") == 0.145
    assert refapi.get_global("proberadius") == 1.4
    assert refapi.get_global("errortolerance") == 1e-6
    assert refapi.get_global("dielectricvertexdensity") == 2.5
    assert refapi.get_global("saltvertexdensity") == 0.5
    assert refapi.get_global("energycalculation", C.c_char) == b"G"
    assert refapi.get_global("outputfrc", C.c_uint) == 1
    assert refapi.get_global("outputmodpdb", C.c_uint) == 1
    assert refapi.get_global("addcoulombic", C.c_uint) == 1
    refapi.set_global("addcoulombic", 0, C.c_uint)


@pytest.mark.parametrize("bad,msg", [("indi=0.5", "Inner dielectric must be >= 1.0"),
                                     ("salt=-1", "Ionic strength must be >= 0.0"),
                                     ("de=0", "Error tolerance must be >= 0.0"),
                                     ("addcoul=maybe", "Invalid parameter boolean expression")])
def test_check_params_errors_exit(tmp_path, bad, msg):
    """error() prints 'ERROR: ...' and exits with status -1 (255), as the
    reference does; run in a child process."""
    f = tmp_path / "p.prm"
    f.write_text("indi=4\nexdi=80\nsalt=0\nprbrad=1.4\nde=1e-6\nenergy(g)\n" + bad + "\n")
    code = ("import sys; sys.path.insert(0, %r); sys.path.insert(0, %r); import refapi; L = refapi.lib(); "
            "L.readParams(%r); L.checkParams(); print('no error')" % (ROOT, os.path.join(ROOT, "tests"),
                                                                    str(f).encode()))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert ("ERROR: " + msg) in r.stdout


def test_read_pdb_columns(tmp_path):
    f = tmp_path / "m.pdb"
    f.write_text("HEADER test\n" + pdb_line("ATOM", 1, "N", "ALA", "A", 5, 1.5, -2.25, 3.125, 1.0, 12.5) +
                 "TER\n" + pdb_line("HETATM", 12, "CA", "HEM", "B", 201, -10.0, 0.001, 99.999) + "END\n")
    p, n = read_entries("readPDB", f)
    assert n == 2
    a, b = p[0], p[1]
    assert (a.record, a.atomnumber, a.atomname, a.residuename, a.chain, a.residuenumber) == \
        (b"ATOM", 1, b"N", b"ALA", b"A", 5)
    assert (a.x, a.y, a.z, a.occupancy, a.temperature) == (1.5, -2.25, 3.125, 1.0, 12.5)
    assert (b.record, b.atomnumber, b.atomname, b.residuename, b.chain, b.residuenumber) == \
        (b"HETATM", 12, b"CA", b"HEM", b"B", 201)
    assert (b.x, b.y, b.z) == (-10.0, 0.001, 99.999)
    refapi.free(p)


def test_read_crd_and_xyzr(tmp_path):
    crd = tmp_path / "m.crd"
    # CHARMM card format: atom, residue number, residue, atom name, x, y, z, segment (chain at 51), weight
    rows = ["%5d%5d %-4s %-4s%10.5f%10.5f%10.5f %-4s %-4s %10.5f\n" % (1, 1, "ALA", "N", 1.0, 2.0, 3.0, "A", "1", 0.5),
            "%5d%5d %-4s %-4s%10.5f%10.5f%10.5f %-4s %-4s %10.5f\n" % (2, 1, "ALA", "CA", -1.0, 0.5, 7.25, "A", "1", 0.0)]
    crd.write_text("* title\n*\n    2\n" + "".join(rows))
    p, n = read_entries("readCRD", crd)
    assert n == 2
    assert (p[0].atomnumber, p[0].residuenumber, p[0].residuename, p[0].atomname) == (1, 1, b"ALA", b"N")
    assert (p[1].x, p[1].y, p[1].z) == (-1.0, 0.5, 7.25)
    assert p[0].chain == b"A"
    xyzr = tmp_path / "m.xyzr"
    xyzr.write_text("0 0 0 1.5\n1 2 3 2.0\n")
    q, m = read_entries("readXYZR", xyzr)
    assert m == 2 and (q[1].x, q[1].y, q[1].z, q[1].radius) == (1.0, 2.0, 3.0, 2.0)


def write_tables(tmp_path):
    siz = tmp_path / "r.siz"
    # atom(6) residue(3) radius(8)
    siz.write_text("atom__res_radius_\n!comment\n" + "".join(
        "%-6s%-3s%8.3f\n" % x for x in [("CA", "ALA", 1.90), ("C", "", 1.70), ("CB", "", 1.80), ("O", "", 1.52),
                                         ("N", "", 1.55)]))
    crg = tmp_path / "c.crg"
    # atom(6) residue(3) resnum(4) chain(1) charge(8)
    crg.write_text("atom__resnumbc_charge_\n" + "".join(
        "%-6s%-3s%-4s%1s%8.3f\n" % x for x in [("N", "ALA", "5", "A", -0.30), ("N", "ALA", "", "", -0.10),
                                                ("CA", "", "", "", 0.25), ("O", "ALA", "5", "", -0.55)]))
    return siz, crg


def test_assign_radii_charges_match_levels(tmp_path):
    siz, crg = write_tables(tmp_path)
    f = tmp_path / "m.pdb"
    f.write_text(pdb_line("ATOM", 1, "N", "ALA", "A", 5, 0, 0, 0) + pdb_line("ATOM", 2, "CA", "ALA", "A", 5, 1, 0, 0) +
                 pdb_line("ATOM", 3, "CB", "GLY", "B", 6, 2, 0, 0) + pdb_line("ATOM", 4, "O", "ALA", "B", 5, 3, 0, 0) +
                 pdb_line("ATOM", 5, "N", "ALA", "B", 9, 4, 0, 0) + pdb_line("ATOM", 6, "SG", "CYS", "A", 7, 5, 0, 0))
    p, n = read_entries("readPDB", f)
    ns, nc = C.c_uint(), C.c_uint()
    ps, pc = C.POINTER(refapi.SIZentry)(), C.POINTER(refapi.CRGentry)()
    L.readSIZ(str(siz).encode(), C.byref(ns), C.byref(ps))
    L.readCRG(str(crg).encode(), C.byref(nc), C.byref(pc))
    assert ns.value == 5 and nc.value == 4
    assert (ps[0].atomlabel, ps[0].residuelabel, ps[0].radius) == (b"CA", b"ALA", 1.9)
    assert (pc[0].atomlabel, pc[0].residuelabel, pc[0].residuenumber, pc[0].chain, pc[0].charge) == \
        (b"N", b"ALA", 5, b"A", -0.3)
    L.assignRadiiCharges(p, n, ps, ns, pc, nc)
    rad = [p[i].radius for i in range(n)]
    chg = [p[i].charge for i in range(n)]
    # N: residue-less "N" prefix; CA: exact (CA, ALA); CB: the first residue-less entry whose label is a
    # prefix of the atom name wins, here "C" (listed before "CB"); O: residue-less "O"; SG: no entry -> 0.0
    assert rad == [1.55, 1.90, 1.70, 1.52, 1.55, 0.0]
    # N (A,5,ALA): level 8; CA: atom-only level 1; CB: none -> 0; O (B,5,ALA): residue+number+atom level 7;
    # N (B,9,ALA): residue+atom level 5, first reached by the (N, ALA, 5, A) entry, which a later entry of
    # the same level does not replace; SG: none -> 0
    assert chg == [-0.30, 0.25, 0.0, -0.55, -0.30, 0.0]


def test_writers(tmp_path):
    f = tmp_path / "m.pdb"
    f.write_text(pdb_line("ATOM", 7, "CA", "ALA", "A", 5, 1.5, -2.25, 3.125))
    p, n = read_entries("readPDB", f)
    p[0].radius, p[0].charge, p[0].potential = 1.9, 0.25, -3.5
    refapi.set_global("innerdielectric", 4.0)
    refapi.set_global("outerdielectric", 80.0)
    L.writeFRC(str(tmp_path / "o.frc").encode(), p, n)
    L.writeMODPDB(str(tmp_path / "o.pdb").encode(), p, n)
    frc = (tmp_path / "o.frc").read_text().splitlines()
    assert frc[0] == "FFTSVDpbe SITE POTENTIAL FILE"
    assert frc[2] == "inner,outer dielectric: %8f %8f" % (4.0, 80.0)
    assert frc[12] == "%9.3f%9.3f%9.3f%8.4f%15.6f  %9.4f%9.4f%9.4f" % (1.5, -2.25, 3.125, 0.25, -3.5, 0, 0, 0)
    assert frc[13].startswith("total energy =")
    mod = (tmp_path / "o.pdb").read_text().splitlines()
    assert mod[3] == "%-6s%5d %4s%1s%3s %1s%4d%1s   %8.3f%8.3f%8.3f%6.2f%7.3f" % (
        "ATOM", 7, "CA", " ", "ALA", "A", 5, " ", 1.5, -2.25, 3.125, 1.9, 0.25)


def test_generate_flat_srf(tmp_path):
    siz, _ = write_tables(tmp_path)
    f = tmp_path / "m.pdb"
    f.write_text(pdb_line("ATOM", 1, "N", "ALA", "A", 1, 0, 0, 0) + pdb_line("ATOM", 2, "CA", "ALA", "A", 1, 1.45, 0, 0) +
                 pdb_line("ATOM", 3, "O", "ALA", "A", 1, 2.2, 1.2, 0))
    refapi.set_global("proberadius", 1.4)
    refapi.set_global("dielectricvertexdensity", 4.0)
    refapi.set_global("ionicstrength", 0.0)
    srf = tmp_path / "op.srf"
    L.generateFlatSRF(str(f).encode(), str(siz).encode(), str(srf).encode())
    s = refapi.read_srf(srf)
    assert s.num[1].value == 1 and s.total.value > 100
    P = panels_of(s, 1, 0).reshape(-1, 3, 3)
    # closed surface: sum of outward area vectors vanishes; enclosed volume
    # (divergence theorem) is positive for outward normals (the reference's
    # panel order after the flip is v0, v2, v1 of the MSMS face)
    n = np.cross(P[:, 2] - P[:, 0], P[:, 1] - P[:, 0]) / 2
    assert np.abs(n.sum(axis=0)).max() < 1e-8 * np.abs(n).sum()
    vol = np.einsum("ij,ij->", P.mean(axis=1), n) / 3
    assert vol > 20.0


def _closed_outward(P):
    """(area-vector sum relative to total area, enclosed volume) of panels
    in the reference's order (v0, v2, v1 of the MSMS face)."""
    n = np.cross(P[:, 2] - P[:, 0], P[:, 1] - P[:, 0]) / 2
    return np.abs(n.sum(axis=0)).max() / np.abs(n).sum(), np.einsum("ij,ij->", P.mean(axis=1), n) / 3


def test_union_mesh_grown_spheres():
    """fftsvd_mesh_union: the surface of the union of spheres grown by g.
    One atom: every vertex at r + g (to the grid's interpolation error) and
    the enclosed volume that of the sphere; two overlapping atoms: closed,
    outward, vertices on the grown spheres (away from their crease)."""
    from fftsvd_amd import meshes
    r, g, h = 1.7, 2.0, 0.35
    V, F = meshes.union_mesh(np.zeros((1, 3)), [r], g, h)
    d = np.linalg.norm(V, axis=1)
    assert np.abs(d - (r + g)).max() < 0.05 * h + 1e-3
    P = meshes.panels_from_mesh(V, F).reshape(-1, 3, 3)
    closure, vol = _closed_outward(P)
    assert closure < 1e-10
    assert abs(vol - 4 / 3 * np.pi * (r + g) ** 3) < 0.01 * vol
    c = np.array([[0, 0, 0], [2.5, 0, 0]])
    V, F = meshes.union_mesh(c, [1.5, 1.2], 1.0, h)
    R = np.array([2.5, 2.2])
    dist = np.abs((np.linalg.norm(V[:, None, :] - c[None], axis=2) - R[None]).min(axis=1))
    # linear interpolation cuts the crease where the spheres meet: there
    # the contour lies inside by up to a quarter cell
    assert np.quantile(dist, 0.95) < 0.05 * h + 1e-3 and dist.max() < 0.3 * h
    closure, vol = _closed_outward(meshes.panels_from_mesh(V, F).reshape(-1, 3, 3))
    assert closure < 1e-10 and vol > 0


def test_generate_flat_srf_stern(tmp_path):
    """With salt and ionexclusionradius > 0, generateFlatSRF also writes the
    ion-exclusion surface (atom spheres grown by the radius) as the salt
    surface of the .srf, enclosing the dielectric."""
    siz, _ = write_tables(tmp_path)
    f = tmp_path / "m.pdb"
    f.write_text(pdb_line("ATOM", 1, "N", "ALA", "A", 1, 0, 0, 0) + pdb_line("ATOM", 2, "CA", "ALA", "A", 1, 1.45, 0, 0) +
                 pdb_line("ATOM", 3, "O", "ALA", "A", 1, 2.2, 1.2, 0))
    refapi.set_global("proberadius", 1.4)
    refapi.set_global("dielectricvertexdensity", 4.0)
    refapi.set_global("saltvertexdensity", 2.0)
    refapi.set_global("ionicstrength", 0.145)
    refapi.set_global("ionexclusionradius", 2.0)
    srf = tmp_path / "op.srf"
    try:
        L.generateFlatSRF(str(f).encode(), str(siz).encode(), str(srf).encode())
    finally:
        refapi.set_global("ionicstrength", 0.0)
        refapi.set_global("ionexclusionradius", 0.0)
    assert srf.read_text().splitlines()[2] == "op.srf.s0"
    s = refapi.read_srf(srf)
    assert s.num[0].value == 1 and s.num[1].value == 1
    assert s.parents[0][0] == 0
    Ps = panels_of(s, 0, 0).reshape(-1, 3, 3)
    Pd = panels_of(s, 1, 0).reshape(-1, 3, 3)
    assert s.total.value == len(Ps) + len(Pd)
    cs, vs = _closed_outward(Ps)
    cd, vd = _closed_outward(Pd)
    assert cs < 1e-8 and cd < 1e-8 and vs > vd > 20.0
    # every salt vertex is outside the dielectric surface's bounding box grown by 1 A
    lo, hi = Pd.reshape(-1, 3).min(axis=0), Pd.reshape(-1, 3).max(axis=0)
    vs_pts = Ps.reshape(-1, 3)
    inside = np.all((vs_pts > lo + 1.0) & (vs_pts < hi - 1.0), axis=1)
    assert not inside.any()


def hollow_shell(n=400, R=7.0):
    """Atom centres on a Fibonacci sphere: a closed shell around a void the
    probe fits in."""
    i = np.arange(n) + 0.5
    phi, th = np.arccos(1 - 2 * i / n), np.pi * (1 + 5 ** 0.5) * i
    return np.c_[R * np.cos(th) * np.sin(phi), R * np.sin(th) * np.sin(phi), R * np.cos(phi)]


def test_generate_flat_srf_with_cavity(tmp_path):
    """A shell of atoms around a void: the mesher's components give one
    dielectric surface (positive enclosed volume) and one dielectric cavity
    (negative), written as meshmaker writes MSMS's components
    (meshmaker.c:119-262,600-690): the dielectric's parent 999 without a
    salt surface, the cavity's parent 0 (all its vertices inside the
    dielectric).  readSRF reads the cavity without the face flip, so its
    panels enclose the void with the normals pointing out of it."""
    xyz = hollow_shell()
    (tmp_path / "m.xyzr").write_text("".join("%.3f %.3f %.3f 1.8\n" % tuple(p) for p in xyz))
    refapi.set_global("proberadius", 1.4)
    refapi.set_global("dielectricvertexdensity", 2.0)
    refapi.set_global("ionicstrength", 0.0)
    srf = tmp_path / "op.srf"
    L.generateFlatSRF(str(tmp_path / "m.xyzr").encode(), b"", str(srf).encode())
    lines = srf.read_text().split("\n")
    assert lines[2:9] == ["", "op.srf.d0", "op.srf.c0", "", "999", "0", ""]
    s = refapi.read_srf(srf)
    assert [x.value for x in s.num] == [0, 1, 1, 0]
    assert s.parents[1][0] == 0
    Pd, Pc = panels_of(s, 1, 0).reshape(-1, 3, 3), panels_of(s, 2, 0).reshape(-1, 3, 3)
    assert s.total.value == len(Pd) + len(Pc)
    closed_d, vol_d = _closed_outward(Pd)
    closed_c, vol_c = _closed_outward(Pc)
    assert closed_d < 1e-8 and closed_c < 1e-8
    # the void: probe centres within 7 - 1.8 - 1.4 = 3.8 of the centre, the
    # cavity surface about 1.4 beyond them (a sphere of 5.2 would hold 589)
    assert 250 < vol_c < 650
    assert vol_d > 4.0 / 3.0 * np.pi * 7.0 ** 3
    rc = np.linalg.norm(Pc.mean(axis=1), axis=1)
    assert rc.max() < 5.6 and rc.min() > 3.5
