"""The reference-style driver (fftsvd_amd/apps/fftsvd_pbe.c, a plain C
program linked against libfftsvd_hip.so through include/FFTSVDpbeAPI.h)
run on DelPhi unit files: parameters, radii and charges assigned from
SIZ/CRG tables, surface from generateFlatSRF, solve, FRC output.  The FRC
potentials are checked against the extended API on the same mesh."""
import os
import subprocess

import numpy as np
import pytest

import fftsvd_amd as fa
from fftsvd_amd import build, meshes
from test_frontend_host import pdb_line

pytestmark = pytest.mark.gpu

NAMES = ["N", "CA", "C", "O"]
RADII = {"N": 1.55, "CA": 1.90, "C": 1.70, "O": 1.52}
CHARGES = {"N": -0.40, "CA": 0.10, "C": 0.55, "O": -0.55}


def write_inputs(d, salt, ionrad=0.0, extra=""):
    xyz, _, _ = meshes.synthetic_globule(natoms=24, seed=5, confine=6.0)
    xyz = np.round(xyz, 3)   # the precision PDB columns carry
    lines = []
    for i, p in enumerate(xyz):
        nm = NAMES[i % 4]
        lines.append(pdb_line("ATOM", i + 1, nm, "ALA", "A", i // 4 + 1, *p))
    (d / "fort.13").write_text("".join(lines) + "END\n")
    (d / "fort.15").write_text("".join(lines) + "END\n")
    (d / "fort.11").write_text("atom__res_radius_\n" + "".join("%-6s%-3s%8.3f\n" % (k, "", v) for k, v in RADII.items()))
    (d / "fort.12").write_text("atom__resnumbc_charge_\n" +
                               "".join("%-6s%-3s%-4s%1s%8.3f\n" % (k, "", "", "", v) for k, v in CHARGES.items()))
    (d / "fort.10").write_text("indi=4.0\nexdi=80.0\nsalt=%g\nionrad=%g\nprbrad=1.4\nde=1e-6\ndvert=2.0\n"
                               "svert=1.0\nenergy(g)\nout(frc)\n%s" % (salt, ionrad, extra))
    return xyz


@pytest.mark.parametrize("salt", [0.0, 0.145])
def test_driver_end_to_end(tmp_path, salt):
    xyz = write_inputs(tmp_path, salt)
    r = subprocess.run([build.PBE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    E_drv = float([l for l in r.stdout.splitlines() if l.startswith("solvation energy")][0].split(":")[1].split()[0])
    frc = (tmp_path / "fort.16").read_text().splitlines()
    rows = [l for l in frc[12:] if not l.startswith("total energy")]
    assert len(rows) == len(xyz)
    pot = np.array([float(l[35:50]) for l in rows])
    q = np.array([CHARGES[NAMES[i % 4]] for i in range(len(xyz))])
    # the same surface through the extended API
    V = np.loadtxt(tmp_path / "operator.srf.d0.vert")
    F = np.loadtxt(tmp_path / "operator.srf.d0.face", dtype=np.int64)[:, :3] - 1
    P = meshes.panels_from_mesh(V, F)
    kappa = np.sqrt(salt) / 3.047
    prob = fa.Problem(P, xyz, q, 4.0, 80.0, kappa)
    x, it, hist, info = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    ref = 561.0 * prob.reaction_potentials(x) / 4.0
    np.testing.assert_allclose(pot, ref, rtol=0, atol=2e-6 * np.abs(ref).max() + 1e-6)   # FRC prints 6 decimals
    E_ref = 0.5 * np.dot(q, ref) * 332.112 / 561.0
    assert abs(E_drv - E_ref) <= 1e-8 * abs(E_ref) + 1e-9
    assert E_ref < 0


def test_driver_stern_layer(tmp_path):
    """salt > 0 and ionrad > 0: generateFlatSRF meshes the ion-exclusion
    surface too and the solve uses the Stern-layer operator; the FRC
    potentials equal the extended API's Stern problem on the same two
    surfaces, and differ from the same dielectric without the layer."""
    salt = 0.145
    xyz = write_inputs(tmp_path, salt, ionrad=2.0)
    r = subprocess.run([build.PBE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    frc = (tmp_path / "fort.16").read_text().splitlines()
    rows = [l for l in frc[12:] if not l.startswith("total energy")]
    pot = np.array([float(l[35:50]) for l in rows])
    q = np.array([CHARGES[NAMES[i % 4]] for i in range(len(xyz))])

    def panels(tag):
        V = np.loadtxt(tmp_path / ("operator.srf.%s.vert" % tag))
        F = np.loadtxt(tmp_path / ("operator.srf.%s.face" % tag), dtype=np.int64)[:, :3] - 1
        return meshes.panels_from_mesh(V, F)

    Pd, Ps = panels("d0"), panels("s0")
    kappa = np.sqrt(salt) / 3.047
    prob = fa.Problem(Pd, xyz, q, 4.0, 80.0, kappa, stern_verts9=Ps)
    x, it, hist, info = prob.solve(prob.rhs(), 1e-6, precision=fa.F64)
    ref = 561.0 * prob.reaction_potentials(x) / 4.0
    np.testing.assert_allclose(pot, ref, rtol=0, atol=2e-6 * np.abs(ref).max() + 1e-6)
    plain = fa.Problem(Pd, xyz, q, 4.0, 80.0, kappa)
    x0, _, _, _ = plain.solve(plain.rhs(), 1e-6, precision=fa.F64)
    ref0 = 561.0 * plain.reaction_potentials(x0) / 4.0
    E, E0 = np.dot(q, ref), np.dot(q, ref0)
    assert E < 0 and E0 < 0 and abs(E - E0) > 1e-4 * abs(E0)


def test_driver_qualocation(tmp_path):
    """usequal=true: the driver runs the qualocation formulation; the FRC
    potentials equal the extended API's qualocation problem on the same
    surface, and the energy agrees with the first-kind formulation to the
    discretisation error (a few percent on this 1060-panel surface, with
    atoms close to it)."""
    xyz = write_inputs(tmp_path, 0.0, extra="usequal=true\n")
    r = subprocess.run([build.PBE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "qualocation operator" in r.stdout
    frc = (tmp_path / "fort.16").read_text().splitlines()
    rows = [l for l in frc[12:] if not l.startswith("total energy")]
    pot = np.array([float(l[35:50]) for l in rows])
    q = np.array([CHARGES[NAMES[i % 4]] for i in range(len(xyz))])
    V = np.loadtxt(tmp_path / "operator.srf.d0.vert")
    F = np.loadtxt(tmp_path / "operator.srf.d0.face", dtype=np.int64)[:, :3] - 1
    P = meshes.panels_from_mesh(V, F)
    qual = fa.Qual(P, xyz, q, 4.0, 80.0)
    x, it, hist, info = qual.solve(qual.rhs(), 1e-6, precision=fa.F64)
    ref = 561.0 * qual.potentials(x) / 4.0
    np.testing.assert_allclose(pot, ref, rtol=0, atol=2e-6 * np.abs(ref).max() + 1e-6)
    first = fa.Problem(P, xyz, q, 4.0, 80.0, 0.0)
    x1, _, _, _ = first.solve(first.rhs(), 1e-6, precision=fa.F64)
    E, E1 = np.dot(q, ref), np.dot(q, 561.0 * first.reaction_potentials(x1) / 4.0)
    assert E < 0 and abs(E - E1) < 0.05 * abs(E1)
