"""Writes globule100k_oracle.json: the FP64 oracle's solve of the
100k-panel synthetic globule (2000-atom seeded walk, SES 1.4 A probe,
eps 4/80, kappa 0.124, GMRES tol 1e-6), the fixture behind
tests/test_gpu_configs.py.  CPU only, single core, about five minutes.

    python tests/golden/make_globule100k.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from fftsvd_amd import meshes  # noqa: E402

SEED, ATOMS, TARGET, KAPPA = 20261202, 2000, 100000, 0.124


def main():
    xyz, radii, q = meshes.synthetic_globule(natoms=ATOMS, seed=SEED)
    # the throughput workload's mesh: spacing rescaled until ~100k faces
    V, F, h = meshes.ses_mesh_target(xyz, radii, TARGET, probe=1.4, spacing=1.3)
    P = meshes.panels_from_mesh(V, F)
    digest = hashlib.sha256(np.ascontiguousarray(P, dtype=np.float64).tobytes()).hexdigest()
    print("mesh: %d panels, spacing %.17g, sha256 %s" % (len(P), h, digest), flush=True)
    t0 = time.time()
    o = oracle.Problem(P, xyz, q, 4.0, 80.0, KAPPA)
    t_build = time.time() - t0
    print("oracle operator built in %.0f s" % t_build, flush=True)
    t0 = time.time()
    sol, it, hist = o.gmres(o.rhs(), 1e-6)
    t_solve = time.time() - t0
    E = o.energy(sol)
    print("oracle GMRES: %d iterations in %.0f s, energy %.12f" % (it, t_solve, E), flush=True)
    out = {"atoms": ATOMS, "seed": SEED, "kappa": KAPPA, "spacing": h, "panels": len(P), "panel_sha256": digest,
           "depth": o.tree(0).depth, "iterations": int(it), "energy_fp64": E, "conversion": 332.112,
           "residual_history_last": float(hist[-1]), "oracle_build_s": t_build, "oracle_gmres_s": t_solve}
    with open(os.path.join(HERE, "globule100k_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
