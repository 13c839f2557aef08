"""Benchmark: PB solve wall-time and BEM matvecs/s on one MI355X.

Workload: synthetic globule -- 2000 atoms from a seeded self-avoiding random
walk (step 1.5 A, non-bonded pairs >= 2.4 A), radii U[1.2, 2.0] A, charges
U[-0.8, 0.8] shifted to an integer net charge, solvent-excluded surface
(1.4 A probe) triangulated to ~100k flat panels by the library's mesher;
eps_in 4, eps_out 80, kappa 0.124 1/A (Yukawa exterior), GMRES tol 1e-6.

A step is one BEM matvec: the full 2N region operator (both trees, both
layers, all regions) in mixed precision (FP32 compressed operator, FP64
vectors), on vectors already resident in HBM.  The JSON line also carries
the FP64 matvec rate, the solve wall-times (setup, GMRES, potentials) in
both modes, the solvation energies and their relative difference, the
roofline of the dominant kernel, a CPU baseline (the FP64 oracle on the
full workload, one core, run in a child process during the GPU run, and
its all-cores rate) and the batched reaction-potential solve
of 64 seeded charge vectors on the same operator (per-column mixed-vs-FP64
energy error, aggregate matvec rate).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# GMRES tolerance of the batched charge-vector solves: at 1e-6 both modes
# stop up to 5e-6 (relative energy) short of the converged answer, which
# is the order of the per-column mixed-vs-FP64 gate itself; at 1e-8 the
# FP32 operator's own error is what the gate sees
MULTI_TOL = 1e-8


def build_workload(target_panels, seed, natoms=2000):
    from fftsvd_amd import meshes

    xyz, radii, q = meshes.synthetic_globule(natoms=natoms, seed=seed)
    spacing = 1.3 * (100000.0 / target_panels) ** 0.5
    V, F, h = meshes.ses_mesh_target(xyz, radii, target_panels, probe=1.4, spacing=spacing)
    P = meshes.panels_from_mesh(V, F)
    return xyz, radii, q, P, h


def _oracle_rate(args):
    """FP64 oracle matvecs on one panel set for about `secs` seconds (single
    core); returns (matvecs per second, setup seconds)."""
    P, xyz, q, secs = args
    import oracle

    t0 = time.time()
    pr = oracle.Problem(P, xyz, q, 4.0, 80.0, 0.124)
    setup = time.time() - t0
    x = np.random.default_rng(1).uniform(-1, 1, 2 * len(P))
    n = 0
    t0 = time.time()
    while True:
        pr.matvec(x)
        n += 1
        if time.time() - t0 > secs and n >= 3:
            break
    return n / (time.time() - t0), setup


def cpu_worker(path):
    """Child process: the oracle's matvec rate, single core, on the full
    workload (panels and charges from `path`); one JSON line on stdout."""
    d = np.load(path)
    rate, setup = _oracle_rate((d["P"], d["xyz"], d["q"], 3.0))
    print(json.dumps({"panels": int(len(d["P"])), "matvecs_per_s": rate, "setup_s": setup}), flush=True)


def start_cpu_baseline(P, xyz, q):
    """Starts the single-core oracle run on the full workload in a child
    process (it shares no state with this one; it loads only the oracle),
    so it runs while the GPU is measured."""
    import subprocess
    import tempfile

    fd, path = tempfile.mkstemp(suffix=".npz")
    os.close(fd)
    np.savez(path, P=P, xyz=xyz, q=q)
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", path],
                            stdout=subprocess.PIPE, text=True)
    return proc, path


def finish_cpu_baseline(proc, path, seed, npan):
    """The oracle's single-core rate on the full workload (from the child),
    and the all-cores rate: one oracle per core (os.sched_getaffinity, at
    most 16) on a 4k-panel globule of the same recipe at once, the ratio of
    their aggregate rate to one core's rate on the same sample carried to
    the full workload."""
    import multiprocessing as mp

    from fftsvd_amd import meshes

    xyz, radii, q = meshes.synthetic_globule(natoms=120, seed=seed, confine=10.0)
    V, F, _ = meshes.ses_mesh_target(xyz, radii, 4000, probe=1.4, spacing=0.8)
    P4 = meshes.panels_from_mesh(V, F)
    cores = min(16, len(os.sched_getaffinity(0)))
    one4, _ = _oracle_rate((P4, xyz, q, 2.0))
    with mp.get_context("spawn").Pool(cores) as pool:
        res = pool.map(_oracle_rate, [(P4, xyz, q, 3.0)] * cores)
    agg4 = sum(r for r, _ in res)
    out, _ = proc.communicate(timeout=600)
    os.unlink(path)
    full = json.loads(out.strip().splitlines()[-1])
    return {"full": full, "cores": cores, "sample_panels": len(P4), "sample_one_core": one4,
            "sample_all_cores": agg4, "all_cores_matvecs_per_s": full["matvecs_per_s"] * agg4 / one4}


STAGE_KERNELS = {"leaf": "k_leaf<float>", "project_source": "k_project<float> + k_source<float>",
                 "translate": "k_translate_pipe<1>", "target": "k_target<float>"}
STAGE_LAUNCHES = {"leaf": 1, "project_source": 2, "translate": 1, "target": 1}   # per tree
# LDS of MI355X: 256 CUs x 128 bytes per clock x 2.4 GHz
LDS_PEAK_GBS = 256 * 128 * 2.4
# calibrated HBM bytes per dispatch from the committed rocprofv3 PMC summary
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "apply_kernels_globule100k_hbm_traffic.txt")


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, each
    scaled by the copy-kernel calibration) from the committed PMC summary, or
    None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("TRAFFIC ") and line[8:].rsplit(" ", 1)[0] == kernel:
                    return float(line.rsplit(" ", 1)[1])
    except OSError:
        pass
    return None


def stage_bytes(st):
    """Algorithmic HBM bytes per matvec of each forward-apply stage in mixed
    mode (4-byte operator entries, 8-byte complex spectrum bins, FP64
    vectors)."""
    n = st["points_applied"]
    return {
        # D blocks, their column indices, the U rows, x (both layers) once, y once
        "leaf": 4.0 * st["near_entries"] + 4.0 * st["near_cols"] + 4.0 * st["ut_entries"] + 24.0 * n,
        # VT and PV once, x (both layers) once, forward spectra written
        "project_source": 4.0 * st["vt_entries"] + 4.0 * st["pv_entries"] + 16.0 * n + 8.0 * st["src_bins"],
        # T spectra and source spectra read once, accumulated spectra written
        "translate": 8.0 * st["t_bins"] + 8.0 * st["src_bins"] + 8.0 * st["tgt_bins"],
        # accumulated spectra read, UTI once
        "target": 8.0 * st["tgt_bins"] + 4.0 * st["uti_entries"],
    }


def stage_rooflines(st, parts_ms):
    out = {}
    for k, b in stage_bytes(st).items():
        ms = parts_ms[k]
        ach = b / (ms * 1e-3) / 1e9
        out[k] = {"kernel": STAGE_KERNELS[k], "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": ach / HBM_PEAK_GBS, "bytes_per_matvec": b, "ms": ms,
                  "launches_per_matvec": st["trees"] * STAGE_LAUNCHES[k]}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--panels", type=int, default=100000)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--multi-rhs", type=int, default=64, help="charge vectors of the batched solve (0: skip)")
    ap.add_argument("--uneven", type=int, default=8, help="molecules of the uneven batch (0: skip)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (got --gpus %d)" % args.gpus)
        sys.exit(2)

    import fftsvd_amd as fa

    t0 = time.time()
    xyz, radii, q, P, h = build_workload(args.panels, args.seed)
    t_mesh = time.time() - t0
    print("# workload: %d atoms, %d panels (grid %.3f A), net charge %+.3f, mesh %.1fs" %
          (len(xyz), len(P), h, q.sum(), t_mesh), flush=True)
    # the CPU baseline's single-core run on the full workload starts now, in
    # a child process, and is collected after the GPU measurements
    cpu_run = None if args.no_cpu else start_cpu_baseline(P, xyz, q)

    # HIP runtime and device initialisation (once per process: context,
    # first stream, code object load) is timed apart from the operator setup
    t0 = time.time()
    dev_name, dev_cu = fa.device_info()
    fa.copy_bandwidth(1 << 20, 1)
    t_init = time.time() - t0
    t0 = time.time()
    prob = fa.Problem(P, xyz, q, 4.0, 80.0, 0.124, want_f32=True)
    t_setup = time.time() - t0
    st = prob.stats()
    print("# setup %.2fs (library %.0f ms), depth %d, near entries %.3g, far entries %.3g" %
          (t_setup, st["setup_ms"], st["depth"], st["near_entries"], st["far_entries"]), flush=True)

    solve = {}
    if not args.no_solve:
        t0 = time.time()
        rhs = prob.rhs()
        t_rhs = time.time() - t0
        sols = {}
        for name, prec in (("fp64", fa.F64), ("mixed", fa.MIXED)):
            t0 = time.time()
            sol, it, hist, info = prob.solve(rhs, 1e-6, precision=prec)
            t_gmres = time.time() - t0
            sols[name] = sol
            t0 = time.time()
            phi = prob.reaction_potentials(sol)
            t_pot = time.time() - t0
            E = 0.5 * float(np.dot(q, phi)) * 332.112 / 4.0
            solve[name] = {"gmres_s": t_gmres, "iterations": it, "refinements": int(info["refinements"]),
                           "final_residual": info["final_residual"], "potentials_s": t_pot, "energy_kcal": E,
                           "matvecs64": int(info["matvecs64"]), "matvecs32": int(info["matvecs32"])}
            print("# %s solve: %d its, %.3fs, E = %.6f kcal/mol" % (name, it, t_gmres, E), flush=True)
        E64, E32 = solve["fp64"]["energy_kcal"], solve["mixed"]["energy_kcal"]
        solve["energy_rel_err_mixed_vs_fp64"] = abs(E32 - E64) / abs(E64)
        for name in ("fp64", "mixed"):
            solve[name]["wall_s"] = t_setup + t_rhs + solve[name]["gmres_s"] + solve[name]["potentials_s"]
        solve["rhs_s"] = t_rhs
        solve["gmres_speedup_mixed"] = solve["fp64"]["gmres_s"] / solve["mixed"]["gmres_s"]

    # the column-batched matvec alone: the batched solve's columns per call,
    # device resident, HIP-event timed
    block = None
    if args.multi_rhs > 0:
        ncb = args.multi_rhs
        block = {"columns": ncb}
        for name, prec, nrep in (("mixed", fa.MIXED, 5), ("fp64", fa.F64, 3)):
            prob.bench_matvec_batch(ncb, prec, 1)
            msb, partsb = prob.bench_matvec_batch(ncb, prec, nrep)
            block[name] = {"ms_per_call": msb, "ms_per_column": msb / ncb, "matvecs_per_s": 1000.0 * ncb / msb,
                           "parts_ms_per_column": {k: v / ncb for k, v in partsb.items()}}
        print("# block matvec, %d columns per call: mixed %.4f ms per column (%.0f matvecs/s), fp64 %.4f ms; "
              "parts mixed %s" % (ncb, block["mixed"]["ms_per_column"], block["mixed"]["matvecs_per_s"],
                                  block["fp64"]["ms_per_column"],
                                  {k: round(v, 4) for k, v in block["mixed"]["parts_ms_per_column"].items()}),
              flush=True)

    # batched reaction-potential solve: seeded i.i.d. U[-0.8, 0.8] charge
    # vectors on the same atoms, one operator; per-column energy error of
    # the mixed mode against FP64 and the aggregate matvec rate of the solves
    multi = None
    if not args.no_solve and args.multi_rhs > 0:
        ncol = args.multi_rhs
        Q = np.column_stack([np.random.default_rng(args.seed + 1 + c).uniform(-0.8, 0.8, len(xyz))
                             for c in range(ncol)])
        t0 = time.time()
        R = prob.rhs_batch(Q)
        t_rhs_b = time.time() - t0
        multi = {"columns": ncol, "rhs_s": t_rhs_b, "gmres_tol": MULTI_TOL}
        E = {}
        for name, prec in (("fp64", fa.F64), ("mixed", fa.MIXED)):
            t0 = time.time()
            X, its, infos = prob.solve_batch(R, MULTI_TOL, precision=prec)
            t_solve = time.time() - t0
            t0 = time.time()
            Phi = prob.potentials_batch(X)
            t_pot = time.time() - t0
            E[name] = 0.5 * np.sum(Q * Phi, axis=0) * 332.112 / 4.0
            nmv = sum(i["matvecs64"] + i["matvecs32"] for i in infos)
            multi[name] = {"solve_s": t_solve, "potentials_s": t_pot, "iterations_mean": float(np.mean(its)),
                           "matvecs": int(nmv), "matvecs_per_s": nmv / t_solve}
            print("# batched %s: %d columns, %.2fs solve (%.0f matvecs/s aggregate), potentials %.2fs" %
                  (name, ncol, t_solve, nmv / t_solve, t_pot), flush=True)
        rel = np.abs(E["mixed"] - E["fp64"]) / np.abs(E["fp64"])
        multi["energy_rel_err_mixed_vs_fp64_max"] = float(rel.max())
        multi["energy_rel_err_mixed_vs_fp64_median"] = float(np.median(rel))
        multi["gmres_speedup_mixed"] = multi["fp64"]["solve_s"] / multi["mixed"]["solve_s"]
        # both modes stop at the MULTI_TOL residual, so part of their
        # difference is the tolerance itself: against FP64 converged to 1e-10
        # (untimed)
        Xc, _, _ = prob.solve_batch(R, 1e-10, precision=fa.F64)
        Ec = 0.5 * np.sum(Q * prob.potentials_batch(Xc), axis=0) * 332.112 / 4.0
        for name in ("fp64", "mixed"):
            multi[name]["energy_rel_err_vs_fp64_tol1e-10_max"] = float((np.abs(E[name] - Ec) / np.abs(Ec)).max())
        print("# batched energies: mixed vs fp64 max %.2e (median %.2e); vs fp64 at 1e-10: fp64 %.2e, mixed %.2e" %
              (rel.max(), np.median(rel), multi["fp64"]["energy_rel_err_vs_fp64_tol1e-10_max"],
               multi["mixed"]["energy_rel_err_vs_fp64_tol1e-10_max"]), flush=True)

    # uneven batch: seeded random-walk molecules with log-uniform atom counts
    # (the globule's recipe and grid, confinement scaled with size), each on
    # its own operator; all mixed solves one after another, then at once
    uneven = None
    if not args.no_solve and args.uneven > 0:
        from fftsvd_amd import meshes
        rng = np.random.default_rng(args.seed + 500)
        counts = np.exp(rng.uniform(np.log(300), np.log(3000), size=args.uneven)).astype(int)
        probs, rhs_u, npan_u = [], [], []
        t0 = time.time()
        for k, na in enumerate(counts):
            xu, ru, qu = meshes.synthetic_globule(natoms=int(na), seed=args.seed + 600 + k,
                                                  confine=26.0 * (na / 2000.0) ** (1 / 3))
            Vu, Fu = meshes.ses_mesh(xu, ru, probe=1.4, spacing=h)
            pk = fa.Problem(meshes.panels_from_mesh(Vu, Fu), xu, qu, 4.0, 80.0, 0.124, want_f32=True)
            probs.append(pk)
            rhs_u.append(pk.rhs())
            npan_u.append(pk.nvec // 2)
        t_build = time.time() - t0
        t0 = time.time()
        seq = [pk.solve(r, 1e-6, precision=fa.MIXED) for pk, r in zip(probs, rhs_u)]
        t_seq = time.time() - t0
        t0 = time.time()
        Xu, itu, infu = fa.solve_many(probs, rhs_u, 1e-6, precision=fa.MIXED)
        t_conc = time.time() - t0
        same = all(np.array_equal(Xu[k], seq[k][0]) for k in range(len(probs)))
        nmv = sum(d["matvecs64"] + d["matvecs32"] for d in infu)
        uneven = {"molecules": len(probs), "atoms": [int(c) for c in counts], "panels": npan_u,
                  "build_s": t_build, "solve_sequential_s": t_seq, "solve_concurrent_s": t_conc,
                  "speedup": t_seq / t_conc, "matvecs": int(nmv), "matvecs_per_s_concurrent": nmv / t_conc,
                  "concurrent_equals_sequential": bool(same)}
        print("# uneven batch: %d molecules (%d..%d panels), mixed solves %.3fs one after another, %.3fs at once "
              "(%.2fx, %.0f matvecs/s), identical %s; meshing + setup %.1fs" %
              (len(probs), min(npan_u), max(npan_u), t_seq, t_conc, t_seq / t_conc, nmv / t_conc, same, t_build),
              flush=True)
        for pk in probs:
            pk.close()

    # gradients of the energy (the adjoint): d/d(charges) and d/d(kappa) at
    # the mixed solution, timed; against central differences (charges: the
    # energy is quadratic, 8 seeded charges, batched FP64 solves at 1e-10;
    # kappa: operators rebuilt at kappa +- h, Richardson over h = 2e-3, 1e-3)
    grads = None
    if not args.no_solve:
        conv = 332.112 / 4.0
        t0 = time.time()
        gq = prob.grad_charges(sols["mixed"], 1e-6, precision=fa.MIXED)
        t_gq = time.time() - t0
        t0 = time.time()
        gk = prob.grad_kappa(sols["mixed"], 1e-6, precision=fa.MIXED)
        t_gk_first = time.time() - t0
        t0 = time.time()
        gk = prob.grad_kappa(sols["mixed"], 1e-6, precision=fa.MIXED)
        t_gk = time.time() - t0
        x64, _, _, _ = prob.solve(prob.rhs(), 1e-10, precision=fa.F64)
        gq64 = prob.grad_charges(x64, 1e-10, precision=fa.F64)
        gk64 = prob.grad_kappa(x64, 1e-10, precision=fa.F64)
        js = np.random.default_rng(args.seed + 100).choice(len(q), 8, replace=False)
        hq = 1e-3
        Qf = np.repeat(q[:, None], 2 * len(js), axis=1)
        for c, j in enumerate(js):
            Qf[j, 2 * c] += hq
            Qf[j, 2 * c + 1] -= hq
        Xf, _, _ = prob.solve_batch(prob.rhs_batch(Qf), 1e-10, precision=fa.F64)
        Ef = 0.5 * np.sum(Qf * prob.potentials_batch(Xf), axis=0) * conv
        fdq = (Ef[0::2] - Ef[1::2]) / (2 * hq)
        dk = {}
        for hk in (2e-3, 1e-3):
            e = []
            for kk in (0.124 + hk, 0.124 - hk):
                pk = fa.Problem(P, xyz, q, 4.0, 80.0, kk, want_f32=False)
                xk, _, _, _ = pk.solve(pk.rhs(), 1e-10, precision=fa.F64)
                e.append(pk.energy(xk))
                pk.close()
                del pk
            dk[hk] = (e[0] - e[1]) / (2 * hk)
        fdk = (4 * dk[1e-3] - dk[2e-3]) / 3
        grads = {"charges_s": t_gq, "kappa_s": t_gk, "kappa_first_call_s": t_gk_first,
                 "dE_dkappa_kcal_per_A": gk, "charge_fd_max_rel": float(np.max(np.abs(gq64[js] - fdq) / np.abs(fdq))),
                 "charge_mixed_vs_fp64_rel": float(np.linalg.norm(gq - gq64) / np.linalg.norm(gq64)),
                 "kappa_fd_rel": abs(gk64 - fdk) / abs(fdk), "kappa_mixed_vs_fp64_rel": abs(gk - gk64) / abs(gk64),
                 "note": "central differences of the energy with the operator rebuilt at kappa +- h at the "
                         "default SVD tolerance 1e-5: re-sampled bases move it by a few 1e-5; at 1e-8 the "
                         "tests find 1e-9"}
        print("# gradients: charges %.3fs, kappa %.3fs (first call %.3fs); dE/dkappa %.6f, vs central "
              "differences %.2e (charges %.2e)" % (t_gq, t_gk, t_gk_first, gk, grads["kappa_fd_rel"],
                                                  grads["charge_fd_max_rel"]), flush=True)

    # matvec throughput, device-resident vectors, HIP-event timed on the library stream
    prob.bench_matvec(fa.MIXED, max(1, args.warmup))
    t0 = time.time()
    ms32, parts32 = prob.bench_matvec(fa.MIXED, args.steps)
    wall32 = time.time() - t0
    ms64, parts64 = prob.bench_matvec(fa.F64, max(10, args.steps // 4))
    print("# matvec mixed %.3f ms, fp64 %.3f ms; parts mixed %s" % (ms32, ms64, parts32), flush=True)

    # rooflines: algorithmic HBM bytes of each apply stage (FP32 storage,
    # every stored entry read once, spectra written and read once, vectors
    # read/written once) over its HIP-event time on the library stream
    rf = stage_rooflines(st, parts32)
    dom = max(rf, key=lambda k: rf[k]["ms"])
    roofline = dict(rf[dom])
    roofline.update({"stage": dom, "traffic": None})
    if STAGE_LAUNCHES[dom] == 1:
        tr = measured_traffic(roofline["kernel"])
        if tr is not None:
            roofline.update({"traffic": tr, "traffic_unit": "bytes per launch",
                             "traffic_source": "profiles/" + os.path.basename(TRAFFIC_FILE),
                             "algorithmic_bytes_per_launch": roofline["bytes_per_matvec"] / st["trees"]})
    # the translation is bound by LDS, not HBM: every complex multiply-add
    # reads one T bin and one source bin (2 x 8 bytes) from LDS
    # reads one T bin and one source bin (2 x 8 bytes) from LDS.  The peak is
    # the nominal 128 B/clk/CU at 2.4 GHz; the measured LDS read rate of this
    # GPU with the kernel's 8-byte reads (fftsvd_lds_bandwidth) is beside it
    lds_bytes = 16.0 * st["translate_macs"]
    lds_rate = lds_bytes / (parts32["translate"] * 1e-3) / 1e9
    lds_meas8 = max(fa.lds_bandwidth(8, 5) for _ in range(2))
    lds_meas16 = max(fa.lds_bandwidth(16, 5) for _ in range(2))
    roofline_lds = {"kernel": STAGE_KERNELS["translate"], "bound": "lds", "achieved": lds_rate,
                    "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": lds_rate / LDS_PEAK_GBS,
                    "measured_peak_8B_reads": lds_meas8, "frac_of_measured": lds_rate / lds_meas8,
                    "measured_peak_16B_reads": lds_meas16,
                    "lds_bytes_per_matvec": lds_bytes, "complex_macs_per_matvec": st["translate_macs"]}
    print("# LDS read rate measured: %.0f GB/s with 8-byte reads, %.0f with 16-byte reads; translation at %.0f "
          "(%.2f of the measured 8-byte rate)" % (lds_meas8, lds_meas16, lds_rate, lds_rate / lds_meas8), flush=True)

    # measured copy bandwidth of this GPU (1 GiB copied 20 times)
    copy_gbs = fa.copy_bandwidth(1 << 30, 20)

    cpu = None
    if cpu_run is not None:
        try:
            cb = finish_cpu_baseline(cpu_run[0], cpu_run[1], args.seed, len(P))
            cpu = {"value": cb["all_cores_matvecs_per_s"], "unit": "matvecs/s", "cores": cb["cores"],
                   "kind": "port",
                   "sample": "FP64-storage oracle (the reference stores its operator in FP32) on the full %d-panel "
                             "workload: %.3f matvecs/s on one core (operator built in %.0f s, in a child process "
                             "during the GPU run); all cores: %d oracles at once on a %d-panel globule of the same "
                             "recipe ran %.2fx one core's rate there, applied to the full-size rate" %
                             (cb["full"]["panels"], cb["full"]["matvecs_per_s"], cb["full"]["setup_s"], cb["cores"],
                              cb["sample_panels"], cb["sample_all_cores"] / cb["sample_one_core"]),
                   "single_core_value": cb["full"]["matvecs_per_s"], "single_core_setup_s": cb["full"]["setup_s"]}
        except Exception as e:  # the baseline never blocks the GPU line
            cpu = {"value": None, "unit": "matvecs/s", "cores": 1, "kind": "port", "sample": "failed: %s" % e}

    line = {
        "metric": "PB solve wall-time & BEM matvecs/s at 1e-6 rel energy err; mixed vs FP64 GPU",
        "value": 1000.0 / ms32,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "synthetic globule, 2000-atom seeded self-avoiding walk, SES 1.4 A probe, "
                               "~100k flat panels, eps 4/80, kappa 0.124, GMRES tol 1e-6",
                   "panels": len(P), "atoms": len(xyz), "seed": args.seed, "mesh_spacing_A": h,
                   "precision": "FP32 compressed operator and Krylov basis; FP64 work vectors, Hessenberg "
                                "and residual refinement"},
        "fp64_matvecs_per_s": 1000.0 / ms64,
        "wall_step_check_ms": wall32 * 1e3 / args.steps,
        "matvec_parts_ms": {"mixed": parts32, "fp64": parts64},
        "setup_s": t_setup,
        "device_init_s": t_init,
        "device": {"name": dev_name, "cus": dev_cu},
        "solve": solve,
        "multi_rhs": multi,
        "block_matvec": block,
        "uneven_batch": uneven,
        "gradients": grads,
        "operator": st,
        "roofline": roofline,
        "roofline_stages": rf,
        "roofline_translate_lds": roofline_lds,
        "hbm_copy_GBs": copy_gbs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
