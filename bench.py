"""Benchmark: harmonic visco-elastic BEM frequency sweep on one MI355X.

Workload (default, soil surface): the BEM part of the soil-pile system -- a truncated half-space
free surface 40 x 40 m at z = 0, 50 x 50 quad9 graded from 0.37 m at the centre to 1.47 m at the
edge (edge/centre size 4; 10,201 nodes, 30,603 complex DOF; the 200 perimeter nodes collocated by
SBIE-MCA, 400 points).  Soil E = 5e7 Pa, nu = 0.4, rho = 1800 kg/m3, 5% hysteretic damping.
Tractions prescribed everywhere: free surface, with a uniform vertical traction of 1 Pa on the
6 x 6 m footprint of the pile group at the centre.  32 frequencies, 0.5 .. 20 Hz; a run of K steps
takes K of them evenly spread over the band (the default, 32 steps, is the whole sweep).
(--workload sphere: the harmonic sphere, quad9 projected cube n = 24, 10,374 complex DOF,
8 frequencies a0 in [0.1, 4].)  Integration tolerance qsi_relative_error = 1e-10 on both paths
(a tolerance of 1e-6 would separate two correct rule choices by ~1e-6 in the solution, far above
the 1e-8 bar).

Two pipelines, each timed over the same frequencies (one step = one frequency, inputs in HBM):
  * learned + mixed (the measured `value`): rules chosen by the quadrature network (with the
    judge: a sample of the network's pairs re-integrated by the adaptive rule every step), complex
    FP64 assembly, complex FP32 LU whose trailing updates run as one FP16 GEMM of hi/lo products
    with FP32 accumulation (--solve fp32: plain complex FP32 updates) + FP64 iterative refinement
    (FP64 LU fallback);
  * plain FP64: the reference's rule estimator, complex FP64 assembly, FP64 LU (zgesv).
The assembly of frequency k+1 runs on a side stream while frequency k is solved.  The relative
difference of the two solutions is reported for every frequency (bar: 1e-8), and rows of the GPU
matrix are compared with the oracle's (C restatement of the reference) outside the timed region.

Prints one JSON line.  value = BEM matrix entries per second through the learned + mixed pipeline
(n_dof^2 x steps / wall time of assembly + solve).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12          # MI355X FP64 vector (AMD public spec)
FLOPS_PER_POINT_HARM = 1500  # harmonic kernel with precalculated data: flops per quadrature point
QNET_DEFAULT = os.path.join(ROOT, "multifebe_amd", "data", "qnet_tau1e-10.npz")
# HBM bytes per launch of k_regular_tile<harmonic> (mean over one assembly's launches), from
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy; None where not measured
TRAFFIC_PER_LAUNCH = {"sphere": None, "soil": 25.78e9}


def workload(name: str):
    from multifebe_amd import mesh as M
    if name == "soil":
        m = M.surface_patch_load_bcs(M.free_surface(40.0, 50, 4.0), half_width=3.0, traction=-1.0)
        E, nu, xi, rho = 5e7, 0.4, 0.05, 1800.0
        mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
        omegas = 2.0 * np.pi * np.linspace(0.5, 20.0, 32)
        desc = (f"soil-pile BEM part: graded quad9 free surface 40 x 40 m, 50 x 50 elements "
                f"0.37 -> 1.47 m ({m.nn} nodes, {m.n_dof} complex DOF, SBIE-MCA on the open edges), "
                f"E 5e7 Pa, nu 0.4, rho 1800, 5% damping, 32 frequencies 0.5-20 Hz")
    else:
        m = M.pressure_hemisphere_bcs(M.sphere(24, 1.0), pressure=1.0)
        E, nu, xi, rho = 1e9, 0.3, 0.05, 2000.0
        mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
        omegas = np.linspace(0.1, 4.0, 8) * np.sqrt(mu_r / rho)
        desc = (f"harmonic sphere, quad9 projected cube n=24 ({m.nn} nodes, {m.n_dof} complex "
                f"DOF), 8 frequencies a0 in [0.1, 4]")
    return m, dict(E=E, nu=nu, xi=xi, rho=rho, mu_r=mu_r, mu_c=mu_c, lam_c=lam_c), omegas, desc


def pick(omegas, steps):
    """`steps` frequencies evenly spread over the sweep (cycled when steps > len(omegas))"""
    n = len(omegas)
    if steps >= n:
        return [float(omegas[k % n]) for k in range(steps)]
    idx = np.round(np.linspace(0, n - 1, steps)).astype(int)
    return [float(omegas[k]) for k in idx]


def oracle_rows(m, p, omega, tau, node_ranges):
    """oracle (C restatement of the reference) rows of A, b for the collocation nodes of
    node_ranges, with the time it took"""
    from oracle import oracle as O
    rows, As, bs = [], [], []
    t0 = time.perf_counter()
    for r in node_ranges:
        A, b = O.assemble(m, lam_c=p["lam_c"], mu_c=p["mu_c"], rho=p["rho"], omega=omega,
                          nu=p["nu"], node_range=r, qsi_re=tau)
        rows.append(np.arange(3 * r[0], 3 * r[1]))
        As.append(np.asarray(A))
        bs.append(b)
    return np.concatenate(rows), np.concatenate(As), np.concatenate(bs), time.perf_counter() - t0


def cpu_baseline(m, p, omega, tau, seconds=20.0):
    """Oracle (C restatement of the reference, OpenMP over elements) + host LAPACK on this host's
    cores: assembly of a bounded block of collocation nodes and an LU of a smaller matrix, both
    scaled to the full problem.  Returns the record and the oracle rows (for the parity check)."""
    import scipy.linalg as sla
    cores = len(os.sched_getaffinity(0))
    n = m.n_dof
    nb = 16
    while True:
        ranges = [(0, nb), (m.nn // 2, m.nn // 2 + nb)]
        rows, Ao, bo, ta = oracle_rows(m, p, omega, tau, ranges)
        if ta > seconds / 3 or 2 * nb >= m.nn:
            break
        nb = min(m.nn // 2, nb * 4)
    t_asm_full = ta * m.nn / (2 * nb)
    ns = 4000
    rng = np.random.default_rng(0)
    As = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns)) + ns * np.eye(ns)
    bs = rng.standard_normal(ns) + 0j
    t0 = time.perf_counter()
    sla.solve(As, bs, check_finite=False)
    tl = time.perf_counter() - t0
    t_lu_full = tl * (n / ns) ** 3
    per_freq = t_asm_full + t_lu_full
    rec = {"value": n * n / per_freq, "unit": "BEM matrix entries/s", "cores": cores,
           "kind": "port",
           "sample": (f"oracle assembly (qsi_relative_error {tau:g}, OpenMP {cores} threads) of "
                      f"{2 * nb}/{m.nn} collocation nodes at one frequency ({ta:.1f} s) scaled to "
                      f"all nodes, plus scipy zgesv on n={ns} ({tl:.2f} s) scaled by (n/{ns})^3: "
                      f"EXTRAPOLATED per-frequency time {per_freq:.1f} s"),
           "assembly_s_per_freq": t_asm_full, "lu_s_per_freq": t_lu_full}
    return rec, (rows, Ao, bo)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32,
                    help="frequencies timed (default: the whole 32-frequency sweep)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["soil", "sphere"], default="soil")
    ap.add_argument("--tau", type=float, default=1e-10, help="qsi_relative_error of both paths")
    ap.add_argument("--qnet", default=QNET_DEFAULT, help="trained network weights (.npz)")
    ap.add_argument("--confidence", type=float, default=0.99)
    ap.add_argument("--judge", type=int, default=256, help="judge sample per step")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 pipeline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-alone", action="store_true",
                    help="skip the extra assembly that times the dominant kernel alone")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--serial", action="store_true",
                    help="assemble and solve one after the other (no side-stream overlap)")
    ap.add_argument("--no-converged", action="store_true",
                    help="skip the accuracy check against a converged quadrature")
    ap.add_argument("--solve", choices=["split16", "fp32"], default="split16",
                    help="factorisation of the learned pipeline (both refined in FP64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from multifebe_amd import bem
    from multifebe_amd.qnet import QNet

    m, p, omegas, desc = workload(args.workload)
    n = m.n_dof
    wdat = np.load(args.qnet)
    net = QNet()
    net.params = wdat["params"]
    qrep = json.loads(str(wdat["report"]))

    t_setup0 = time.perf_counter()
    learned = bem.BemSystem(m, bem.Settings(qsi_relative_error=args.tau))
    learned.set_qnet(net, confidence=args.confidence, judge_qsi_relative_error=1e-12,
                     judge_samples=args.judge)
    plain = bem.BemSystem(m, bem.Settings(qsi_relative_error=args.tau))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup0
    ncp = learned.collocation()[0].shape[0]
    omegas_run = pick(omegas, args.steps)
    lam, mu_c, rho, nu = p["lam_c"], p["mu_c"], p["rho"], p["nu"]
    mixed = bem.PREC_SPLIT16_REFINE if args.solve == "split16" else bem.PREC_FP32_REFINE
    print(f"# {desc}; steps at f = {[round(w / 2 / np.pi, 3) for w in omegas_run]} Hz",
          file=sys.stderr, flush=True)
    # warm-up: workspaces, kernels, the network and both solvers once per pipeline
    for w in range(args.warmup):
        learned.sweep_harela(omegas_run[:2], lam, mu_c, rho, nu, mixed, overlap=not args.serial)
        if not args.no_plain:
            plain.sweep_harela(omegas_run[:2], lam, mu_c, rho, nu, bem.PREC_FP64,
                               overlap=not args.serial)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = learned.sweep_harela(omegas_run, lam, mu_c, rho, nu, mixed, overlap=not args.serial)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"# learned + mixed sweep {wall:.2f} s", file=sys.stderr, flush=True)
    st = learned.stats()
    tm = learned.timings()
    recs = [(x, a, t, rep) for x, rep, a, t, _ in res]
    # the dominant kernel over every assembly of the timed sweep
    reg_ms = sum(r[4]["timings"]["regular_ms"] for r in res)
    launches = sum(r[4]["timings"]["regular_launches"] for r in res)
    reg_points = sum(r[4]["stats"]["regular_points"] for r in res)
    judges = [(r[4]["stats"]["judge_max_err"], r[4]["stats"]["judge_ref_max_err"]) for r in res]
    # the same kernel with nothing beside it: one more learned assembly after the timed sweep
    # (not part of the step time), for the kernel's own throughput next to the in-sweep figure
    alone = None
    if not args.no_alone:
        At1, b1 = learned.assemble_harela(omegas_run[-1], lam, mu_c, rho, nu)
        torch.cuda.synchronize()
        tm1, st1 = learned.timings(), learned.stats()
        del At1, b1
        alone_ach = st1["regular_points"] * FLOPS_PER_POINT_HARM / (tm1["regular_ms"] * 1e-3)
        alone = {"launches": tm1["regular_launches"],
                 "avg_launch_ms": tm1["regular_ms"] / max(tm1["regular_launches"], 1),
                 "achieved": alone_ach / 1e12, "frac": alone_ach / PEAK_FP64,
                 "note": "one learned assembly alone after the timed sweep (nothing else on the "
                         "GPU)"}

    plain_out = None
    errs = []
    if not args.no_plain:
        t0 = time.perf_counter()
        pres = plain.sweep_harela(omegas_run, lam, mu_c, rho, nu, bem.PREC_FP64,
                                  overlap=not args.serial)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        print(f"# plain FP64 sweep {pwall:.2f} s", file=sys.stderr, flush=True)
        for r, q in zip(res, pres):
            errs.append((torch.linalg.vector_norm(r[0] - q[0]) / torch.linalg.vector_norm(q[0])).item())
        lu_flops = 8.0 / 3.0 * n ** 3
        pst = plain.stats()
        plain_out = {
            "ms_per_step": 1e3 * pwall / args.steps,
            "assembly_ms_per_step": 1e3 * sum(q[2] for q in pres) / args.steps,
            "solve_ms_per_step": 1e3 * sum(q[3] for q in pres) / args.steps,
            "lu_tflops": lu_flops / (sum(q[3] for q in pres) / args.steps) / 1e12,
            "quad_points_per_step": pst["points"], "pairs_adaptive": pst["adaptive"],
            "speedup_of_learned_mixed": pwall / wall,
            "rel_err_learned_mixed_vs_plain_fp64": errs,
            "max_rel_err": max(errs),
        }

    t_asm = sum(r[1] for r in recs)
    t_sol = sum(r[2] for r in recs)
    # dominant assembly kernel: precalculated-rule pairs, FP64-VALU bound; algorithmic work =
    # quadrature points of those pairs x flops per point of the harmonic kernel.  Its duration is
    # the HIP-event time of its launches (recorded on the assembly's stream), summed over every
    # assembly of the timed sweep.
    achieved = reg_points * FLOPS_PER_POINT_HARM / (reg_ms * 1e-3)
    roofline = {"kernel": "k_regular_tile<true> (harmonic, precalculated-rule pairs, scatter into A)",
                "bound": "mfma",
                "compute_pipe": "FP64 VALU (MI355X FP64 vector peak = FP64 matrix peak = 78.6 TFLOP/s)",
                "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": TRAFFIC_PER_LAUNCH[args.workload],
                "traffic_unit": "bytes per launch (mean)",
                "launches": launches, "avg_launch_ms": reg_ms / max(launches, 1),
                "regular_points": reg_points, "flops_per_point": FLOPS_PER_POINT_HARM,
                "timed": "every assembly of the timed sweep, each beside the LU of the previous "
                         "frequency",
                "alone": alone}
    # accuracy against a converged quadrature at the highest frequency of the run: every pair
    # that the reference would integrate with a precalculated rule takes the 12 x 12 rule (the
    # reference's rule choice does not look at the wavenumber), near pairs the adaptive rule
    converged = None
    if not args.no_converged and not args.no_plain:
        k_hi = int(np.argmax(omegas_run))
        w_hi = omegas_run[k_hi]
        conv = bem.BemSystem(m, bem.Settings(qsi_relative_error=args.tau, precalsets=(12,)))
        Atc, bc = conv.assemble_harela(w_hi, lam, mu_c, rho, nu)
        xc, _ = bem.gesv_device(Atc, bc, bem.PREC_FP64)
        del Atc
        nx = torch.linalg.vector_norm(xc).item()
        converged = {"frequency_hz": w_hi / 2 / np.pi,
                     "rel_err_learned_mixed": (torch.linalg.vector_norm(res[k_hi][0] - xc) / nx).item(),
                     "rel_err_plain_fp64": (torch.linalg.vector_norm(pres[k_hi][0] - xc) / nx).item(),
                     "reference": "FP64 LU of the system with 12 x 12 Gauss-Legendre on every "
                                  "precalculated-rule pair, adaptive rule elsewhere"}
        conv.close()
    cpu = None
    parity = {"rel_err_learned_mixed_vs_plain_fp64_max": max(errs) if errs else None,
              "backward_error_max": max(r[3]["backward_error"] for r in recs)}
    if not args.no_cpu:
        cpu, (rows, Ao, bo) = cpu_baseline(m, p, omegas_run[0], args.tau, args.cpu_seconds)
        # the same rows of the GPU matrix (plain rules, the oracle's decisions) at that frequency
        At, b = plain.assemble_harela(omegas_run[0], lam, mu_c, rho, nu)
        A_rows = At.T[torch.as_tensor(rows, device="cuda")].cpu().numpy()
        b_rows = b[torch.as_tensor(rows, device="cuda")].cpu().numpy()
        del At, b
        parity["A_rows_checked"] = int(len(rows))
        parity["A_err_vs_oracle"] = float(np.max(np.abs(A_rows - Ao) / np.abs(Ao).max(axis=1, keepdims=True)))
        parity["b_err_vs_oracle"] = float(np.max(np.abs(b_rows - bo)) / max(np.max(np.abs(bo)), 1e-300))
    out = {
        "metric": "BEM entries/s + solve time at 30k DOF; learned+mixed vs plain FP64, 1e-8 err",
        "value": n * n * args.steps / wall,
        "unit": "BEM matrix entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c128 assembly, c64 LU (trailing updates: f16 x3 products, f32 accumulation) + "
                  "c128 refinement" if args.solve == "split16" else
                  "c128 assembly, c64 LU + c128 refinement"),
        "data": "synthetic",
        "config": {"workload": desc + f"; qsi_relative_error {args.tau:g}",
                   "n_dof": n, "collocation_points": ncp, "rules": "learned",
                   "solve": ("FP32 LU with split FP16 trailing updates + FP64 refinement"
                             if args.solve == "split16" else "FP32 LU + FP64 refinement"),
                   "frequencies_hz": [w / 2 / np.pi for w in omegas_run]},
        "time_to_solution_s_per_step": wall / args.steps,
        "assembly_ms_per_step": 1e3 * t_asm / args.steps,
        "solve_ms_per_step": 1e3 * t_sol / args.steps,
        "pipeline": ("assembly and solve one after the other" if args.serial else
                     "assembly of frequency k+1 on a side stream overlaps the solve of k"),
        "assembly_entries_per_s": n * n * args.steps / t_asm,
        "lu_tflops_per_step": 8.0 / 3.0 * n ** 3 / (t_sol / args.steps) / 1e12,
        "setup_s": t_setup,
        "pairs": st["pairs"], "pairs_singular": st["singular"], "pairs_adaptive": st["adaptive"],
        "quad_points_per_step": st["points"],
        "judge": {"pairs": st["judge_pairs"], "max_err_per_step": [j[0] for j in judges],
                  "reference_estimator_max_err_same_pairs": [j[1] for j in judges],
                  "tolerance": args.tau},
        "qnet": {"file": os.path.relpath(args.qnet, ROOT), "confidence": args.confidence,
                 "holdout_accuracy": qrep.get("holdout_accuracy"),
                 "training_rows": qrep.get("rows")},
        "kernel_ms_last_step": tm,
        "refinement_steps": [r[3]["refinement_steps"] for r in recs],
        "fell_back": [r[3]["fell_back"] for r in recs],
        "parity": parity,
        "plain_fp64": plain_out,
        "accuracy_vs_converged_quadrature": converged,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    learned.close()
    plain.close()
    net.close()


if __name__ == "__main__":
    main()
