"""Benchmark: harmonic visco-elastic BEM frequency sweep on one MI355X.

Workload (time-harmonic sphere): cube with 24 intervals per edge projected onto a sphere of radius
1 m, quad9 (3458 nodes, 10374 complex DOF); E = 1e9 (1 + 0.1 i) Pa, nu = 0.3, rho = 2000 kg/m3;
unit normal pressure on z > 0, z <= 0 fixed; frequencies a0 = omega R / c_s evenly spaced in
[0.1, 4].  One step = one frequency: complex FP64 assembly of A, b on the GPU followed by the
dense solve (default: FP32 LU on MFMA + FP64 iterative refinement, FP64 fallback).

Prints one JSON line.  value = BEM matrix entries per second through the whole pipeline
(n_dof^2 x steps / wall time of assembly + solve, inputs resident on the GPU).
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
PEAK_FP32_MFMA = 157.3e12    # MI355X FP32 matrix
FLOPS_PER_POINT_HARM = 1500  # harmonic kernel with precalculated data: flops per quadrature point


def workload(n_int=24):
    from multifebe_amd import mesh as M
    m = M.pressure_hemisphere_bcs(M.sphere(n_int, 1.0), pressure=1.0)
    E, nu, xi, rho = 1e9, 0.3, 0.05, 2000.0   # E(1 + 0.1i) = mu (1 + 2 i xi) with xi = 0.05
    mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
    cs = np.sqrt(mu_r / rho)
    return m, dict(E=E, nu=nu, xi=xi, rho=rho, mu_r=mu_r, mu_c=mu_c, lam_c=lam_c, cs=cs)


def cpu_baseline(m, p, omega, seconds=20.0):
    """Oracle (C restatement of the reference, OpenMP over elements) + host LAPACK on this host's
    cores: assembly of a bounded block of collocation nodes and an LU of a smaller matrix, both
    scaled to the full problem."""
    from oracle import oracle as O
    import scipy.linalg as sla
    cores = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    n = m.n_dof
    # assembly: grow the node block until it takes a few seconds
    nb = 16
    while True:
        t0 = time.perf_counter()
        O.assemble(m, lam_c=p["lam_c"], mu_c=p["mu_c"], rho=p["rho"], omega=omega, nu=p["nu"],
                   node_range=(0, nb))
        ta = time.perf_counter() - t0
        if ta > seconds / 3 or nb >= m.nn:
            break
        nb = min(m.nn, nb * 4)
    t_asm_full = ta * m.nn / nb
    # LU: complex FP64 solve of a smaller dense matrix, scaled by n^3
    ns = 3000
    rng = np.random.default_rng(0)
    As = rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns)) + ns * np.eye(ns)
    bs = rng.standard_normal(ns) + 0j
    t0 = time.perf_counter()
    sla.solve(As, bs, check_finite=False)
    tl = time.perf_counter() - t0
    t_lu_full = tl * (n / ns) ** 3
    per_freq = t_asm_full + t_lu_full
    return {"value": n * n / per_freq, "unit": "BEM matrix entries/s", "cores": cores,
            "kind": "port",
            "sample": (f"oracle assembly of {nb}/{m.nn} collocation nodes at one frequency "
                       f"({ta:.1f} s) scaled to all nodes, plus scipy zgesv on n={ns} ({tl:.1f} s) "
                       f"scaled by (n/{ns})^3; extrapolated per-frequency time {per_freq:.1f} s"),
            "assembly_s_per_freq": t_asm_full, "lu_s_per_freq": t_lu_full}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--n-int", type=int, default=24, help="intervals per cube edge (24: 10374 DOF)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-fp64-compare", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from multifebe_amd import bem

    m, p = workload(args.n_int)
    n = m.n_dof
    a0s = np.linspace(0.1, 4.0, 8)
    omegas = a0s * p["cs"] / 1.0
    mode = bem.PREC_FP32_REFINE if args.mode == "mixed" else bem.PREC_FP64

    t_setup0 = time.perf_counter()
    sysm = bem.BemSystem(m)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup0
    At, b = sysm.alloc(True)
    Aw = torch.empty_like(At)
    bw = torch.empty_like(b)

    def step(k, keep=False):
        om = float(omegas[k % len(omegas)])
        ta = time.perf_counter()
        sysm.assemble_harela(om, p["lam_c"], p["mu_c"], p["rho"], p["nu"], At, b)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        bw.copy_(b)
        if mode == bem.PREC_FP64:
            Aw.copy_(At)
            x, rep = bem.gesv_device(Aw, bw, mode)
        else:
            x, rep = bem.gesv_device(At, bw, mode)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        return tb - ta, tc - tb, rep, sysm.timings()

    for w in range(args.warmup):
        step(w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = []
    for k in range(args.steps):
        recs.append(step(k))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    t_asm = sum(r[0] for r in recs)
    t_sol = sum(r[1] for r in recs)
    st = sysm.stats()
    tm = recs[-1][3]
    # dominant assembly kernel: precalculated-rule pairs, FP64-VALU bound; algorithmic work =
    # quadrature points x flops per point of the harmonic kernel with precalculated data
    reg_ms = float(np.mean([r[3]["regular_ms"] for r in recs]))
    launches = recs[-1][3]["regular_launches"]
    reg_flops = st["regular_points"] * FLOPS_PER_POINT_HARM
    achieved = reg_flops / (reg_ms * 1e-3)
    roofline = {"kernel": "k_regular<harmonic> (precalculated-rule pairs)", "bound": "fp64-valu",
                "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64, "traffic": None,
                "launches_per_step": launches, "avg_launch_ms": reg_ms / max(launches, 1),
                "flops_per_point": FLOPS_PER_POINT_HARM}
    # plain FP64 path on the last frequency: time and solution error of the mixed path
    fp64 = None
    if mode != bem.PREC_FP64 and not args.no_fp64_compare:
        k = (args.steps - 1) % len(omegas)
        sysm.assemble_harela(float(omegas[k]), p["lam_c"], p["mu_c"], p["rho"], p["nu"], At, b)
        bw.copy_(b)
        xm, _ = bem.gesv_device(At, bw, mode)
        xm = xm.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sysm.assemble_harela(float(omegas[k]), p["lam_c"], p["mu_c"], p["rho"], p["nu"], At, b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Aw.copy_(At)
        bw.copy_(b)
        x64, rep64 = bem.gesv_device(Aw, bw, bem.PREC_FP64)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        err = (torch.linalg.vector_norm(xm - x64) / torch.linalg.vector_norm(x64)).item()
        lu_flops = 8.0 / 3.0 * n ** 3
        fp64 = {"ms_per_step": 1e3 * (t2 - t0), "solve_ms": 1e3 * (t2 - t1),
                "lu_tflops": lu_flops / ((t2 - t1) * 1e12),
                "speedup_of_mixed": (t2 - t0) / (wall / args.steps),
                "rel_err_mixed_vs_fp64": err}
    mixed_lu_tflops = 8.0 / 3.0 * n ** 3 / (t_sol / args.steps) / 1e12
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(m, p, float(omegas[0]), args.cpu_seconds)
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
        "dtype": "c128 assembly, c64 LU + c128 refinement" if mode else "c128",
        "data": "synthetic",
        "config": {"workload": f"harmonic sphere, quad9 projected cube n={args.n_int} "
                               f"({m.nn} nodes, {n} complex DOF), 8 frequencies a0 in [0.1, 4]",
                   "n_dof": n, "solve": args.mode},
        "assembly_ms_per_step": 1e3 * t_asm / args.steps,
        "solve_ms_per_step": 1e3 * t_sol / args.steps,
        "assembly_entries_per_s": n * n * args.steps / t_asm,
        "setup_s": t_setup,
        "pairs": st["pairs"], "pairs_singular": st["singular"], "pairs_adaptive": st["adaptive"],
        "quad_points_per_step": st["points"],
        "kernel_ms_last_step": tm,
        "refinement_steps": [r[2]["refinement_steps"] for r in recs],
        "fell_back": [r[2]["fell_back"] for r in recs],
        "backward_error": max(r[2]["backward_error"] for r in recs),
        "lu_tflops_per_step": mixed_lu_tflops,
        "plain_fp64": fp64,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
