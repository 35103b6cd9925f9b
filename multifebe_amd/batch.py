"""Batches of independent uneven problems: many closed quad9 boundaries (cubes, projected-cube
spheres, capped cylinders) at several materials and frequencies each, assembled by one replay of a
captured HIP graph (every (problem, frequency) job into its own matrix) and then solved.

    python -m multifebe_amd.batch [--count 48] [--seed 20261203] [--mode split16|fp32|fp64] [--tune]

The seeded generator follows the batched configuration: n uniform over the even values in
[4, 16] intervals per edge (about 300 to 4,600 DOF), E log-uniform in [1e8, 1e10] Pa, nu uniform
in [0.2, 0.35], hysteretic damping uniform in [0, 0.1], rho = 2000 kg/m3, 1 to 16 frequencies with
a0 = omega R / c_s uniform in [0.1, 5] (R = 1 m: the radius of spheres and cylinders, the side of
cubes).  Loads: cubes bottom fixed and top compressed; spheres and cylinders fixed below z = 0
and under unit pressure above.  With --tune the solves use, per size class, the precision mode
and schedule that the autotuner (autotune.py) found fastest on the class's largest system.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import autotune, bem
from . import mesh as M


@dataclass
class Problem:
    shape: str
    n: int
    mesh: M.Mesh
    E: float
    nu: float
    xi: float
    rho: float
    omegas: list = field(default_factory=list)

    @property
    def n_dof(self) -> int:
        return self.mesh.n_dof


def problems(count: int = 48, seed: int = 20261203):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        shape = ("cube", "sphere", "cylinder")[rng.integers(3)]
        n = int(2 * rng.integers(2, 9))                    # 4, 6, ..., 16
        if shape == "cube":
            m = M.static_cube_bcs(M.cube(n), traction=1.0)
        elif shape == "sphere":
            m = M.pressure_hemisphere_bcs(M.sphere(n), pressure=1.0)
        else:
            m = M.pressure_hemisphere_bcs(M.cylinder(n), pressure=1.0)
        E = float(10.0 ** rng.uniform(8.0, 10.0))
        nu = float(rng.uniform(0.2, 0.35))
        xi = float(rng.uniform(0.0, 0.1))
        rho = 2000.0
        nf = int(rng.integers(1, 17))
        cs = np.sqrt(E / (2.0 * (1.0 + nu)) / rho)
        omegas = [float(a0 * cs / 1.0) for a0 in rng.uniform(0.1, 5.0, nf)]
        out.append(Problem(shape, n, m, E, nu, xi, rho, omegas))
    return out


def jobs_of(systems, probs, buffers):
    """one assembly job per (problem, frequency), each into its own buffers"""
    jobs = []
    for s, p, bufs in zip(systems, probs, buffers):
        mu_r, mu_c, lam_c = M.elastic_constants(p.E, p.nu, p.xi)
        for w, (At, b) in zip(p.omegas, bufs):
            jobs.append(dict(system=s, At=At, b=b, omega=w, lam=lam_c, mu_c=mu_c, rho=p.rho,
                             nu=p.nu))
    return jobs


def run(probs, mode: int = bem.PREC_SPLIT16_REFINE, graph: bool = True, solve: bool = True,
        tune: bool = False):
    """Create the meshes, assemble every (problem, frequency) job (one graph launch, or one batch
    call), solve each system (with `tune`: per size class the autotuner's choice of mode and
    schedule, `mode` otherwise); returns (jobs, solutions, reports, record, kept objects)"""
    torch = bem._torch()
    t0 = time.perf_counter()
    systems = [bem.BemSystem(p.mesh) for p in probs]
    buffers = [[s.alloc(True) for _ in p.omegas] for s, p in zip(systems, probs)]
    jobs = jobs_of(systems, probs, buffers)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    g = None
    t0 = time.perf_counter()
    if graph:
        g = bem.BatchGraph(jobs)
        torch.cuda.synchronize()
    t_capture = time.perf_counter() - t0
    t0 = time.perf_counter()
    if graph:
        g.launch()
        g.check()
    else:
        bem.assemble_batch(jobs)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    xs, reps = [], []
    table, t_tune = None, 0.0
    if solve and tune:
        t0 = time.perf_counter()
        table = autotune.tune(jobs)
        t_tune = time.perf_counter() - t0
    t0 = time.perf_counter()
    if solve and tune:
        xs, reps = autotune.solve(jobs, table)
    elif solve:
        for j in jobs:
            x, rep = bem.gesv_device(j["At"], j["b"].clone(), mode)
            xs.append(x)
            reps.append(rep)
        torch.cuda.synchronize()
    t_sol = time.perf_counter() - t0
    rec = {"problems": len(probs), "jobs": len(jobs),
           "dof_total": int(sum(p.n_dof * len(p.omegas) for p in probs)),
           "entries_total": float(sum(p.n_dof ** 2 * len(p.omegas) for p in probs)),
           "setup_s": t_setup, "capture_s": t_capture, "assembly_s": t_asm, "solve_s": t_sol,
           "fell_back": int(sum(r["fell_back"] for r in reps))}
    if table is not None:
        rec["tune_s"] = t_tune
        rec["schedule"] = autotune.summary(table)
    return jobs, xs, reps, rec, (systems, buffers, g)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=48)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--mode", choices=["split16", "fp32", "fp64"], default="split16")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tune", action="store_true",
                    help="per size class, the autotuner's mode and schedule for the solves")
    args = ap.parse_args(argv)
    mode = {"split16": bem.PREC_SPLIT16_REFINE, "fp32": bem.PREC_FP32_REFINE,
            "fp64": bem.PREC_FP64}[args.mode]
    probs = problems(args.count, args.seed)
    # warm-up on the two smallest problems (kernels, solver workspaces)
    small = sorted(probs, key=lambda p: p.n_dof)[:2]
    run([Problem(p.shape, p.n, p.mesh, p.E, p.nu, p.xi, p.rho, p.omegas[:1]) for p in small],
        mode, graph=not args.no_graph)
    _, _, _, rec, keep = run(probs, mode, graph=not args.no_graph, tune=args.tune)
    rec["entries_per_s"] = rec["entries_total"] / (rec["assembly_s"] + rec["solve_s"])
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
