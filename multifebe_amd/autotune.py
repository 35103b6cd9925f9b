"""Autotuner of the batched configuration's dense solves.

For each class of system size, it picks the precision mode and the solve schedule that solve the
class's largest system fastest. The schedule covers look-ahead on/off, the columns per
second-stream trailing-update GEMM, and the parts per row block of the persistent triangular
solves. Only candidates that solve that system to FP64 accuracy are eligible: backward error
<= 1e-14 and no fallback.

    table = autotune.tune(jobs)             # jobs as batch.jobs_of builds them, assembled
    xs, reps = autotune.solve(jobs, table)  # every job with its class's choice

The reference solves every system with LAPACK ?gesv (src/solve_lse_c.f90:124) and has nothing
to tune. Here the choice changes speed only: FP64 mode keeps LAPACK's pivots under every
schedule, and the mixed modes refine to the same accuracy bar.
"""
from __future__ import annotations

import math
import time

from . import bem

# upper bounds (DOF) of the size classes
CLASS_EDGES = (1024, 2048, 4096, 8192, 16384, 1 << 30)
MODES = {"fp64": bem.PREC_FP64, "fp32": bem.PREC_FP32_REFINE,
         "split16": bem.PREC_SPLIT16_REFINE}
MAX_BACKWARD_ERROR = 1e-14


def size_class(n_dof: int) -> int:
    for e in CLASS_EDGES:
        if n_dof <= e:
            return e
    return CLASS_EDGES[-1]


def candidates(modes=("fp64", "fp32", "split16")):
    """(mode, look-ahead, chunk, triangular-solve parts); the chunk width only matters with
    look-ahead on"""
    out = []
    for m in modes:
        for la, ch in ((1, 0), (1, 256), (0, 0)):
            for tp in (4, 1):
                out.append({"mode": m, "lookahead": la, "la_chunk": ch, "trsv_parts": tp})
    return out


def backward_error(At, b, x) -> float:
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) for A = At.T, on the GPU"""
    torch = bem._torch()
    A = At.T
    inf = float("inf")
    r = b - A @ x
    den = (torch.linalg.matrix_norm(A, ord=inf) * torch.linalg.vector_norm(x, ord=inf)
           + torch.linalg.vector_norm(b, ord=inf))
    return (torch.linalg.vector_norm(r, ord=inf) / den).item()


def _apply(choice):
    bem.set_solver_schedule(choice["lookahead"], choice["la_chunk"], choice["trsv_parts"])


def time_candidate(At, b, cand, repeats: int = 3):
    """best wall time (s) of `repeats` solves after one untimed solve (it grows the workspaces),
    the backward error of the last solution and the OR of the fallback flags"""
    torch = bem._torch()
    _apply(cand)
    best, fell, x = math.inf, 0, None
    for k in range(repeats + 1):
        A = At.clone()
        x = b.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, rep = bem.gesv_device(A, x, MODES[cand["mode"]])
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        if k:
            best = min(best, t)
        fell |= rep["fell_back"]
        del A
    return best, backward_error(At, b, x), fell


def tune(jobs, cands=None, repeats: int = 3) -> dict:
    """{class edge: {"choice", "ms", "n_dof", "tried": [(candidate, ms, backward error,
    fell_back)]}}, timed on each class's largest system; the schedule is reset afterwards"""
    cands = candidates() if cands is None else cands
    by_class = {}
    for j in jobs:
        by_class.setdefault(size_class(j["system"].n_dof), []).append(j)
    table = {}
    try:
        for edge in sorted(by_class):
            j = max(by_class[edge], key=lambda q: q["system"].n_dof)
            tried = []
            for c in cands:
                t, be, fell = time_candidate(j["At"], j["b"], c, repeats)
                tried.append((dict(c), t * 1e3, be, int(fell)))
            ok = [r for r in tried if r[3] == 0 and r[2] <= MAX_BACKWARD_ERROR]
            if not ok:
                raise RuntimeError(f"autotune: no candidate solves the {edge}-DOF class's "
                                   f"system to {MAX_BACKWARD_ERROR}")
            best = min(ok, key=lambda r: r[1])
            table[edge] = {"choice": best[0], "ms": best[1], "n_dof": j["system"].n_dof,
                           "tried": tried}
    finally:
        bem.set_solver_schedule(-1, -1, -1)
    return table


def solve(jobs, table, default=None):
    """every job solved with its class's choice (classes missing from the table: `default`, a
    candidate dict, or split16 with the built-in schedule); b is left unchanged"""
    torch = bem._torch()
    default = default or {"mode": "split16", "lookahead": -1, "la_chunk": -1, "trsv_parts": -1}
    xs, reps = [None] * len(jobs), [None] * len(jobs)
    order = sorted(range(len(jobs)), key=lambda q: size_class(jobs[q]["system"].n_dof))
    try:
        cur = None
        for q in order:
            j = jobs[q]
            entry = table.get(size_class(j["system"].n_dof))
            choice = entry["choice"] if entry else default
            if choice is not cur:
                _apply(choice)
                cur = choice
            A = j["At"] if choice["mode"] != "fp64" else j["At"].clone()
            xs[q], reps[q] = bem.gesv_device(A, j["b"].clone(), MODES[choice["mode"]])
        torch.cuda.synchronize()
    finally:
        bem.set_solver_schedule(-1, -1, -1)
    return xs, reps


def summary(table) -> dict:
    """the choices and their times, without the list of everything tried"""
    return {str(e): {"n_dof": v["n_dof"], "ms": round(v["ms"], 3), **v["choice"]}
            for e, v in table.items()}
