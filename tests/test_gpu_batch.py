"""Batched uneven problems (cubes, spheres, cylinders at several materials and frequencies): the
assembly of every (problem, frequency) job captured once into a HIP graph and replayed, against the
per-job assembly, the oracle and LAPACK; the seeded 48-problem configuration assembled by one graph
launch and solved, every solution checked by its residual.

Graph replays are checked several times, with other GPU work and synchronisations in between and
new materials / frequencies written between launches."""
import numpy as np
import pytest
import scipy.linalg as sla

from multifebe_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    from multifebe_amd import batch, bem
    bem._torch()
    return batch, bem


def small_batch(batch):
    """the seeded generator's first two problems of each shape with n <= 8, at most two
    frequencies each"""
    out, per = [], {}
    for p in batch.problems(48, 20261203):
        if p.n <= 8 and per.get(p.shape, 0) < 2:
            per[p.shape] = per.get(p.shape, 0) + 1
            p.omegas = p.omegas[:2]
            out.append(p)
    return sorted(out, key=lambda p: p.n_dof)


def eager(bem, job):
    s = job["system"]
    return s.assemble_harela(job["omega"], job["lam"], job["mu_c"], job["rho"], job["nu"])


def test_graph_replays_match_eager_oracle_and_lapack(B, oracle):
    batch, bem = B
    import torch
    probs = small_batch(batch)
    assert {p.shape for p in probs} == {"cube", "sphere", "cylinder"}
    systems = [bem.BemSystem(p.mesh) for p in probs]
    buffers = [[s.alloc(True) for _ in p.omegas] for s, p in zip(systems, probs)]
    jobs = batch.jobs_of(systems, probs, buffers)
    print("jobs", len(jobs), "dof", [j["system"].n_dof for j in jobs], flush=True)
    g = bem.BatchGraph(jobs)
    for rep in range(3):
        if rep:
            # new frequencies and materials for the same jobs, then other GPU work in between
            for j in jobs:
                j["omega"] *= 1.3
                j["mu_c"] *= 1.1
                j["lam"] *= 1.1
            g.set_params(jobs)
            other, _ = eager(bem, jobs[0])
            (other @ other[:, :3].contiguous()).sum().item()
            torch.cuda.synchronize()
        g.launch()
        g.check()
        torch.cuda.synchronize()
        for q, j in enumerate(jobs):
            Ae, be = eager(bem, j)
            A, b = j["At"].T.cpu().numpy(), j["b"].cpu().numpy()
            ea = np.max(np.abs(A - Ae.T.cpu().numpy())) / np.max(np.abs(A))
            eb = np.max(np.abs(b - be.cpu().numpy())) / np.max(np.abs(b))
            assert ea <= 1e-13 and eb <= 1e-13, (rep, q, ea, eb)
        print("replay", rep, "equal to the per-job assemblies", flush=True)
    # the oracle on the first jobs (the smallest meshes), LAPACK on every job
    for q, j in enumerate(jobs):
        A, b = j["At"].T.cpu().numpy(), j["b"].cpu().numpy()
        if q < 2:
            p = probs[[k for k, s in enumerate(systems) if s is j["system"]][0]]
            Ao, bo = oracle.assemble(p.mesh, lam_c=j["lam"], mu_c=j["mu_c"], rho=j["rho"],
                                     omega=j["omega"], nu=j["nu"])
            ea = np.max(np.abs(A - Ao) / np.abs(Ao).max(axis=1, keepdims=True))
            eb = np.max(np.abs(b - bo)) / np.max(np.abs(bo))
            print("job", q, p.shape, p.n, "A err vs oracle", ea, "b err", eb, flush=True)
            assert ea <= 1e-10 and eb <= 1e-10
        xr = sla.solve(A, b)
        for mode, tol in ((bem.PREC_FP64, 1e-10), (bem.PREC_SPLIT16_REFINE, 1e-8)):
            x, rep = bem.gesv_device(j["At"].clone(), j["b"].clone(), mode)
            err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
            assert err <= tol and rep["fell_back"] == 0, (q, mode, err, rep)
    # the plain batch call (no graph) gives the same matrices
    saved = [(j["At"].clone(), j["b"].clone()) for j in jobs]
    bem.assemble_batch(jobs)
    torch.cuda.synchronize()
    for (A0, b0), j in zip(saved, jobs):
        assert torch.max(torch.abs(A0 - j["At"])).item() <= 1e-13 * torch.max(torch.abs(A0)).item()
    g.close()


def test_seeded_48_problem_batch(B):
    """the batched configuration at its own size: 48 problems, one graph launch for all their
    frequencies, every system solved (split FP16 LU + FP64 refinement) and checked by its residual"""
    batch, bem = B
    import torch
    probs = batch.problems(48, 20261203)
    jobs, xs, reps, rec, keep = batch.run(probs, bem.PREC_SPLIT16_REFINE, graph=True)
    print(rec, flush=True)
    assert rec["jobs"] == sum(len(p.omegas) for p in probs) and rec["fell_back"] == 0
    worst = 0.0
    for j, x in zip(jobs, xs):
        A = j["At"].T
        r = j["b"] - A @ x
        be = (torch.linalg.vector_norm(r, ord=float("inf"))
              / (torch.linalg.matrix_norm(A, ord=float("inf")) * torch.linalg.vector_norm(x, ord=float("inf"))
                 + torch.linalg.vector_norm(j["b"], ord=float("inf")))).item()
        worst = max(worst, be)
    print("largest backward error", worst, flush=True)
    assert worst <= 1e-14


def test_autotuned_solves(B):
    """the autotuner on a seeded subset of the batch: one entry per size class, its choice one of
    the candidates and the fastest of those that refine to FP64 accuracy; every system then solved
    with its class's choice to backward error <= 1e-14; the default schedule restored"""
    from multifebe_amd import autotune
    batch, bem = B
    import torch
    probs = [p for p in batch.problems(48, 20261203) if p.n in (6, 12)][:4]
    for p in probs:
        p.omegas = p.omegas[:2]
    jobs, _, _, rec, keep = batch.run(probs, graph=True, solve=False)
    torch.cuda.synchronize()
    cands = autotune.candidates()
    table = autotune.tune(jobs, repeats=2)
    print(autotune.summary(table), flush=True)
    classes = {autotune.size_class(j["system"].n_dof) for j in jobs}
    assert set(table) == classes
    for edge, e in table.items():
        assert e["choice"] in cands and len(e["tried"]) == len(cands)
        ok = [r for r in e["tried"] if r[3] == 0 and r[2] <= autotune.MAX_BACKWARD_ERROR]
        assert e["ms"] == min(r[1] for r in ok)
    assert bem.solver_schedule() == {"lookahead": -1, "la_chunk": -1, "trsv_parts": -1}
    xs, reps = autotune.solve(jobs, table)
    for j, x, r in zip(jobs, xs, reps):
        be = autotune.backward_error(j["At"], j["b"], x)
        assert r["fell_back"] == 0 and be <= 1e-14, (j["system"].n_dof, be, r)
    assert bem.solver_schedule() == {"lookahead": -1, "la_chunk": -1, "trsv_parts": -1}
