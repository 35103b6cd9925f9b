"""Frequency sweep of the harmonic system (BemSystem.sweep_harela): the overlapped pipeline (the
assembly of frequency k+1 on a side stream while frequency k is solved) gives the solutions of
the plain sequence assemble -> solve at every frequency."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_overlapped_sweep_matches_sequential_solves():
    import torch
    from multifebe_amd import bem, mesh as M
    m = M.pressure_hemisphere_bcs(M.sphere(4, 1.0), pressure=1.0)
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    rho, nu = 2000.0, 0.3
    cs = np.sqrt(mu_r / rho)
    omegas = [0.5 * cs, 1.7 * cs, 3.1 * cs]
    sysm = bem.BemSystem(m)
    for mode in (bem.PREC_FP64, bem.PREC_FP32_REFINE):
        res = sysm.sweep_harela(omegas, lam_c, mu_c, rho, nu, mode, overlap=True)
        assert len(res) == len(omegas)
        for w, (x, rep, t_asm, t_sol, rec) in zip(omegas, res):
            At, b = sysm.assemble_harela(w, lam_c, mu_c, rho, nu)
            xr, _ = bem.gesv_device(At, b.clone(), bem.PREC_FP64)
            err = (torch.linalg.vector_norm(x - xr) / torch.linalg.vector_norm(xr)).item()
            print(mode, w / cs, err, rep["refinement_steps"], rec["timings"]["regular_launches"],
                  flush=True)
            assert err <= 1e-10
            assert rec["stats"]["pairs"] == m.nn * m.ne
            assert rec["timings"]["regular_ms"] > 0
    sysm.close()
