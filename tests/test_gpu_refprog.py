"""GPU assembly + solve against the reference program's own solutions (tests/golden/refprog.npz:
MultiFEBE's driver, built unmodified, on open surfaces with SBIE-MCA edges and on closed
surfaces) and against the oracle's matrices.

Bars: collocation points and pair classes identical to the oracle's; A within 1e-10 of the
largest entry of its row and b within 1e-10 of max |b| (oracle); solutions (FP64 LU, and FP32 LU
+ FP64 refinement) within 1e-8 of the reference program's, relative to the largest unknown."""
import numpy as np
import pytest

from multifebe_amd import mesh as M
from test_refprog_golden import CASES, load_case, unknowns

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bem():
    from multifebe_amd import bem as B
    B._torch()
    return B


@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_reference_program(bem, oracle, golden_dir, name):
    m, (E, nu, rho, xi), omegas, u, t = load_case(golden_dir, name)
    s = bem.BemSystem(m)
    xc, crow, cel = s.collocation()
    oxc, ocrow, ocel, _, _ = oracle.collocation(m)
    assert np.array_equal(crow, ocrow) and np.array_equal(cel, ocel)
    assert np.max(np.abs(xc - oxc)) <= 1e-15 * max(1.0, np.abs(m.x).max())
    cg, co = s.pair_classes(), oracle.pair_classes(m)
    bad = np.argwhere(cg != co)
    for e, q in bad[:20]:
        print("class differs: element", e, "point", q, "(node", crow[q], "element", cel[q], ")",
              "GPU", cg[e, q], "oracle", co[e, q], "x", xc[q], "element nodes", m.x[m.conn[e]].tolist(),
              flush=True)
    assert len(bad) == 0
    oms = [None] if len(omegas) == 0 else list(omegas)
    for k, om in enumerate(oms):
        if om is None:
            mu = E / (2 * (1 + nu))
            At, b = s.assemble_staela(mu, nu)
            Ao, bo = oracle.assemble(m, mu=mu, nu=nu)
            xr = unknowns(m, u, t)
        else:
            mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
            At, b = s.assemble_harela(om, lam_c, mu_c, rho, nu)
            Ao, bo = oracle.assemble(m, lam_c=lam_c, mu_c=mu_c, rho=rho, omega=om, nu=nu)
            xr = unknowns(m, u[k], t[k])
        A = At.T.cpu().numpy()
        ea = float(np.max(np.abs(A - Ao) / np.abs(Ao).max(axis=1, keepdims=True)))
        eb = float(np.max(np.abs(b.cpu().numpy() - bo)) / np.max(np.abs(bo)))
        assert ea <= 1e-10 and eb <= 1e-10, (ea, eb)
        for mode in (bem.PREC_FP64, bem.PREC_FP32_REFINE):
            x, rep = bem.gesv_device(At.clone(), b.clone(), mode)
            err = np.max(np.abs(x.cpu().numpy() - xr)) / np.max(np.abs(xr))
            print(name, om, "mode", mode, "A err", ea, "x err vs reference program", err, flush=True)
            assert err <= 1e-8
    s.close()


@pytest.mark.parametrize("name", CASES)
def test_gpu_interior_points_match_reference_program(bem, oracle, golden_dir, name):
    """interior displacements on the GPU (from the reference program's boundary solution) against
    the reference program's own interior values, and pair classes against the oracle's"""
    import os
    m, (E, nu, rho, xi), omegas, u, t = load_case(golden_dir, name)
    g = np.load(os.path.join(golden_dir, "refprog.npz"))
    pts, upr = g[f"{name}/points"], g[f"{name}/u_points"]
    s = bem.BemSystem(m)
    P = s.points(pts)
    assert np.array_equal(P.pair_classes(), oracle.pair_classes_points(m, pts))
    if len(omegas) == 0:
        mu = E / (2 * (1 + nu))
        up = P.displacement(u, t, mu=mu, nu=nu).cpu().numpy()
        err = np.max(np.abs(up - upr)) / np.max(np.abs(upr))
        print(name, "GPU interior u err", err, flush=True)
        assert err <= 1e-10
    else:
        mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
        for k, om in enumerate(omegas):
            up = P.displacement(u[k], t[k], omega=om, lam=lam_c, mu_c=mu_c, rho=rho).cpu().numpy()
            err = np.max(np.abs(up - upr[k])) / np.max(np.abs(upr[k]))
            print(name, om, "GPU interior u err", err, flush=True)
            assert err <= 1e-10
    # a point on a boundary node takes the node's displacement
    P2 = s.points(m.x[[5, 17]])
    uu = u if len(omegas) == 0 else u[0]
    tt = t if len(omegas) == 0 else t[0]
    kw = dict(mu=E / (2 * (1 + nu)), nu=nu) if len(omegas) == 0 else \
        dict(omega=omegas[0], lam=M.elastic_constants(E, nu, xi)[2],
             mu_c=M.elastic_constants(E, nu, xi)[1], rho=rho)
    assert np.array_equal(P2.displacement(uu, tt, **kw).cpu().numpy(), uu[[5, 17]])
    P.close()
    P2.close()
    s.close()
