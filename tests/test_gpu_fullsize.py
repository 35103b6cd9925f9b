"""The benchmark configurations at their own sizes, GPU against the oracle (C restatement of the
reference, pinned to the compiled reference's golden vectors) and LAPACK:

  * static cube: side 1 m, 12 intervals per edge (866 nodes, 2,598 DOF), E = 1e9 Pa, nu = 0.25,
    bottom fixed, top normal traction -1e3 Pa, other faces free;
  * harmonic sphere: radius 1 m, 24 intervals per edge (3,458 nodes, 10,374 complex DOF),
    E = 1e9 (1 + 0.1 i) Pa, nu = 0.3, rho = 2000 kg/m3, unit pressure on z > 0, z <= 0 fixed,
    at a0 = 0.1 and 4.

Bars: identical integration class of every pair; A within 1e-10 of the largest entry of its row,
b within 1e-10 of max |b|; solutions within 1e-8 (FP64 and FP32 + FP64 refinement) of LAPACK's
solution of the oracle's system."""
import numpy as np
import pytest
import scipy.linalg as sla

from multifebe_amd import mesh as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bem():
    from multifebe_amd import bem as B
    B._torch()
    return B


def _compare(sysm, oracle_cls, At, b, Ao, bo, label):
    cg = sysm.pair_classes()
    assert np.array_equal(cg, oracle_cls), f"{label}: pair classes differ"
    A = At.T.cpu().numpy()
    bb = b.cpu().numpy()
    ea = float(np.max(np.abs(A - Ao) / np.abs(Ao).max(axis=1, keepdims=True)))
    eb = float(np.max(np.abs(bb - bo)) / np.max(np.abs(bo)))
    print(f"{label}: classes {np.bincount(cg.ravel())}, A err {ea:.2e}, b err {eb:.2e}", flush=True)
    assert ea <= 1e-10 and eb <= 1e-10


def test_static_cube12_full_size(bem, oracle):
    m = M.static_cube_bcs(M.cube(12), traction=1e3)
    assert m.n_dof == 2598
    E, nu = 1e9, 0.25
    mu = E / (2 * (1 + nu))
    Ao, bo = oracle.assemble(m, mu=mu, nu=nu)
    s = bem.BemSystem(m)
    At, b = s.assemble_staela(mu, nu)
    _compare(s, oracle.pair_classes(m), At, b, Ao, bo, "cube12")
    xo = sla.solve(Ao, bo)
    x, rep = bem.gesv_device(At.clone(), b.clone(), bem.PREC_FP64)
    ex = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    print(f"cube12 FP64 solution err vs LAPACK on the oracle system {ex:.2e}", flush=True)
    assert ex <= 1e-8
    s.close()


@pytest.mark.parametrize("a0", [0.1, 4.0])
def test_harmonic_sphere24_full_size(bem, oracle, a0):
    m = M.pressure_hemisphere_bcs(M.sphere(24), pressure=1.0)
    assert m.n_dof == 10374
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    omega = a0 * np.sqrt(mu_r / 2000.0) / 1.0
    Ao, bo = oracle.assemble(m, lam_c=lam_c, mu_c=mu_c, rho=2000.0, omega=omega, nu=0.3)
    s = bem.BemSystem(m)
    At, b = s.assemble_harela(omega, lam_c, mu_c, 2000.0, 0.3)
    _compare(s, oracle.pair_classes(m), At, b, Ao, bo, f"sphere24 a0={a0}")
    xo = sla.solve(Ao, bo)
    for mode in (bem.PREC_FP64, bem.PREC_FP32_REFINE):
        x, rep = bem.gesv_device(At.clone(), b.clone(), mode)
        ex = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
        print(f"sphere24 a0={a0} mode {mode}: solution err vs LAPACK {ex:.2e} {rep}", flush=True)
        assert ex <= 1e-8
    s.close()
