"""The learned-quadrature workload's geometry: a thin slab 1 x 1 x 0.01 m (quad9 of 0.05 m, 3,522
nodes, 880 elements; opposite faces 0.01 m apart, so most pairs across the slab are
quasi-singular) bounding the unbounded region around it, the slab moved rigidly (u prescribed),
and 20,000 points at distances log-uniform in [1e-3, 1e-1] m from the surface.  Static and
harmonic (a0 = omega L / c_s = 3 with L = 1 m).

GPU against the oracle: pair classes of a sample of the points identical; boundary solution (GPU
assembly + FP64 LU) within 1e-8 of LAPACK on the oracle's system; displacements at a sample of
the points within 1e-8 of the oracle's (from the oracle's boundary solution)."""
import numpy as np
import pytest
import scipy.linalg as sla

from multifebe_amd import mesh as M

pytestmark = pytest.mark.gpu
E, NU, RHO, XI = 1e9, 0.3, 2000.0, 0.05


@pytest.fixture(scope="module")
def bem():
    from multifebe_amd import bem as B
    B._torch()
    return B


def slab_problem():
    m = M.slab(0.05, 1.0, 0.01, exterior=True)
    m.ctype = np.zeros((m.nn, 3), dtype=np.int32)       # u prescribed: a rigid translation
    m.cval = np.tile(np.array([0.0, 0.0, 1e-3]), (m.nn, 1))
    pts = M.points_off_surface(m, 20000, 1e-3, 1e-1, outward=False)
    return m, pts


@pytest.mark.parametrize("harmonic", [False, True])
def test_slab_interior_points(bem, oracle, harmonic):
    m, pts = slab_problem()
    mu_r, mu_c, lam_c = M.elastic_constants(E, NU, XI)
    omega = 3.0 * np.sqrt(mu_r / RHO) / 1.0
    s = bem.BemSystem(m)
    if harmonic:
        At, b = s.assemble_harela(omega, lam_c, mu_c, RHO, NU)
        Ao, bo = oracle.assemble(m, lam_c=lam_c, mu_c=mu_c, rho=RHO, omega=omega, nu=NU)
        kw_gpu = dict(omega=omega, lam=lam_c, mu_c=mu_c, rho=RHO)
        kw_orc = dict(omega=omega, lam_c=lam_c, mu_c=mu_c, rho=RHO)
    else:
        At, b = s.assemble_staela(mu_r, NU)
        Ao, bo = oracle.assemble(m, mu=mu_r, nu=NU)
        kw_gpu = kw_orc = dict(mu=mu_r, nu=NU)
    ea = float(np.max(np.abs(At.T.cpu().numpy() - Ao) / np.abs(Ao).max(axis=1, keepdims=True)))
    xo = sla.solve(Ao, bo)
    x, rep = bem.gesv_device(At, b.clone(), bem.PREC_FP64)
    ex = np.linalg.norm(x.cpu().numpy() - xo) / np.linalg.norm(xo)
    print("slab", "harmonic" if harmonic else "static", "A err", ea, "x err", ex, flush=True)
    assert ea <= 1e-10 and ex <= 1e-8
    u, t = s.boundary_fields(x)
    uo, to = s.boundary_fields(xo)
    P = s.points(pts)
    st = P.stats()
    print("20k points: pairs", st["pairs"], "adaptive", st["adaptive"], "precalculated", st["pre"],
          flush=True)
    up = P.displacement(u, t, **kw_gpu).cpu().numpy()
    sample = np.arange(0, len(pts), 100)                 # 200 points through the oracle
    cls = P.pair_classes()[:, sample]
    assert np.array_equal(cls, oracle.pair_classes_points(m, pts[sample]))
    upo = oracle.internal_points(m, pts[sample], uo, to, **kw_orc)
    err = np.max(np.abs(up[sample] - upo)) / np.max(np.abs(upo))
    print("interior displacement err vs oracle", err, flush=True)
    assert err <= 1e-8
    P.close()
    s.close()
