"""GPU assembly (libmfbem, HIP on gfx950) against the reference's golden vectors and the oracle.

Bar: per-entry error <= 1e-10 x (largest entry of the row) for A, <= 1e-10 x max|b| for b, and the
integration class of every pair identical to the oracle's decision sequence."""
import os

import numpy as np
import pytest

from multifebe_amd import mesh as M

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def bem():
    from multifebe_amd import bem as B
    B._torch()
    return B


def to_np(At, b):
    return At.T.cpu().numpy(), b.cpu().numpy()


def rowrel(A, Ar):
    return float(np.max(np.abs(A - Ar) / np.abs(Ar).max(axis=1, keepdims=True)))


def golden(golden_dir):
    return np.load(os.path.join(golden_dir, "assemble.npz"))


@pytest.mark.parametrize("harm", [0, 1])
def test_pair_blocks_match_reference_golden(bem, golden_dir, harm):
    # 315 (element, point) pairs of the compiled reference's sbie_auto: flat, curved, distorted
    # elements; on-element (singular), near (adaptive) and far (precalculated) points
    g = np.load(os.path.join(golden_dir, "pairs.npz"))
    sel = np.where((g["harm"] == harm) & (g["rev"] == 0))[0]
    for omega in sorted(set(g["omega"][sel])):
        idx = sel[g["omega"][sel] == omega]
        if harm:
            h, gg, cls = bem.pair_blocks(g["X"][idx], g["xi"][idx], omega=float(omega),
                                         lam=complex(g["lam_c"]), mu_c=complex(g["mu_c"]), rho=float(g["rho"]))
        else:
            h, gg, cls = bem.pair_blocks(g["X"][idx], g["xi"][idx], mu=float(g["mu"]), nu=float(g["nu"]))
        tr = g["trace"][idx]
        # reference trace -> class code: ps (1-based) -> ps-1; adaptive -> 8; singular -> 9
        ref_cls = np.where(tr[:, 0] == 1, 9, np.where(tr[:, 3] > 0, tr[:, 3] - 1, 8))
        worst = []
        for j, i in enumerate(idx):
            hr = g["h"][i].reshape(3, 3, 9).transpose(2, 1, 0)
            gr = g["g"][i].reshape(3, 3, 9).transpose(2, 1, 0)
            eh = np.max(np.abs(h[j] - hr)) / np.max(np.abs(hr))
            eg = np.max(np.abs(gg[j] - gr)) / np.max(np.abs(gr))
            worst.append((max(eh, eg), int(cls[j]), int(i)))
        worst.sort(reverse=True)
        print("harm", harm, "omega", omega, "worst", worst[:5], flush=True)
        assert np.array_equal(cls, ref_cls)
        assert worst[0][0] <= TOL


@pytest.mark.parametrize("name", ["cube4", "sphere4", "sphere8", "cylinder6"])
def test_pair_classes_match_oracle(bem, oracle, name):
    m = {"cube4": M.cube(4), "sphere4": M.sphere(4), "sphere8": M.sphere(8),
         "cylinder6": M.cylinder(6)}[name]
    m = M.static_cube_bcs(m)
    sysm = bem.BemSystem(m)
    cg = sysm.pair_classes()
    co = oracle.pair_classes(m)
    print(name, "classes", np.bincount(cg.ravel()), flush=True)
    assert np.array_equal(cg, co)


@pytest.mark.parametrize("n", [2, 4])
def test_static_matches_reference_golden(bem, golden_dir, n):
    g = golden(golden_dir)
    m = M.static_cube_bcs(M.cube(n))
    At, b = bem.BemSystem(m).assemble_staela(float(g["mu"]), float(g["nu"]))
    A, b = to_np(At, b)
    Ar, br = g["sta_cube%d_A" % n], g["sta_cube%d_b" % n]
    print("cube", n, "A err", rowrel(A, Ar), flush=True)
    assert rowrel(A, Ar) <= TOL
    assert np.max(np.abs(b - br)) <= TOL * np.abs(br).max()


@pytest.mark.parametrize("a0", ["0.1", "4"])
def test_harmonic_matches_reference_golden(bem, golden_dir, a0):
    g = golden(golden_dir)
    key = "har_sphere4_a" + a0
    m = M.pressure_hemisphere_bcs(M.sphere(4))
    At, b = bem.BemSystem(m).assemble_harela(float(g[key + "_omega"]), complex(g["lam_c"]),
                                             complex(g["mu_c"]), float(g["rho"]), float(g["nuh"]))
    A, b = to_np(At, b)
    Ar, br = g[key + "_A"], g[key + "_b"]
    print(key, "A err", rowrel(A, Ar), flush=True)
    assert rowrel(A, Ar) <= TOL
    assert np.max(np.abs(b - br)) <= TOL * np.abs(br).max()


@pytest.mark.parametrize("name", ["sphere8", "cylinder6", "cube8"])
def test_static_matches_oracle(bem, oracle, name):
    m = {"sphere8": M.sphere(8), "cylinder6": M.cylinder(6), "cube8": M.cube(8)}[name]
    m = M.static_cube_bcs(m)
    mu, nu = 4e8, 0.25
    At, b = bem.BemSystem(m).assemble_staela(mu, nu)
    A, b = to_np(At, b)
    Ao, bo = oracle.assemble(m, mu=mu, nu=nu)
    print(name, "A err", rowrel(A, Ao), flush=True)
    assert rowrel(A, Ao) <= TOL
    assert np.max(np.abs(b - bo)) <= TOL * np.abs(bo).max()


@pytest.mark.parametrize("a0", [0.5, 3.0])
def test_harmonic_matches_oracle(bem, oracle, a0):
    m = M.pressure_hemisphere_bcs(M.sphere(8))
    E, nu, xi, rho = 1e9, 0.3, 0.05, 2000.0
    mu_r, mu_c, lam_c = M.elastic_constants(E, nu, xi)
    omega = a0 * np.sqrt(mu_r / rho)
    At, b = bem.BemSystem(m).assemble_harela(omega, lam_c, mu_c, rho, nu)
    A, b = to_np(At, b)
    Ao, bo = oracle.assemble(m, lam_c=lam_c, mu_c=mu_c, rho=rho, omega=omega, nu=nu)
    print("sphere8 a0", a0, "A err", rowrel(A, Ao), flush=True)
    assert rowrel(A, Ao) <= TOL
    assert np.max(np.abs(b - bo)) <= TOL * np.abs(bo).max()


def test_stats_and_repeatability(bem):
    m = M.pressure_hemisphere_bcs(M.sphere(6))
    s = bem.BemSystem(m)
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    A1, b1 = s.assemble_harela(300.0, lam_c, mu_c, 2000.0, 0.3)
    A1 = A1.clone(); b1 = b1.clone()
    A2, b2 = s.assemble_harela(300.0, lam_c, mu_c, 2000.0, 0.3)
    st = s.stats()
    assert st["pairs"] == m.nn * m.ne
    assert st["singular"] == 9 * m.ne - 0 or st["singular"] > 0
    assert st["pre"] + st["adaptive"] + st["singular"] == st["pairs"]
    d = (A1 - A2).abs().max().item() / A1.abs().max().item()
    assert d <= 1e-14


# flat quad9 [-1, 1]^2 in z = 0, reference local node order
_XI_Q9 = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)]


@pytest.mark.parametrize("harm", [0, 1])
def test_deep_adaptive_trees_have_no_leaf_cap(bem, oracle, harm):
    """points 1e-5 .. 1e-8 above the centre of a flat element: the reference's subdivision
    (ext_adp) produces 172 leaves at qsi_relative_error 1e-10 (more than one batch of 128); the
    GPU integrates them all, in the recursion's order"""
    X = np.array([[a, b, 0.0] for a, b in _XI_Q9])
    pts = np.array([[0.0, 0.0, 1e-5], [0.0, 0.0, 1e-8], [1.0, 0.0, 1e-6], [0.3, 0.2, 1e-4]])
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    omega = 3.0 * np.sqrt(mu_r / 2000.0)
    s = bem.Settings(qsi_relative_error=1e-10)
    kw = dict(omega=omega, lam=lam_c, mu_c=mu_c, rho=2000.0) if harm else dict(mu=mu_r, nu=0.3)
    h, g, cls = bem.pair_blocks(np.repeat(X[None], len(pts), 0), pts, settings=s, **kw)
    leaves = []
    for j, x_i in enumerate(pts):
        okw = dict(omega=omega, lam_c=lam_c, mu_c=mu_c, rho=2000.0) if harm else dict(mu=mu_r, nu=0.3)
        hr, gr, tr, dtr = oracle.pair(X, x_i, qsi_re=1e-10, **okw)
        leaves.append(int(dtr[3]))
        eh = np.max(np.abs(h[j] - hr)) / np.max(np.abs(hr))
        eg = np.max(np.abs(g[j] - gr)) / np.max(np.abs(gr))
        print("point", x_i, "leaves", leaves[-1], "class", cls[j], "err", eh, eg, flush=True)
        assert cls[j] == 8 and eh <= TOL and eg <= TOL
    assert max(leaves) > 128
