"""GPU dense solve (blocked LU: FP64, FP32 + FP64 refinement, FP32 with split FP16 trailing
updates + FP64 refinement) against LAPACK (numpy / scipy)."""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bem():
    from multifebe_amd import bem as B
    B._torch()
    return B


def dev(M):
    import torch
    # column-major storage: the tensor holds M^T row-major
    return torch.from_numpy(np.ascontiguousarray(M.T)).cuda()


def rand(n, cplx, seed, diag=0.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    if cplx:
        A = A + 1j * rng.standard_normal((n, n))
    return A + diag * np.eye(n)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 2, 31, 32, 33, 255, 256, 257, 700, 1300])
def test_fp64_matches_lapack(bem, n, cplx):
    import torch
    A = rand(n, cplx, n)
    b = rand(n, cplx, n + 1)[:, 0]
    At = dev(A)
    bt = torch.from_numpy(b.copy()).cuda()
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    x, rep = bem.gesv_device(At, bt, bem.PREC_FP64, ipiv=ipiv)
    xr = np.linalg.solve(A, b)
    err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    print(n, cplx, "err", err, rep, flush=True)
    assert err <= 1e-11 * max(1.0, np.linalg.cond(A) / 1e3)
    # same pivot sequence as LAPACK getrf (1-based)
    lu, piv = sla.lu_factor(A)
    assert np.array_equal(ipiv.cpu().numpy(), piv + 1)
    # same factors
    LU = At.T.cpu().numpy()
    assert np.max(np.abs(LU - lu)) <= 1e-10 * np.max(np.abs(lu))


@pytest.mark.parametrize("cplx", [False, True])
def test_multiple_rhs(bem, cplx):
    import torch
    n, k = 600, 5
    A = rand(n, cplx, 3)
    B = rand(n, cplx, 4)[:, :k]
    At = dev(A)
    Bt = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()  # (k, n): column-major n x k
    x, rep = bem.gesv_device(At, Bt, bem.PREC_FP64)
    X = x.cpu().numpy().T
    Xr = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) <= 1e-11


def test_singular_matrix_reports_info(bem):
    import torch
    n = 300
    A = rand(n, False, 9)
    A[:, 77] = 0.0           # exactly singular: zero column -> zero pivot at column 78
    b = np.ones(n)
    with pytest.raises(Exception) as ei:
        bem.gesv_device(dev(A), torch.from_numpy(b).cuda(), bem.PREC_FP64)
    assert "singular" in str(ei.value)


@pytest.mark.parametrize("cplx", [False, True])
def test_mixed_precision_refines_to_fp64(bem, cplx):
    import torch
    n = 1500
    A = rand(n, cplx, 21, diag=3 * np.sqrt(n))
    b = rand(n, cplx, 22)[:, 0]
    At = dev(A)
    x, rep = bem.gesv_device(At, torch.from_numpy(b.copy()).cuda(), bem.PREC_FP32_REFINE)
    xr = np.linalg.solve(A, b)
    err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    print("mixed", cplx, err, rep, flush=True)
    assert rep["precision_mode"] == bem.PREC_FP32_REFINE and rep["fell_back"] == 0
    assert rep["refinement_steps"] >= 1
    assert err <= 1e-12
    # the FP64 matrix is left intact in mixed mode
    assert np.array_equal(At.T.cpu().numpy(), A)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("diag", [3.0, 0.0])
def test_split16_refines_to_fp64(bem, cplx, diag):
    """FP32 LU whose trailing updates are one FP16 GEMM of hi/lo products: refined to FP64
    accuracy in a few steps, on a diagonally dominant and on a plain Gaussian matrix (cond ~1e4);
    n = 2000 spans 8 panels (look-ahead, short last panel)"""
    import torch
    n = 2000
    A = rand(n, cplx, 31, diag=diag * np.sqrt(n))
    b = rand(n, cplx, 32)[:, 0]
    At = dev(A)
    x, rep = bem.gesv_device(At, torch.from_numpy(b.copy()).cuda(), bem.PREC_SPLIT16_REFINE)
    xr = np.linalg.solve(A, b)
    err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    x32, rep32 = bem.gesv_device(dev(A), torch.from_numpy(b.copy()).cuda(), bem.PREC_FP32_REFINE)
    print("split16", cplx, diag, err, rep, "fp32 steps", rep32["refinement_steps"], flush=True)
    assert rep["precision_mode"] == bem.PREC_SPLIT16_REFINE and rep["fell_back"] == 0
    assert 1 <= rep["refinement_steps"] <= rep32["refinement_steps"] + 3
    assert err <= 1e-11 * max(1.0, np.linalg.cond(A) / 1e3)
    assert np.array_equal(At.T.cpu().numpy(), A)


@pytest.mark.parametrize("mode", ["split16", "fp32"])
def test_expired_waits_are_recovered(bem, mode):
    """the persistent kernels' waits are bounded; when one expires (forced from the host: spin
    limit 0, so a wait that does not find its tag at once expires) the mixed solve redoes the
    factorisation with one launch per column and the triangular solves with one launch per block
    step, reports it (fell_back bit 2) and returns the same solution; with the limit restored,
    nothing is redone"""
    import os
    import torch
    pm = {"split16": bem.PREC_SPLIT16_REFINE, "fp32": bem.PREC_FP32_REFINE}[mode]
    n = 3000                      # 12 panel workgroups: waits on other workgroups' posts
    A = rand(n, True, 41, diag=3.0 * np.sqrt(n))
    b = rand(n, True, 42)[:, 0]
    xr = np.linalg.solve(A, b)
    os.environ["MFB_SPIN_LIMIT"] = "0"
    try:
        x, rep = bem.gesv_device(dev(A), torch.from_numpy(b.copy()).cuda(), pm)
    finally:
        del os.environ["MFB_SPIN_LIMIT"]
    err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)
    print("expired waits", mode, err, rep, flush=True)
    assert rep["fell_back"] & 2 and rep["precision_mode"] == pm
    assert err <= 1e-11
    x2, rep2 = bem.gesv_device(dev(A), torch.from_numpy(b.copy()).cuda(), pm)
    err2 = np.linalg.norm(x2.cpu().numpy() - xr) / np.linalg.norm(xr)
    assert rep2["fell_back"] == 0 and err2 <= 1e-11


def test_mixed_precision_falls_back_when_ill_conditioned(bem):
    import torch
    n = 400
    rng = np.random.default_rng(5)
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.logspace(0, -12, n)            # cond 1e12: FP32 LU cannot be refined
    A = (U * s) @ V.T
    b = rng.standard_normal(n)
    x, rep = bem.gesv_device(dev(A), torch.from_numpy(b.copy()).cuda(), bem.PREC_FP32_REFINE)
    print("fallback", rep, flush=True)
    assert rep["fell_back"] == 1 and rep["precision_mode"] == bem.PREC_FP64
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr) <= 1e-3


def test_bem_system_mixed_matches_fp64_lapack(bem, oracle):
    # harmonic BEM matrix of the sphere: FP32 LU + FP64 refinement within 1e-8 of LAPACK zgesv
    import torch
    from multifebe_amd import mesh as M
    m = M.pressure_hemisphere_bcs(M.sphere(8))
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    omega = 2.0 * np.sqrt(mu_r / 2000.0)
    s = bem.BemSystem(m)
    At, b = s.assemble_harela(omega, lam_c, mu_c, 2000.0, 0.3)
    A = At.T.cpu().numpy()
    bb = b.cpu().numpy()
    xr = sla.solve(A, bb)
    x32, rep = bem.gesv_device(At.clone(), b.clone(), bem.PREC_FP32_REFINE)
    x16, rep16 = bem.gesv_device(At.clone(), b.clone(), bem.PREC_SPLIT16_REFINE)
    x64, rep64 = bem.gesv_device(At.clone(), b.clone(), bem.PREC_FP64)
    e32 = np.linalg.norm(x32.cpu().numpy() - xr) / np.linalg.norm(xr)
    e16 = np.linalg.norm(x16.cpu().numpy() - xr) / np.linalg.norm(xr)
    e64 = np.linalg.norm(x64.cpu().numpy() - xr) / np.linalg.norm(xr)
    print("bem sphere8", m.n_dof, "mixed err", e32, rep, "split16 err", e16, rep16, "fp64 err", e64,
          flush=True)
    assert e32 <= 1e-8 and e64 <= 1e-10
    assert e16 <= 1e-8 and rep16["fell_back"] == 0


def test_solve_lse_c_dropin_and_reuse(bem):
    n = 500
    A = rand(n, True, 31)
    b = rand(n, True, 32)[:, :1]
    Af = np.asfortranarray(A.copy())
    bf = np.asfortranarray(b.copy())
    ipiv = np.zeros(n, np.int32)
    dummy = np.zeros(1, np.complex128)
    r = np.zeros(n); c = np.zeros(n)
    rep = bem.solve_lse_c(n, Af, ipiv, 1, dummy, "N", r, c, 1, bf, factorize=True)
    assert np.linalg.norm(bf[:, 0] - np.linalg.solve(A, b[:, 0])) <= 1e-11 * np.linalg.norm(bf)
    lu, piv = sla.lu_factor(A)
    assert np.array_equal(ipiv, piv + 1)
    assert np.max(np.abs(Af - lu)) <= 1e-10 * np.max(np.abs(lu))
    # factorize = .false. re-solve with a new right-hand side
    b2 = np.asfortranarray(rand(n, True, 33)[:, :1])
    x2 = b2.copy(order="F")
    bem.solve_lse_c(n, Af, ipiv, 1, dummy, "N", r, c, 1, x2, factorize=False)
    assert np.linalg.norm(x2[:, 0] - np.linalg.solve(A, b2[:, 0])) <= 1e-11 * np.linalg.norm(x2)
    # the condition-number stage is reported as unsupported, not silently skipped
    with pytest.raises(Exception):
        bem.solve_lse_c(n, Af, ipiv, 1, dummy, "N", r, c, 1, x2, factorize=True, condition=True)


@pytest.mark.parametrize("cplx", [False, True])
def test_solve_lse_reuses_the_kept_factor(bem, cplx):
    """factorize = 0 (src/multifebe.f90:119) solves with the factor kept on the device: only the
    triangular solves run (no factorisation time), whatever the host A now holds"""
    n = 700
    fn = bem.solve_lse_c if cplx else bem.solve_lse_r
    dt = np.complex128 if cplx else np.float64
    A = rand(n, cplx, 51)
    Af = np.asfortranarray(A.astype(dt))
    ipiv = np.zeros(n, np.int32)
    dummy = np.zeros(1, dt)
    r = np.zeros(n); c = np.zeros(n)
    b = np.asfortranarray(rand(n, cplx, 52)[:, :2].astype(dt))
    rep = fn(n, Af, ipiv, 1, dummy, "N", r, c, 2, b, factorize=True)
    lu, piv = sla.lu_factor(A)
    assert np.array_equal(ipiv, piv + 1)
    assert np.max(np.abs(Af - lu)) <= 1e-10 * np.max(np.abs(lu))
    assert rep["factor_ms"] > 0
    for seed in (53, 54):
        b2 = rand(n, cplx, seed)[:, :1]
        x2 = np.asfortranarray(b2.astype(dt))
        rep2 = fn(n, Af, ipiv, 1, dummy, "N", r, c, 1, x2, factorize=False)
        assert rep2["factor_ms"] == 0
        xr = sla.lu_solve((lu, piv), b2[:, 0])
        assert np.linalg.norm(x2[:, 0] - xr) <= 1e-11 * np.linalg.norm(xr)
    # another array (or another mode) has no kept factor: an error, not a wrong answer
    other = np.asfortranarray(A.astype(dt))
    with pytest.raises(Exception) as ei:
        fn(n, other, ipiv, 1, dummy, "N", r, c, 1, np.asfortranarray(b2.astype(dt)), factorize=False)
    assert "factorize = 0" in str(ei.value)
    with pytest.raises(Exception):
        fn(n, Af, ipiv, 1, dummy, "N", r, c, 1, np.asfortranarray(b2.astype(dt)), factorize=False,
           precision_mode=bem.PREC_FP32_REFINE)


@pytest.mark.parametrize("cplx", [False, True])
def test_solve_lse_mixed_mode_factors_and_reuse(bem, cplx):
    """mixed mode through the host drop-in: FP64-accurate solutions on factorize = 1 and 0, and A /
    ipiv returned as LAPACK-format factors of A (FP32 accuracy) that LAPACK's getrs can use"""
    n = 900
    fn = bem.solve_lse_c if cplx else bem.solve_lse_r
    dt = np.complex128 if cplx else np.float64
    A = rand(n, cplx, 61, diag=3 * np.sqrt(n))
    Af = np.asfortranarray(A.astype(dt))
    ipiv = np.zeros(n, np.int32)
    dummy = np.zeros(1, dt)
    r = np.zeros(n); c = np.zeros(n)
    b = rand(n, cplx, 62)[:, 0]
    x = np.asfortranarray(b.astype(dt).reshape(n, 1))
    rep = fn(n, Af, ipiv, 1, dummy, "N", r, c, 1, x, factorize=True,
             precision_mode=bem.PREC_FP32_REFINE)
    assert rep["precision_mode"] == bem.PREC_FP32_REFINE and rep["refinement_steps"] >= 1
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x[:, 0] - xr) <= 1e-12 * np.linalg.norm(xr)
    xl = sla.lu_solve((np.array(Af), ipiv - 1), b)
    assert np.linalg.norm(xl - xr) <= 1e-4 * np.linalg.norm(xr)
    b2 = rand(n, cplx, 63)[:, 0]
    x2 = np.asfortranarray(b2.astype(dt).reshape(n, 1))
    rep2 = fn(n, Af, ipiv, 1, dummy, "N", r, c, 1, x2, factorize=False,
              precision_mode=bem.PREC_FP32_REFINE)
    assert rep2["factor_ms"] == 0 and rep2["precision_mode"] == bem.PREC_FP32_REFINE
    xr2 = np.linalg.solve(A, b2)
    assert np.linalg.norm(x2[:, 0] - xr2) <= 1e-12 * np.linalg.norm(xr2)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("case", ["none", "rows", "cols", "both"])
def test_solve_lse_scaling_matches_lapack_geequ(bem, cplx, case):
    """lse_scaling (src/solve_lse_c.f90:81-117): r, c equal to LAPACK ?geequ's, the equilibration
    LAPACK ?laqge chooses, factors of the scaled matrix, and the unscaled solution"""
    import scipy.linalg.lapack as L
    n = 400
    rng = np.random.default_rng(71)
    A0 = rand(n, cplx, 72, diag=4.0)
    rs = np.ones(n) if case in ("none", "cols") else 10.0 ** rng.uniform(-6, 0, n)
    cs = 10.0 ** rng.uniform(-6, 0, n) if case == "both" else np.ones(n)
    if case == "cols":
        cs[rng.choice(n, 10, replace=False)] = 1e-3
    A = (rs[:, None] * A0) * cs[None, :]
    dt = np.complex128 if cplx else np.float64
    fn = bem.solve_lse_c if cplx else bem.solve_lse_r
    geequ = L.zgeequ if cplx else L.dgeequ
    r_ref, c_ref, rowcnd, colcnd, amax, info = geequ(A)
    assert info == 0
    Af = np.asfortranarray(A.astype(dt))
    ipiv = np.zeros(n, np.int32)
    dummy = np.zeros(1, dt)
    r = np.zeros(n); c = np.zeros(n)
    eq = bytearray(b"N")
    b = rand(n, cplx, 73)[:, 0]
    x = np.asfortranarray(b.astype(dt).reshape(n, 1))
    fn(n, Af, ipiv, 1, dummy, eq, r, c, 1, x, factorize=True, scaling=True)
    assert np.allclose(r, r_ref, rtol=1e-15, atol=0) and np.allclose(c, c_ref, rtol=1e-15, atol=0)
    expect = {"none": "N", "rows": "R", "cols": "C", "both": "B"}[case]
    assert eq.decode() == expect, (eq, rowcnd, colcnd)
    S = A.copy()
    if expect in ("R", "B"):
        S = r_ref[:, None] * S
    if expect in ("C", "B"):
        S = S * c_ref[None, :]
    lu, piv = sla.lu_factor(S)
    assert np.array_equal(ipiv, piv + 1)
    assert np.max(np.abs(Af - lu)) <= 1e-10 * np.max(np.abs(lu))
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x[:, 0] - xr) <= 1e-10 * np.linalg.norm(xr)
    # re-solve with the kept factor and the caller's r, c, equed (factorize = .false.)
    b2 = rand(n, cplx, 74)[:, 0]
    x2 = np.asfortranarray(b2.astype(dt).reshape(n, 1))
    fn(n, Af, ipiv, 1, dummy, eq, r, c, 1, x2, factorize=False, scaling=True)
    xr2 = np.linalg.solve(A, b2)
    assert np.linalg.norm(x2[:, 0] - xr2) <= 1e-10 * np.linalg.norm(xr2)


@pytest.mark.parametrize("cplx", [False, True])
def test_tall_panels_match_lapack(bem, cplx):
    """n = 4100: narrow panels of up to 17 row workgroups (the persistent panel's grid barrier
    across workgroups), pivots and factors against LAPACK"""
    import torch
    n = 4100
    A = rand(n, cplx, 41)
    b = rand(n, cplx, 42)[:, 0]
    At = dev(A)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    x, rep = bem.gesv_device(At, torch.from_numpy(b.copy()).cuda(), bem.PREC_FP64, ipiv=ipiv)
    lu, piv = sla.lu_factor(A)
    assert np.array_equal(ipiv.cpu().numpy(), piv + 1)
    LU = At.T.cpu().numpy()
    assert np.max(np.abs(LU - lu)) <= 1e-9 * np.max(np.abs(lu))
    xr = sla.lu_solve((lu, piv), b)
    assert np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr) <= 1e-10


def _lapack_match_in_child(env_extra, label):
    """FP64 LU of a complex n = 1100 matrix in a child process with env_extra set: pivots equal
    to LAPACK's, factors and the solution of one right-hand side within 1e-10"""
    import subprocess
    import sys
    import os
    code = (
        "import numpy as np, scipy.linalg as sla, torch\n"
        "from multifebe_amd import bem\n"
        "rng = np.random.default_rng(5)\n"
        "n = 1100\n"
        "A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))\n"
        "At = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()\n"
        "ipiv = torch.zeros(n, dtype=torch.int32, device='cuda')\n"
        "x, rep = bem.gesv_device(At, torch.ones(n, dtype=torch.complex128, device='cuda'), "
        "bem.PREC_FP64, ipiv=ipiv)\n"
        "lu, piv = sla.lu_factor(A)\n"
        "assert np.array_equal(ipiv.cpu().numpy(), piv + 1)\n"
        "assert np.max(np.abs(At.T.cpu().numpy() - lu)) <= 1e-10 * np.max(np.abs(lu))\n"
        "xr = sla.lu_solve((lu, piv), np.ones(n))\n"
        "assert np.linalg.norm(x.cpu().numpy() - xr) <= 1e-10 * np.linalg.norm(xr)\n"
        f"print('{label} ok', flush=True)\n")
    env = dict(os.environ, **env_extra)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and f"{label} ok" in r.stdout


def test_column_panel_mode_matches_lapack():
    """the one-launch-per-column panel (MFB_PANEL=column), no look-ahead"""
    _lapack_match_in_child({"MFB_PANEL": "column"}, "column mode")


def test_no_lookahead_matches_lapack():
    """persistent panels with the whole trailing update on one stream (MFB_LOOKAHEAD=0)"""
    _lapack_match_in_child({"MFB_LOOKAHEAD": "0"}, "no look-ahead")


def test_lookahead_small_chunks_matches_lapack():
    """look-ahead with the side-stream update split into 256-column GEMMs"""
    _lapack_match_in_child({"MFB_LA_CHUNK": "256"}, "look-ahead chunks")


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("scaling", [False, True])
def test_solve_lse_condition_and_refine_match_lapack_gesvx(bem, cplx, scaling):
    """the condition and refine stages of solve_lse (?gecon, ?gerfs on the FP64 factors) against
    LAPACK ?gesvx, which runs the same ?geequ/?laqge, ?getrf, ?gecon, ?getrs and ?gerfs steps;
    then a factorize = 0 re-solve with refinement"""
    import scipy.linalg.lapack as L
    n, nrhs = 700, 2
    rng = np.random.default_rng(17)
    A = rand(n, cplx, 41)
    A[:, :n // 3] *= 1e3                       # badly scaled columns (equilibration acts)
    B = rng.standard_normal((n, nrhs)) + (1j * rng.standard_normal((n, nrhs)) if cplx else 0)
    dt = np.complex128 if cplx else np.float64
    gesvx = L.zgesvx if cplx else L.dgesvx
    solve = bem.solve_lse_c if cplx else bem.solve_lse_r
    out = gesvx(A, B, fact=b"E" if scaling else b"N")
    As, equed_l, x_l, rcond_l, ferr_l, berr_l = out[0], out[3], out[7], out[8], out[9], out[10]
    Af = np.asfortranarray(A.astype(dt))
    Ao = np.zeros((n, n), dtype=dt, order="F")
    b = np.asfortranarray(B.astype(dt))
    ipiv = np.zeros(n, dtype=np.int32)
    r, c = np.zeros(n), np.zeros(n)
    eq = bytearray(b"N")
    rep = solve(n, Af, ipiv, n, Ao, eq, r, c, nrhs, b, factorize=True, scaling=scaling,
                condition=True, refine=True)
    print("condition/refine", cplx, scaling, rep, "lapack rcond", rcond_l, "ferr", ferr_l, "berr",
          berr_l, flush=True)
    assert bytes(eq) == bytes(equed_l)
    Ascaled = As if scaling else A
    assert np.max(np.abs(Ao - Ascaled)) <= 1e-15 * np.max(np.abs(Ascaled))
    assert abs(rep["rcond"] - rcond_l) <= 1e-6 * rcond_l
    assert np.max(np.abs(b - x_l)) <= 1e-10 * np.max(np.abs(x_l))
    assert rep["backward_error"] <= 4 * np.finfo(float).eps
    if not scaling:   # ?gesvx divides ferr by the column scaling's ratio; solve_lse does not
        assert 0.5 <= rep["ferr"] / np.max(ferr_l) <= 2.0
    # re-solve with the kept factor: refine needs the kept Ao; the condition stage runs only
    # with factorize = 1
    b2 = np.asfortranarray((B[:, :1] * 2.0 + 1.0).astype(dt))
    rep2 = solve(n, Af, ipiv, n, Ao, eq, r, c, 1, b2, factorize=False, scaling=scaling,
                 condition=True, refine=True)
    x2 = np.linalg.solve(A, (B[:, :1] * 2.0 + 1.0))
    assert rep2["factor_ms"] == 0 and rep2["rcond"] == 0
    assert np.max(np.abs(b2 - x2)) <= 1e-9 * np.max(np.abs(x2))


def test_split16_lookahead_chunks_refines():
    """split FP16 trailing updates with the side-stream update in 256-column GEMMs"""
    import os
    import subprocess
    import sys
    code = ("import numpy as np, torch\n"
            "from multifebe_amd import bem\n"
            "bem._torch()\n"
            "n = 1800\n"
            "rng = np.random.default_rng(3)\n"
            "A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))\n"
            "b = rng.standard_normal(n) + 0j\n"
            "At = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()\n"
            "x, rep = bem.gesv_device(At, torch.from_numpy(b.copy()).cuda(), "
            "bem.PREC_SPLIT16_REFINE)\n"
            "xr = np.linalg.solve(A, b)\n"
            "err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)\n"
            "print(err, rep, flush=True)\n"
            "assert rep['fell_back'] == 0 and err <= 1e-10\n"
            "print('split16 chunks ok', flush=True)\n")
    env = dict(os.environ, MFB_LA_CHUNK="256")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "split16 chunks ok" in r.stdout


def test_unsplit_triangular_solve_matches_lapack():
    """one workgroup per row block in the persistent triangular solves (MFB_TRSV_SPLIT=0)"""
    _lapack_match_in_child({"MFB_TRSV_SPLIT": "0"}, "unsplit solve")


@pytest.mark.parametrize("parts", ["2", "0"])
def test_fp32_triangular_solves_with_two_parts_and_unsplit(parts):
    """the FP32 / complex FP32 persistent triangular solves of the refinement with each row block
    split over 2 workgroups (MFB_TRSV_SPLIT=2; the default at the largest sizes, where 4 parts do
    not fit co-resident) and unsplit (=0): same refinement as the default 4 parts, short last
    block (n = 1800 = 7 x 256 + 8)"""
    import os
    import subprocess
    import sys
    code = ("import numpy as np, torch\n"
            "from multifebe_amd import bem\n"
            "bem._torch()\n"
            "n = 1800\n"
            "rng = np.random.default_rng(5)\n"
            "for cplx in (False, True):\n"
            "    A = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)\n"
            "    A += 3 * np.sqrt(n) * np.eye(n)\n"
            "    b = rng.standard_normal(n) + 0j if cplx else rng.standard_normal(n)\n"
            "    xr = np.linalg.solve(A, b)\n"
            "    for mode in (bem.PREC_FP32_REFINE, bem.PREC_SPLIT16_REFINE):\n"
            "        At = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()\n"
            "        x, rep = bem.gesv_device(At, torch.from_numpy(b.copy()).cuda(), mode)\n"
            "        err = np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr)\n"
            "        print(cplx, mode, err, rep, flush=True)\n"
            "        assert rep['fell_back'] == 0 and 1 <= rep['refinement_steps'] <= 4\n"
            "        assert err <= 1e-12\n"
            "print('trsv parts ok', flush=True)\n")
    env = dict(os.environ, MFB_TRSV_SPLIT=parts)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, "-c", code], env=env, cwd=root, capture_output=True,
                       text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and "trsv parts ok" in r.stdout
