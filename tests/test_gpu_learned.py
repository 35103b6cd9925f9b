"""Learned rule choice end to end: labels from the reference's adaptive rule, training on the GPU,
assembly with the network, the judge, and what the network gains over the estimator."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAU = 1e-6


@pytest.fixture(scope="module")
def trained():
    from multifebe_amd import learned as L, mesh as M
    samples = [L.Sample(M.sphere(6), harmonic_a0=(1.0, 3.0), far_pairs=1500),
               L.Sample(M.cube(6), harmonic_a0=(2.0,), far_pairs=1500),
               L.Sample(M.cylinder(4), harmonic_a0=(2.5,), far_pairs=1500)]
    ts = L.build_training_set(samples, TAU, tau_ref=1e-12, seed=1)
    net, rep = L.train(ts, epochs=40, batch=1024, lr=2e-3, seed=2)
    print("training", rep, flush=True)
    return net, rep


def test_training_report(trained):
    net, rep = trained
    assert rep["rows"] > 5000
    assert rep["loss_last"] < 0.6 * rep["loss_first"]
    assert rep["holdout_accuracy"] > 0.6
    # a confident choice is rarely cheaper than the label
    assert rep["under_resolved@0.99"] < 0.05


def test_learned_assembly_judge_and_gain(trained):
    import torch
    from multifebe_amd import bem, mesh as M
    net, _ = trained
    m = M.pressure_hemisphere_bcs(M.sphere(8))
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    omega = 2.0 * np.sqrt(mu_r / 2000.0)
    sysm = bem.BemSystem(m, bem.Settings(qsi_relative_error=TAU))
    At_p, b_p = sysm.assemble_harela(omega, lam_c, mu_c, 2000.0, 0.3)
    A_p, b_p = At_p.clone(), b_p.clone()
    st_p = sysm.stats()
    sysm.set_qnet(net, confidence=0.99, judge_qsi_relative_error=1e-12, judge_samples=512)
    At_l, b_l = sysm.assemble_harela(omega, lam_c, mu_c, 2000.0, 0.3)
    st_l = sysm.stats()
    print("plain", st_p, "\nlearned", st_l, flush=True)
    assert st_l["learned"] == 1 and st_l["judge_pairs"] >= 500
    # the network meets the tolerance on the judge's sample (a few-fold margin for the rare miss)
    assert st_l["judge_max_err"] < 5 * TAU
    # and does at least as well there as the reference's estimator on the same pairs
    assert st_l["judge_max_err"] <= max(st_l["judge_ref_max_err"], TAU)
    # the two matrices agree to the integration tolerance
    d = (At_l - A_p).abs().max().item() / A_p.abs().max().item()
    print("matrix difference", d, flush=True)
    assert d < 1e-4
    x_p = torch.linalg.solve(A_p.T, b_p)
    x_l = torch.linalg.solve(At_l.T, b_l)
    assert ((x_l - x_p).norm() / x_p.norm()).item() < 1e-4
    # back to the estimator: the same matrix as before (to the rounding of the atomic additions)
    sysm.set_qnet(None)
    At_q, _ = sysm.assemble_harela(omega, lam_c, mu_c, 2000.0, 0.3)
    assert (At_q - A_p).abs().max().item() <= 1e-14 * A_p.abs().max().item()
    sysm.close()


def test_judge_action_promotes_short_rules(trained):
    """with the judge's action on, rules whose judged pairs miss the tolerance are promoted and
    the assembly is run again: the judge's error after the promotions is below its error before,
    the matrix moves towards a converged quadrature, and the promotions persist"""
    from multifebe_amd import bem, mesh as M
    net, _ = trained
    tau = 1e-10   # tighter than the training tolerance: the network's rules miss it on some pairs
    m = M.pressure_hemisphere_bcs(M.sphere(8))
    mu_r, mu_c, lam_c = M.elastic_constants(1e9, 0.3, 0.05)
    omega = 3.0 * np.sqrt(mu_r / 2000.0)
    args = (omega, lam_c, mu_c, 2000.0, 0.3)
    conv = bem.BemSystem(m, bem.Settings(qsi_relative_error=1e-12, precalsets=(12,)))
    Ac = conv.assemble_harela(*args)[0].T.cpu().numpy()
    sysm = bem.BemSystem(m, bem.Settings(qsi_relative_error=tau))
    sysm.set_qnet(net, confidence=0.99, judge_qsi_relative_error=1e-12, judge_samples=512)
    A0 = sysm.assemble_harela(*args)[0].T.cpu().numpy()
    st0 = sysm.stats()
    sysm.set_judge_action(True)
    A1 = sysm.assemble_harela(*args)[0].T.cpu().numpy()
    st1 = sysm.stats()
    A2 = sysm.assemble_harela(*args)[0].T.cpu().numpy()
    st2 = sysm.stats()
    rowmax = np.abs(Ac).max(axis=1, keepdims=True)
    e0 = np.max(np.abs(A0 - Ac) / rowmax)
    e1 = np.max(np.abs(A1 - Ac) / rowmax)
    print("before", st0, "\nafter", st1, "\nnext", st2, "\nerr vs converged", e0, e1, flush=True)
    assert st0["judge_max_err"] > tau and st0["judge_promoted"] == 0
    assert st1["judge_promoted"] >= 1
    assert abs(st1["judge_first_max_err"] - st0["judge_max_err"]) <= 1e-3 * st0["judge_max_err"]
    assert st1["judge_max_err"] < st1["judge_first_max_err"]
    assert e1 <= e0 * (1 + 1e-6)
    assert st1["points"] > st0["points"]
    # the promotions persist: the next assembly starts from them
    assert st2["judge_first_max_err"] <= st1["judge_max_err"] * (1 + 1e-9)
    assert np.max(np.abs(A2 - A1) / rowmax) <= 1e-12 or st2["judge_promoted"] >= 1
    sysm.close()
    conv.close()
