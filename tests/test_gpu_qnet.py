"""GPU quadrature network (mfbem_qnet.hip) against its CPU reference (oracle/qnet.py)."""
import numpy as np
import pytest

from oracle import qnet as Q

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def QN():
    from multifebe_amd import qnet
    return qnet


def test_gradient_matches_oracle(QN):
    rng = np.random.default_rng(5)
    net = QN.QNet(seed=11)
    p = net.params.astype(np.float64)
    n = 300  # several workgroups of 64 and a padded last one
    X = rng.normal(size=(n, Q.N_IN)).astype(np.float32)
    y = rng.integers(0, Q.N_OUT, size=n).astype(np.int32)
    g, loss = net.gradient(X, y)
    g_ref = Q.gradient(p, X.astype(np.float64), y)
    l_ref = Q.loss(p, X.astype(np.float64), y)
    print("grad rel err", np.abs(g - g_ref).max() / np.abs(g_ref).max(), "loss", loss, l_ref)
    assert abs(loss - l_ref) <= 1e-5 * abs(l_ref)
    assert np.abs(g - g_ref).max() <= 2e-5 * np.abs(g_ref).max()
    net.close()


def test_inference_matches_oracle(QN):
    rng = np.random.default_rng(6)
    net = QN.QNet(seed=12)
    p = net.params.astype(np.float64)
    X = rng.normal(size=(1000, Q.N_IN)).astype(np.float32)
    probs, cls = net.infer(X, confidence=0.8)
    P_ref, _ = Q.forward(p, X.astype(np.float64))
    probs = probs.cpu().numpy()
    assert np.abs(probs - P_ref).max() <= 1e-5
    c_ref = Q.choose(P_ref, 0.8)
    cum = np.cumsum(P_ref, axis=1)
    differ = np.nonzero(cls.cpu().numpy() != c_ref)[0]
    # a different choice only where the cumulative probability is within rounding of the level
    for r in differ:
        assert np.min(np.abs(cum[r] - 0.8)) < 1e-5
    net.close()


def synthetic(n, rng):
    X = rng.uniform(-1, 1, size=(n, Q.N_IN)).astype(np.float32)
    y = np.clip(np.floor(4.5 - 4.0 * X[:, 0] + 0.8 * X[:, 1] * X[:, 2]), 0, Q.N_OUT - 1)
    return X, y.astype(np.int32)


def test_training_learns_and_is_repeatable(QN):
    rng = np.random.default_rng(7)
    X, y = synthetic(40000, rng)
    nets = [QN.QNet(seed=3), QN.QNet(seed=3)]
    losses = [net.train(X, y, epochs=12, batch=1024, lr=3e-3, seed=9) for net in nets]
    print("losses", losses[0])
    assert losses[0][-1] < 0.25 * losses[0][0]
    assert np.array_equal(nets[0].params, nets[1].params)
    Xt, yt = synthetic(5000, np.random.default_rng(8))
    probs, _ = nets[0].infer(Xt, confidence=0.5)
    acc = (probs.cpu().numpy().argmax(axis=1) == yt).mean()
    print("held-out accuracy", acc)
    assert acc > 0.9
    for net in nets:
        net.close()
