"""The CPU reference of the quadrature network (oracle/qnet.py), pinned by finite differences."""
import numpy as np

from oracle import qnet as Q


def test_parameter_count_matches_layout():
    assert Q.n_params() == 5 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 9 + 9


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = rng.normal(scale=0.3, size=Q.n_params())
    X = rng.normal(size=(7, Q.N_IN))
    y = rng.integers(0, Q.N_OUT, size=7)
    g = Q.gradient(p, X, y)
    idx = np.concatenate([rng.choice(Q.n_params(), 40, replace=False),
                          [0, 5 * 64 + 3, Q.n_params() - 1, Q.n_params() - 10]])
    h = 1e-6
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (Q.loss(p + e, X, y) - Q.loss(p - e, X, y)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-6 * abs(fd), (k, fd, g[k])


def test_probabilities_and_choice():
    rng = np.random.default_rng(4)
    p = rng.normal(scale=0.3, size=Q.n_params())
    X = rng.normal(size=(50, Q.N_IN))
    P, _ = Q.forward(p, X)
    assert np.allclose(P.sum(axis=1), 1.0) and (P >= 0).all()
    c = Q.choose(P, 0.9)
    cum = np.cumsum(P, axis=1)
    for r in range(50):
        assert cum[r, c[r]] >= 0.9 or c[r] == Q.N_OUT - 1
        assert c[r] == 0 or cum[r, c[r] - 1] < 0.9
    # confidence 1 with rounding below 1 falls through to the last class only when needed
    assert (Q.choose(np.eye(Q.N_OUT)[[2, 5]], 1.0) == [2, 5]).all()
