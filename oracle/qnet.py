"""CPU reference of the learned quadrature-rule network (float64 numpy).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the GPU network (multifebe_amd/csrc/mfbem_qnet.hip)
-- forward probabilities, the class choice and the gradient of the loss.  It is pinned by central
finite differences of its own loss (tests/test_qnet_oracle.py).

Network: N_IN inputs -> 3 tanh layers of 64 -> N_OUT logits, softmax cross-entropy (mean over
rows).  Parameters: one flat array; layer l stores W_l transposed (Wt[i][o], o fastest, shape
n_in x n_out) followed by b_l (n_out), for l = 0..3.
"""
import numpy as np

N_IN, N_HID, N_OUT = 5, 64, 9
LAYERS = [(N_IN, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_OUT)]


def n_params():
    return sum(i * o + o for i, o in LAYERS)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, k = [], 0
    for i, o in LAYERS:
        Wt = p[k:k + i * o].reshape(i, o)
        k += i * o
        b = p[k:k + o]
        k += o
        out.append((Wt, b))
    return out


def forward(p, X):
    """probabilities (n, N_OUT) and the activations needed by backward"""
    layers = unpack(p)
    a = [np.asarray(X, dtype=np.float64)]
    for l, (Wt, b) in enumerate(layers):
        z = a[-1] @ Wt + b
        a.append(np.tanh(z) if l < 3 else z)
    z = a[-1]
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True), a


def loss(p, X, y):
    P, _ = forward(p, X)
    return float(-np.mean(np.log(np.maximum(P[np.arange(len(y)), y], 1e-300))))


def gradient(p, X, y):
    """d loss / d p (same flat layout)"""
    P, a = forward(p, X)
    n = len(y)
    d = P.copy()
    d[np.arange(n), y] -= 1.0
    d /= n
    layers = unpack(p)
    grads = [None] * 4
    for l in range(3, -1, -1):
        Wt, b = layers[l]
        grads[l] = (a[l].T @ d, d.sum(axis=0))
        if l > 0:
            d = (d @ Wt.T) * (1.0 - a[l] ** 2)
    return np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])


def choose(P, confidence):
    """cheapest class whose cumulative probability reaches confidence (last class otherwise)"""
    c = np.cumsum(P, axis=1)
    hit = c >= confidence
    return np.where(hit.any(axis=1), hit.argmax(axis=1), P.shape[1] - 1)
