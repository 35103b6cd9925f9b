"""mfb_qnet_gradient_noalloc (fixed device scratch, no allocation) against mfb_qnet_gradient
through the C ABI: the same kernels and summation order, so the gradient and the loss are
bit-identical; two calls in a row, the second smaller, and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grad(fn, net, X, y):
    from multifebe_amd import _lib
    g = np.zeros(_lib.lib().mfb_qnet_param_count(), dtype=np.float32)
    loss = C.c_float()
    rc = fn(net.handle, C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()), int(X.shape[0]),
            C.c_void_p(g.ctypes.data), C.byref(loss))
    return rc, g, loss.value


def test_noalloc_gradient_matches_allocating_gradient():
    import torch
    from multifebe_amd import _lib
    from multifebe_amd.qnet import QNet
    lib = _lib.lib()
    net = QNet(seed=7)
    rng = np.random.default_rng(11)
    for n in (9000, 777):   # second call smaller: the scratch keeps the first call's partials
        X = torch.from_numpy(rng.standard_normal((n, 5)).astype(np.float32)).cuda()
        y = torch.from_numpy(rng.integers(0, 9, n).astype(np.int32)).cuda()
        rc0, g0, l0 = _grad(lib.mfb_qnet_gradient, net, X, y)
        rc1, g1, l1 = _grad(lib.mfb_qnet_gradient_noalloc, net, X, y)
        print(n, "loss", l0, l1, "max |dg|", float(np.max(np.abs(g0 - g1))), flush=True)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(g0, g1) and l0 == l1 and np.abs(g0).max() > 0
    X = torch.zeros((16385, 5), dtype=torch.float32, device="cuda")
    y = torch.zeros(16385, dtype=torch.int32, device="cuda")
    rc, _, _ = _grad(lib.mfb_qnet_gradient_noalloc, net, X, y)
    assert rc == -6 and "16384" in lib.mfb_last_error().decode()
    net.close()
