"""Learned quadrature-rule selection network on the MI355X library (host-side interface).

The network replaces the reference's fitted Gauss-point estimator
(fbem_qs_n_estimation_standard, lib/fbem/src/quasisingular_integration.f90:294-715) in the rule
choice of sbie_auto (lib/fbem/src/bem_staela3d.f90:549-587): from the features of a (collocation
point, element) pair it predicts the class of the cheapest rule that meets the integration
tolerance.  Classes 0..7: precalculated Gauss-Legendre rule of order 2..9; class 8: the adaptive
rule.  Training, gradients and inference run in the library's HIP kernels; tensors here are torch
CUDA tensors (device memory only).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

N_INPUTS = 5
N_CLASSES = 9
ADAPTIVE_CLASS = 8


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the quadrature network needs a visible GPU")
    return torch


def _dev(t, dtype):
    torch = _torch()
    t = torch.as_tensor(t, dtype=dtype, device="cuda").contiguous()
    return t


class QNet:
    """FP32 MLP: 5 features -> 3 x 64 tanh -> 9 logits (weights live in device memory)."""

    def __init__(self, seed: int = 0):
        _torch()  # initialise the device through torch first, so that both share the context
        self._h = C.c_void_p()
        check(lib().mfb_qnet_create(C.c_ulonglong(seed), C.byref(self._h)), "mfb_qnet_create")
        _lib.track(self)

    def close(self):
        if self._h:
            lib().mfb_qnet_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @staticmethod
    def param_count() -> int:
        return lib().mfb_qnet_param_count()

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        check(lib().mfb_qnet_get_params(self._h, out.ctypes.data_as(C.c_void_p)), "get_params")
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count()
        check(lib().mfb_qnet_set_params(self._h, p.ctypes.data_as(C.c_void_p)), "set_params")

    def gradient(self, X, y):
        """(gradient of the mean cross-entropy, loss) over all rows"""
        torch = _torch()
        X = _dev(X, torch.float32)
        y = _dev(y, torch.int32)
        g = np.zeros(self.param_count(), np.float32)
        loss = C.c_float()
        check(lib().mfb_qnet_gradient(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()),
                                      X.shape[0], g.ctypes.data_as(C.c_void_p), C.byref(loss)),
              "mfb_qnet_gradient")
        return g, loss.value

    def train(self, X, y, epochs=20, batch=4096, lr=2e-3, seed=0):
        """Adam over shuffled mini-batches; returns the mean mini-batch loss of every epoch"""
        torch = _torch()
        X = _dev(X, torch.float32)
        y = _dev(y, torch.int32)
        losses = np.zeros(epochs, np.float32)
        check(lib().mfb_qnet_train(self._h, C.c_void_p(X.data_ptr()), C.c_void_p(y.data_ptr()),
                                   X.shape[0], epochs, batch, C.c_float(lr), C.c_ulonglong(seed),
                                   losses.ctypes.data_as(C.c_void_p)), "mfb_qnet_train")
        return losses

    def infer(self, X, confidence=0.99):
        """(probabilities n x 9, chosen class per row) as torch CUDA tensors"""
        torch = _torch()
        X = _dev(X, torch.float32)
        n = X.shape[0]
        probs = torch.empty((n, N_CLASSES), dtype=torch.float32, device="cuda")
        cls = torch.empty(n, dtype=torch.int32, device="cuda")
        check(lib().mfb_qnet_infer(self._h, C.c_void_p(X.data_ptr()), n, C.c_float(confidence),
                                   C.c_void_p(probs.data_ptr()), C.c_void_p(cls.data_ptr()),
                                   C.c_void_p(torch.cuda.current_stream().cuda_stream)),
              "mfb_qnet_infer")
        return probs, cls
