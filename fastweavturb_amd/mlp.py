"""ctypes interface to the surrogate network of libwovengpu.so (wt_mlp_* in include/wovengpu.h).

A multilayer perceptron with GELU hidden layers and a linear output, trained and evaluated on
the GPU by the library's BF16 MFMA kernels. There is no CPU fallback: without the built library
`MLP(...)` raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .api import WovenError, load_library

_bound = False


def _lib():
    global _bound
    L = load_library()
    if not _bound:
        vp, i32, i64, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64
        fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.wt_mlp_create.argtypes = [C.POINTER(vp), i32, ip, u64, i32]
        L.wt_mlp_destroy.argtypes = [vp]
        L.wt_mlp_destroy.restype = None
        L.wt_mlp_last_error.argtypes = [vp]
        L.wt_mlp_last_error.restype = C.c_char_p
        L.wt_mlp_num_params.argtypes = [vp]
        L.wt_mlp_num_params.restype = i64
        L.wt_mlp_get_params.argtypes = [vp, fp]
        L.wt_mlp_set_params.argtypes = [vp, fp]
        L.wt_mlp_forward.argtypes = [vp, i32, fp, fp]
        L.wt_mlp_loss_grad.argtypes = [vp, i32, fp, fp, fp, dp, fp]
        L.wt_mlp_train_step.argtypes = [vp, i32, fp, fp, fp, C.c_float, dp]
        L.wt_mlp_set_data.argtypes = [vp, i32, fp, fp, fp]
        L.wt_mlp_train.argtypes = [vp, i32, i32, C.c_float, C.c_float, u64, dp]
        _bound = True
    return L


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """sizes = [inputs, hidden..., outputs]; weights initialised from `seed` on the host."""

    def __init__(self, sizes, seed: int = 0, device: int = 0):
        self._L = _lib()
        self.sizes = [int(s) for s in sizes]
        arr = np.ascontiguousarray(self.sizes, dtype=np.int32)
        h = C.c_void_p()
        rc = self._L.wt_mlp_create(C.byref(h), len(arr) - 1, arr.ctypes.data_as(C.POINTER(C.c_int32)),
                                   int(seed), device)
        if rc != 0:
            raise WovenError(rc, self._L.wt_mlp_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc < 0:
            raise WovenError(rc, self._L.wt_mlp_last_error(self._h).decode())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._L.wt_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self) -> int:
        return int(self._L.wt_mlp_num_params(self._h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, dtype=np.float32)
        self._check(self._L.wt_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, flat):
        a = _f32(flat)
        if a.size != self.num_params:
            raise ValueError(f"need {self.num_params} parameters, got {a.size}")
        self._check(self._L.wt_mlp_set_params(self._h, _fp(a)))

    def forward(self, X) -> np.ndarray:
        X = _f32(X)
        Y = np.zeros((X.shape[0], self.sizes[-1]), dtype=np.float32)
        self._check(self._L.wt_mlp_forward(self._h, X.shape[0], _fp(X), _fp(Y)))
        return Y

    def loss_grad(self, X, T, mask=None):
        X, T, M = _f32(X), _f32(T), _f32(mask)
        loss = C.c_double()
        g = np.zeros(self.num_params, dtype=np.float32)
        self._check(self._L.wt_mlp_loss_grad(self._h, X.shape[0], _fp(X), _fp(T), _fp(M), C.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, X, T, mask=None, lr: float = 1e-3) -> float:
        X, T, M = _f32(X), _f32(T), _f32(mask)
        loss = C.c_double()
        self._check(self._L.wt_mlp_train_step(self._h, X.shape[0], _fp(X), _fp(T), _fp(M), float(lr), C.byref(loss)))
        return loss.value

    def set_data(self, X, T, mask=None):
        X, T, M = _f32(X), _f32(T), _f32(mask)
        self._check(self._L.wt_mlp_set_data(self._h, X.shape[0], _fp(X), _fp(T), _fp(M)))

    def train(self, epochs: int, batch: int = 256, lr: float = 2e-3, lr_final: float = 2e-5,
              seed: int = 0) -> np.ndarray:
        hist = np.zeros(epochs)
        self._check(self._L.wt_mlp_train(self._h, epochs, batch, float(lr), float(lr_final), int(seed),
                                         hist.ctypes.data_as(C.POINTER(C.c_double))))
        return hist
