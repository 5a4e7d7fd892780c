"""ctypes layer over the learned-screener network (nw_mlp_* in include/nwick.h)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import NwickError, _ptr, load, sync_torch

N_HIDDEN = 128


def _setup(L):
    if getattr(L, "_mlp_ready", False):
        return
    vp = ct.c_void_p
    L.nw_mlp_create.argtypes = [ct.POINTER(vp), ct.c_int, ct.c_int, ct.c_uint64]
    L.nw_mlp_num_params.argtypes = [vp, ct.POINTER(ct.c_longlong)]
    L.nw_mlp_set_params.argtypes = [vp, vp, ct.c_longlong]
    L.nw_mlp_get_params.argtypes = [vp, vp, ct.c_longlong]
    L.nw_mlp_forward.argtypes = [vp, vp, ct.c_longlong, vp]
    L.nw_mlp_grad.argtypes = [vp, vp, vp, vp, ct.c_longlong, vp, vp]
    L.nw_mlp_step.argtypes = [vp, vp, vp, vp, ct.c_longlong, ct.c_float, vp]
    L.nw_mlp_train.argtypes = [vp, vp, vp, vp, ct.c_longlong, ct.c_int, ct.c_int, ct.c_float, ct.c_uint64, vp]
    L.nw_mlp_last_error.argtypes = [vp]
    L.nw_mlp_last_error.restype = ct.c_char_p
    L.nw_mlp_destroy.argtypes = [vp]
    L.nw_mlp_destroy.restype = None
    L._mlp_ready = True


def _f32(a):
    if a is None or hasattr(a, "data_ptr"):
        return a
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """n_in-128-128-1 ReLU network with BF16 weights, trained on the GPU."""

    def __init__(self, n_in=24, seed=0, device=0):
        self.L = load()
        _setup(self.L)
        h = ct.c_void_p()
        rc = self.L.nw_mlp_create(ct.byref(h), device, n_in, seed)
        if rc != 0:
            raise NwickError(rc, "nw_mlp_create failed")
        self.h, self.n_in = h, n_in
        n = ct.c_longlong()
        self._check(self.L.nw_mlp_num_params(self.h, ct.byref(n)))
        self.n_params = n.value

    def _check(self, rc):
        if rc != 0:
            raise NwickError(rc, self.L.nw_mlp_last_error(self.h).decode())

    @property
    def params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        self._check(self.L.nw_mlp_get_params(self.h, _ptr(p), self.n_params))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(self.L.nw_mlp_set_params(self.h, _ptr(p), len(p)))

    def forward(self, X, out=None):
        X = _f32(X)
        n = X.shape[0]
        y = np.empty(n, dtype=np.float32) if out is None else out
        sync_torch(X, y)
        self._check(self.L.nw_mlp_forward(self.h, _ptr(X), n, _ptr(y)))
        return y

    def grad(self, X, t, w=None):
        X, t, w = _f32(X), _f32(t), _f32(w)
        loss = np.zeros(1, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        sync_torch(X, t, w)
        self._check(self.L.nw_mlp_grad(self.h, _ptr(X), _ptr(t), _ptr(w), X.shape[0], _ptr(loss), _ptr(g)))
        return float(loss[0]), g

    def step(self, X, t, w=None, lr=1e-3):
        X, t, w = _f32(X), _f32(t), _f32(w)
        loss = np.zeros(1, dtype=np.float32)
        sync_torch(X, t, w)
        self._check(self.L.nw_mlp_step(self.h, _ptr(X), _ptr(t), _ptr(w), X.shape[0], float(lr), _ptr(loss)))
        return float(loss[0])

    def train(self, X, t, w=None, epochs=5, batch=4096, lr=1e-3, seed=0):
        X, t, w = _f32(X), _f32(t), _f32(w)
        el = np.zeros(epochs, dtype=np.float32)
        sync_torch(X, t, w)
        self._check(self.L.nw_mlp_train(self.h, _ptr(X), _ptr(t), _ptr(w), X.shape[0], int(epochs), int(batch),
                                        float(lr), int(seed), _ptr(el)))
        return el

    def close(self):
        if getattr(self, "h", None):
            self.L.nw_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
