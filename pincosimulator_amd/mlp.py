"""The warm start's network on the GPU (twmc_mlp_* of libtwmc_hip.so): a
fully connected MLP with exact-erf GELU hidden layers, FP32, trained by
minibatch Adam and evaluated by the library's own HIP kernels."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, cols):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a.reshape(-1, cols)


class MLP:
    def __init__(self, n_in, n_out, n_hidden=256, n_hidden_layers=3, seed=0, device=0):
        self._L = _lib.load()
        self.widths = [n_in] + [n_hidden] * n_hidden_layers + [n_out]
        cfg = _lib.MlpConfig(n_in=n_in, n_hidden=n_hidden, n_hidden_layers=n_hidden_layers, n_out=n_out,
                             seed=seed)
        ctx = C.c_void_p()
        check(self._L.twmc_mlp_create(C.byref(cfg), device, C.byref(ctx)), None, mlp=True)
        self._ctx = ctx
        self.n_params = int(self._L.twmc_mlp_n_params(ctx))

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.twmc_mlp_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        check(self._L.twmc_mlp_get_params(self._ctx, _fp(out)), self._ctx, mlp=True)
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        check(self._L.twmc_mlp_set_params(self._ctx, _fp(p)), self._ctx, mlp=True)

    def forward(self, x):
        x = _f32(x, self.widths[0])
        y = np.zeros((x.shape[0], self.widths[-1]), dtype=np.float32)
        check(self._L.twmc_mlp_forward(self._ctx, _fp(x), x.shape[0], _fp(y)), self._ctx, mlp=True)
        return y

    def loss_grad(self, x, t, w=None):
        x = _f32(x, self.widths[0])
        t = _f32(t, self.widths[-1])
        wp = C.POINTER(C.c_float)()
        if w is not None:
            w = _f32(w, self.widths[-1])
            wp = _fp(w)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_double(0.0)
        check(self._L.twmc_mlp_loss_grad(self._ctx, _fp(x), _fp(t), wp, x.shape[0], C.byref(loss), _fp(g)),
              self._ctx, mlp=True)
        return loss.value, g

    def train(self, x, t, w=None, epochs=100, batch=256, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8,
              shuffle_seed=1):
        x = _f32(x, self.widths[0])
        t = _f32(t, self.widths[-1])
        wp = C.POINTER(C.c_float)()
        if w is not None:
            w = _f32(w, self.widths[-1])
            wp = _fp(w)
        hist = np.zeros(max(int(epochs), 1))
        o = _lib.MlpTrainOpts(epochs=int(epochs), batch=int(batch), lr=lr, beta1=beta1, beta2=beta2, eps=eps,
                              shuffle_seed=shuffle_seed)
        check(self._L.twmc_mlp_train(self._ctx, _fp(x), _fp(t), wp, x.shape[0], C.byref(o),
                                     hist.ctypes.data_as(C.POINTER(C.c_double))), self._ctx, mlp=True)
        return hist[:int(epochs)]
