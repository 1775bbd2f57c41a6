"""The calibration emulator of libcovidsim_hip (cs_mlp_* of include/covidsim_hip.h).

A multilayer perceptron n_in -> n_hidden -> n_hidden -> n_out trained and
evaluated on the GPU with the library's BF16 MFMA kernels.  Every method is
one C ABI call; errors raise CSError.  No fallback: the library and a GPU
are required.
"""
import ctypes as C

import numpy as np

from .engine import CSError, lib

_declared = False


def _lib():
    global _declared
    L = lib()
    if not _declared:
        P, vp = C.POINTER, C.c_void_p
        F, D, I = C.c_float, C.c_double, C.c_int
        L.cs_mlp_create.argtypes = [P(vp), I, I, I, I, C.c_uint64]
        L.cs_mlp_create.restype = I
        L.cs_mlp_destroy.argtypes = [vp]
        L.cs_mlp_destroy.restype = None
        L.cs_mlp_last_error.argtypes = [vp]
        L.cs_mlp_last_error.restype = C.c_char_p
        L.cs_mlp_n_params.argtypes = [vp]
        L.cs_mlp_n_params.restype = C.c_int64
        L.cs_mlp_get_params.argtypes = [vp, P(F)]
        L.cs_mlp_set_params.argtypes = [vp, P(F)]
        L.cs_mlp_forward.argtypes = [vp, P(F), I, P(F)]
        L.cs_mlp_forward_noalloc.argtypes = [vp, vp, I, vp]
        L.cs_mlp_forward_noalloc.restype = I
        L.cs_mlp_forward_trace.argtypes = [vp, P(F), I, P(F), P(F), P(C.c_uint16), P(C.c_uint16), P(F)]
        L.cs_mlp_loss_grad.argtypes = [vp, P(F), P(F), I, P(D), P(F)]
        L.cs_mlp_set_data.argtypes = [vp, P(F), P(F), I]
        L.cs_mlp_train.argtypes = [vp, I, I, F, C.c_uint64, P(D)]
        for n in ("get_params", "set_params", "forward", "forward_trace", "loss_grad", "set_data", "train"):
            getattr(L, "cs_mlp_" + n).restype = I
        _declared = True
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Emulator:
    def __init__(self, n_in, n_hidden, n_out, seed=1, device=0):
        self.L = _lib()
        self.h = C.c_void_p()
        self.n_in, self.n_hidden, self.n_out = n_in, n_hidden, n_out
        rc = self.L.cs_mlp_create(C.byref(self.h), int(device), int(n_in), int(n_hidden), int(n_out), int(seed))
        if rc:
            msg = self.L.cs_mlp_last_error(self.h).decode() if self.h else "create failed"
            self.close()
            raise CSError(rc, msg)

    def _check(self, rc):
        if rc:
            raise CSError(rc, self.L.cs_mlp_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.L.cs_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self):
        return int(self.L.cs_mlp_n_params(self.h))

    def params(self):
        out = np.zeros(self.n_params, np.float32)
        self._check(self.L.cs_mlp_get_params(self.h, _p(out, C.c_float)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._check(self.L.cs_mlp_set_params(self.h, _p(p, C.c_float)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        y = np.zeros((len(x), self.n_out), np.float32)
        self._check(self.L.cs_mlp_forward(self.h, _p(x, C.c_float), len(x), _p(y, C.c_float)))
        return y

    def forward_device(self, x_dev, batch, y_dev):
        """The allocation-free forward: x_dev and y_dev are device addresses
        ([batch][n_in] and [batch][n_out] FP32) the caller owns."""
        self._check(self.L.cs_mlp_forward_noalloc(self.h, C.c_void_p(int(x_dev)), int(batch), C.c_void_p(int(y_dev))))

    def forward_trace(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        b, H = len(x), self.n_hidden
        z1, z2 = np.zeros((b, H), np.float32), np.zeros((b, H), np.float32)
        h1, h2 = np.zeros((b, H), np.uint16), np.zeros((b, H), np.uint16)
        y = np.zeros((b, self.n_out), np.float32)
        self._check(self.L.cs_mlp_forward_trace(self.h, _p(x, C.c_float), b, _p(z1, C.c_float), _p(z2, C.c_float),
                                                _p(h1, C.c_uint16), _p(h2, C.c_uint16), _p(y, C.c_float)))
        return z1, h1, z2, h2, y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, self.n_out)
        loss = C.c_double()
        g = np.zeros(self.n_params, np.float32)
        self._check(self.L.cs_mlp_loss_grad(self.h, _p(x, C.c_float), _p(t, C.c_float), len(x), C.byref(loss), _p(g, C.c_float)))
        return loss.value, g

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, self.n_in)
        t = np.ascontiguousarray(t, np.float32).reshape(-1, self.n_out)
        self._check(self.L.cs_mlp_set_data(self.h, _p(x, C.c_float), _p(t, C.c_float), len(x)))

    def train(self, epochs, batch, lr, seed=7):
        losses = np.zeros(max(epochs, 1))
        self._check(self.L.cs_mlp_train(self.h, int(epochs), int(batch), float(lr), int(seed), _p(losses, C.c_double)))
        return losses[:epochs]
