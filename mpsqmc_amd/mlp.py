"""The learned proposal's network (``mqe_mlp_*`` of the C ABI).

An MLP n_in -> hidden -> hidden -> n_out with ReLU in FP32 whose forward,
backward and Adam kernels run on the GPU (``csrc/mqe_mlp.hip``).  For the
learned proposal it maps MQE_NFEAT = 23 per-walker, per-coupling features to the
MQE_NLABEL = 16 normalised two-site elements combos/O; see
``Engine.collect_samples``, ``Engine.train_proposal`` and
``Engine.set_proposal``.  The reference has no learned component.
"""
import ctypes

import numpy as np

from ._lib import check, lib

NFEAT = 23
NLABEL = 16

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


class MLP:
    def __init__(self, n_in=NFEAT, hidden=128, n_out=NLABEL, seed=1):
        self.n_in, self.hidden, self.n_out = int(n_in), int(hidden), int(n_out)
        h = ctypes.c_void_p()
        check(lib().mqe_mlp_create(self.n_in, self.hidden, self.n_out, int(seed), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().mqe_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return int(lib().mqe_mlp_num_params(self.h))

    def params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(lib().mqe_mlp_get_params(self.h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p):
        p = _f32(p)
        if p.size != self.num_params:
            raise ValueError("wrong parameter count")
        check(lib().mqe_mlp_set_params(self.h, p.ctypes.data_as(_fp)))

    def scaling(self):
        a = [np.zeros(self.n_in, np.float32), np.zeros(self.n_in, np.float32), np.zeros(self.n_out, np.float32),
             np.zeros(self.n_out, np.float32)]
        check(lib().mqe_mlp_get_scaling(self.h, *[x.ctypes.data_as(_fp) for x in a]))
        return tuple(a)

    def set_scaling(self, x_mean, x_std, y_mean, y_std):
        a = [_f32(x_mean), _f32(x_std), _f32(y_mean), _f32(y_std)]
        check(lib().mqe_mlp_set_scaling(self.h, *[x.ctypes.data_as(_fp) for x in a]))

    def forward(self, X):
        X = _f32(X, (-1, self.n_in))
        Y = np.zeros((X.shape[0], self.n_out), np.float32)
        check(lib().mqe_mlp_forward(self.h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp)))
        return Y

    def loss_grad(self, X, Y):
        X = _f32(X, (-1, self.n_in))
        Y = _f32(Y, (-1, self.n_out))
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_double()
        check(lib().mqe_mlp_loss_grad(self.h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp),
                                      ctypes.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def fit(self, X, Y, epochs=5, batch=1024, lr=1e-3, standardize=True):
        X = _f32(X, (-1, self.n_in))
        Y = _f32(Y, (-1, self.n_out))
        loss = ctypes.c_double()
        check(lib().mqe_mlp_fit(self.h, X.shape[0], X.ctypes.data_as(_fp), Y.ctypes.data_as(_fp), int(epochs),
                                int(batch), float(lr), 1 if standardize else 0, ctypes.byref(loss)))
        return loss.value
