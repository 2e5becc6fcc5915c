"""Host random stream of the parity contract (Random.cpp:14-59 with a fixed seed).

One std::mt19937 stream, u = x * 2^-32.  Per step the propagation consumes
n_walkers * n_couplings draws (walker-major, couplings in lexicographic order),
then population control consumes one draw per walker whose scaled weight lies
outside [0.25, 1.5], in walker order.
"""
import ctypes

import numpy as np

from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


class Random:
    def __init__(self, seed):
        h = ctypes.c_void_p()
        check(lib().mqe_rng_create(ctypes.c_uint(seed & 0xFFFFFFFF), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().mqe_rng_destroy(self.h)
            self.h = None

    def uniforms(self, n):
        out = np.zeros(n)
        check(lib().mqe_rng_uniforms(self.h, n, out.ctypes.data_as(_dp)))
        return out

    def peek(self, n):
        out = np.zeros(n)
        check(lib().mqe_rng_peek(self.h, n, out.ctypes.data_as(_dp)))
        return out

    def discard(self, k):
        check(lib().mqe_rng_discard(self.h, k))

    def rand(self):
        return float(self.uniforms(1)[0])
