"""mqe_overlaps_noalloc (the transfer-matrix overlaps in a static device
workspace, for callers that cannot allocate) against mqe_overlaps on the same
engine: two calls in a row, the second with a smaller input, bit-identical in
FP64 and FP32 (the same kernel runs on the same data)."""
import ctypes

import numpy as np
import pytest

from mpsqmc_amd import systems
from mpsqmc_amd._lib import MqeError, MqeMps, check, lib
from mpsqmc_amd.engine import Engine

from .common import random_trial

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _noalloc(eng, vd, data):
    vd = np.ascontiguousarray(vd, dtype=np.int32)
    data = np.ascontiguousarray(data, dtype=np.float64)
    m = MqeMps(eng.L, 2, vd.ctypes.data_as(_ip), data.ctypes.data_as(_dp))
    out = np.zeros(eng.num_walkers())
    check(lib().mqe_overlaps_noalloc(eng.h, ctypes.byref(m), out.ctypes.data_as(_dp)))
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_overlaps_match(precision):
    s = systems.chain(16)
    tvd, tdata = random_trial(16, 16, 3)
    eng = Engine(s, tvd, tdata, n_desired=64, capacity=256, max_bond=16, precision=precision, seed=4)
    vds, datas = zip(*[random_trial(16, 16, 10 + i) for i in range(64)])
    eng.load_walkers(list(vds), list(datas))
    big = random_trial(16, 24, 5)      # first call: a probe of bond 24
    small = random_trial(16, 6, 6)     # second call: a smaller probe
    for vd, data in (big, small):
        want = eng.overlaps(vd, data)
        got = _noalloc(eng, vd, data)
        np.testing.assert_array_equal(got, want)
    # the workspace keeps the last call's contents: a third call on the first input again
    np.testing.assert_array_equal(_noalloc(eng, *big), eng.overlaps(*big))
    eng.close()


def test_noalloc_rejects_oversized_input():
    # one walker's transfer matrices at bond 1500 (4 x 1500^2 doubles = 72 MB) exceed the 64 MiB workspace
    from mpsqmc_amd import mps as M
    s = systems.chain(24)
    tvd, tdata = random_trial(24, 4, 3)
    eng = Engine(s, tvd, tdata, n_desired=2, capacity=8, max_bond=4, precision="fp64", seed=4)
    eng.load_copies(tvd, tdata, 2)
    vd, data = M.random_mps(24, 1500, 2, 7)  # rejected before any staging or kernel
    with pytest.raises(MqeError):
        _noalloc(eng, vd, data)
    np.testing.assert_allclose(eng.overlaps(tvd, tdata), 1.0, rtol=1e-12)  # the engine is still usable
    eng.close()


def test_noalloc_overlaps_in_chunks():
    # D = 128 in FP64: 512 KiB of transfer matrices per walker, so the 64 MiB
    # workspace takes the 200 walkers in two launches
    s = systems.chain(20)
    tvd, tdata = random_trial(20, 128, 8)
    eng = Engine(s, tvd, tdata, n_desired=200, capacity=400, max_bond=128, precision="fp64", seed=4)
    wv, wd = random_trial(20, 128, 9)
    eng.load_walkers([tvd, wv] * 100, [tdata, wd] * 100)
    probe = random_trial(20, 64, 10)
    np.testing.assert_array_equal(_noalloc(eng, *probe), eng.overlaps(*probe))
    eng.close()
