"""Population snapshot and exact resume (mqe_save_state / mqe_load_state)."""
import numpy as np
import pytest

from mpsqmc_amd import systems
from mpsqmc_amd.engine import Engine
from tests.common import random_trial


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_snapshot_resumes_the_walk_exactly(tmp_path, require_gpu, precision):
    s = systems.chain(12)
    vd, data = random_trial(12, 8, 5)

    def make():
        e = Engine(s, vd, data, n_desired=64, precision=precision, seed=11)
        e.set_canonical_interval(3)  # the snapshot sits between two re-orthonormalisations
        return e

    a = make()
    a.load_copies(vd, data, 64)
    a.prepare()
    a.run(1, 5)
    path = str(tmp_path / "walkers.state")
    a.save_state(path, step=5)
    a.run(6, 4)
    ta = a.read_trace(6, 4)
    sa = a.scalars()
    a.close()

    b = make()
    assert b.load_state(path) == 5
    b.run(6, 4)
    tb = b.read_trace(6, 4)
    sb = b.scalars()
    b.close()
    np.testing.assert_array_equal(ta, tb)
    for k in ("w", "O", "eh", "EL"):
        np.testing.assert_array_equal(sa[k], sb[k])


@pytest.mark.gpu
def test_snapshot_rejects_other_parameters(tmp_path, require_gpu):
    from mpsqmc_amd._lib import MqeError
    s = systems.chain(12)
    vd, data = random_trial(12, 8, 5)
    a = Engine(s, vd, data, n_desired=16, precision="fp64", seed=1)
    a.load_copies(vd, data, 16)
    a.prepare()
    path = str(tmp_path / "w.state")
    a.save_state(path)
    a.close()
    b = Engine(s, vd, data, n_desired=16, precision="fp32", seed=1)
    with pytest.raises(MqeError):
        b.load_state(path)
    b.close()
