"""Step control of the walker engine: device random numbers of explicit calls,
graph replay of mqe_run, population growth after a resume, and the capacity
error path.

- mqe_propagate / mqe_population_control without host draws: the k-th call
  draws the numbers of step k of mqe_run, so a walk driven call by call
  (the reference's MPSQMC2::Walk order, MPSQMC2.cpp:465-511) equals mqe_run.
- mqe_run(..., use_graph=1) replays a captured graph; every replay equals the
  eager launches bit for bit.
"""
import numpy as np
import pytest

from mpsqmc_amd import systems
from mpsqmc_amd._lib import MqeError, check, lib
from mpsqmc_amd.engine import Engine

from .common import random_trial

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _engine(precision, L=12, D=8, n=32, seed=11, n_load=None, canon=3, dtau=0.01):
    s = systems.chain(L, dtau=dtau)
    vd, data = random_trial(L, D, 5)
    e = Engine(s, vd, data, n_desired=n, capacity=4 * n, max_bond=D, precision=precision, seed=seed)
    e.set_canonical_interval(canon)
    if n_load != 0:
        e.load_copies(vd, data, n if n_load is None else n_load)
        e.prepare()
    return e


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_device_draws_differ_between_calls(precision):
    e = _engine(precision)
    e.propagate(None)
    a = e.indices().copy()
    e.propagate(None)
    b = e.indices().copy()
    e.close()
    assert a.shape == b.shape
    assert np.mean(a != b) > 0.5, "two propagations drew the same grid points"


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_call_by_call_walk_equals_run(precision):
    steps = 6
    a = _engine(precision)
    rows = []
    for step in range(1, steps + 1):
        n = a.num_walkers()
        sw = a.propagate(None)
        avg = sw / n
        scaling = a.n_desired / (n * avg)  # MPSQMC2.cpp:480-484
        EP, om = a.energy(step, True)
        rows.append((step, n, EP, np.log(scaling) / a.system.dtau, om, sw))
        a.population_control(scaling, None)
    sa = a.scalars()
    a.close()
    b = _engine(precision)
    b.run(1, steps)
    tb = b.read_trace(1, steps)
    sb = b.scalars()
    b.close()
    ra = np.array(rows)
    for col in (0, 1, 2, 4, 5):
        np.testing.assert_array_equal(ra[:, col], tb[:, col])
    np.testing.assert_allclose(ra[:, 3], tb[:, 3], rtol=1e-14, atol=1e-12)  # host and device log
    for k in ("w", "O", "eh", "EL"):
        np.testing.assert_array_equal(sa[k], sb[k])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_graph_replay_equals_eager(precision):
    k = 5
    a = _engine(precision, n_load=16)  # the population grows: copies inside the replayed steps
    a.run(1, k)
    a.run(k + 1, 3)
    a.run(k + 4, k)
    ta = a.read_trace(1, 2 * k + 3)
    sa = a.scalars()
    a.close()
    b = _engine(precision, n_load=16)
    b.run(1, k, use_graph=True)
    b.sync()
    tb1 = b.read_trace(1, k)
    b.run(k + 1, 3)            # eager steps between two replays of the same graph
    check(lib().mqe_step_graph(b.h, k + 4, k))
    tb = b.read_trace(1, 2 * k + 3)
    sb = b.scalars()
    b.close()
    np.testing.assert_array_equal(tb1, ta[:k])
    np.testing.assert_array_equal(tb, ta)
    assert tb[-1, 1] > 16
    for key in ("w", "O", "eh", "EL"):
        np.testing.assert_array_equal(sa[key], sb[key])


def test_graph_replay_rejects_the_learned_proposal():
    e = _engine("fp32")
    e.collect_samples(1000)
    with pytest.raises(MqeError):
        e.run(1, 2, use_graph=True)
    e.collect_samples(0)
    e.run(1, 2, use_graph=True)
    e.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_resume_with_population_growth(tmp_path, precision):
    # 8 walkers loaded for 32 desired, snapshot before the first step: population
    # control splits every walker after the resume, so copies land in slots the
    # snapshot never held (their boundary environments must be set)
    a = _engine(precision, n_load=8)
    path = str(tmp_path / "w.state")
    a.save_state(path, step=0)
    a.run(1, 4)
    ta = a.read_trace(1, 4)
    sa = a.scalars()
    a.close()
    b = _engine(precision, n_load=0)  # a fresh engine: no walkers loaded before the snapshot
    assert b.load_state(path) == 0
    assert b.num_walkers() == 8
    b.run(1, 4)
    tb = b.read_trace(1, 4)
    sb = b.scalars()
    b.close()
    assert tb[0, 1] == 8 and tb[1, 1] > 8, "no walker copies after the resume"
    np.testing.assert_array_equal(ta, tb)
    for k in ("w", "O", "eh", "EL"):
        np.testing.assert_array_equal(sa[k], sb[k])


def test_capacity_overflow_leaves_the_population():
    s = systems.chain(8)
    vd, data = random_trial(8, 4, 3)
    e = Engine(s, vd, data, n_desired=1000, capacity=16, max_bond=4, precision="fp64", seed=2)
    e.load_copies(vd, data, 8)
    e.prepare()
    sw = e.propagate(None)
    before = e.scalars()
    with pytest.raises(MqeError):
        e.population_control(1000.0 / sw)  # about 125 copies of each walker: far beyond 16 slots
    after = e.scalars()
    assert e.num_walkers() == 8
    for k in ("w", "O", "eh", "EL"):
        np.testing.assert_array_equal(before[k], after[k])
    with pytest.raises(MqeError):
        e.propagate(None)
    e.close()


def test_long_walk_every_walker_re_orthonormalises():
    # the benchmark walk (L = 32, D = 32 DMRG trial, 2048 walkers) over 1920 steps:
    # staggered by slot alone, a line of copies moved to new slots could go
    # without re-orthonormalisation for hundreds of steps, and its FP32 gauge
    # drifted until a PDF norm vanished near step 1800; the per-walker counter
    # bounds every walker's interval to the default 64
    from mpsqmc_amd import mps as M
    s = systems.chain(32, dtau=0.01)
    vd, data, _ = M.dmrg(s, 32, max_sweeps=6, tol=1e-8, seed=20131015)
    e = Engine(s, vd, data, n_desired=2048, capacity=6144, max_bond=32, precision="fp32", seed=20131015)
    e.load_copies(vd, data, 2048)
    e.prepare()
    for first in range(1, 1921, 320):
        e.run(first, 320)
        e.sync()  # raises on a zero PDF norm or a non-finite weight
    tr = e.read_trace(1, 1920)
    e.close()
    assert np.all(np.isfinite(tr))
    assert abs(tr[-200:, 2].mean() - (-13.99732)) < 1e-4
