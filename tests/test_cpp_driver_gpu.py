"""The C++ host classes (MPSQMC2 / Walker / MPSstate over the C ABI) run a
Walk whose per-step trace equals the oracle's MPSQMC2::Walk on the same trial
file and the same random stream (FP64 engine, host RNG contract)."""
import os
import subprocess

import numpy as np
import pytest

from mpsqmc_amd import mps as M
from mpsqmc_amd import systems

from .common import oracle_population, random_trial

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

WALK = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mpsqmc_amd", "mpsqmc_walk")


def _run(args):
    out = subprocess.run([WALK] + args, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    lines = [ln.split() for ln in out.stdout.strip().splitlines()]
    E0 = float(lines[0][1])
    rows = np.array([[float(x) for x in ln] for ln in lines[1:]])
    return E0, rows


@pytest.mark.parametrize("shape", ["chain", "snake"])
def test_cpp_walk_matches_oracle(tmp_path, shape):
    if shape == "chain":
        s, extra, D = systems.chain(10), [], 6
    else:
        s, extra, D = systems.snake_lattice(3, 3), ["--snake", "3x3"], 4
    vd, data = random_trial(s.L, D, 41)
    path = str(tmp_path / "trial.mps")
    M.write_mps(path, vd, data)
    N, steps, seed = 12, 4, 777
    E0, rows = _run(["--trial", path, "--N", str(N), "--steps", str(steps), "--seed", str(seed)] + extra)
    pop = oracle_population(s, vd, data)
    pop.init_walkers(N, D)
    E0o, tr = pop.walk(steps, seed, N, max(1000, 3 * N))
    assert abs(E0 - E0o) < 1e-10 * abs(E0o)
    assert np.array_equal(rows[:, 1], tr[:, 1])
    for col in (2, 3, 5):
        assert np.allclose(rows[:, col], tr[:, col], rtol=1e-10, atol=1e-12), col
    assert np.allclose(rows[:, 4], tr[:, 4], rtol=1e-6, atol=1e-12)


def test_cpp_walk_compressed_walkers(tmp_path):
    # DW < DT: walkers are the trial compressed to DW (MPSQMC2.cpp:404-409)
    s = systems.chain(10)
    vd, data = random_trial(10, 8, 43)
    path = str(tmp_path / "trial.mps")
    M.write_mps(path, vd, data)
    E0, rows = _run(["--trial", path, "--N", "8", "--steps", "3", "--seed", "5", "--DW", "4"])
    pop = oracle_population(s, vd, data)
    pop.init_walkers(8, 4)
    E0o, tr = pop.walk(3, 5, 8, 1000)
    assert abs(E0 - E0o) < 1e-9 * abs(E0o)
    assert np.allclose(rows[:, 2], tr[:, 2], rtol=1e-9)


def test_cpp_device_rng_walk_follows_the_seed(tmp_path):
    # MPSQMC2 with Rng::Device keys the device random numbers with the caller's
    # seed: its call-by-call walk (propagate, energy, population control) equals
    # mqe_run on the same trial and seed, and another seed walks differently
    from mpsqmc_amd.engine import Engine
    s = systems.chain(10)
    vd, data = random_trial(10, 6, 47)
    path = str(tmp_path / "trial.mps")
    M.write_mps(path, vd, data)
    N, steps = 16, 5
    rows = {}
    for seed in (101, 202):
        E0, r = _run(["--trial", path, "--N", str(N), "--steps", str(steps), "--seed", str(seed), "--device-rng"])
        rows[seed] = r
    # the driver left-normalises the trial it reads (MPSQMC2.cpp:239): same routine here
    _, _, _, tv, td = M.read_mps(path)
    td, _ = M.left_normalize(tv, td)
    eng = Engine(s, tv, td, n_desired=N, precision="fp64", seed=101)
    eng.load_copies(tv, td, N)
    eng.prepare()
    eng.run(1, steps)
    tr = eng.read_trace(1, steps)
    eng.close()
    assert np.array_equal(rows[101][:, 1], tr[:, 1])
    for col in (2, 3, 4, 5):
        assert np.allclose(rows[101][:, col], tr[:, col], rtol=1e-12, atol=1e-14), col
    assert not np.allclose(rows[202][:, 5], tr[:, 5], rtol=1e-9), "the device draws ignore the seed"
