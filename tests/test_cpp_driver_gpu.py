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
