"""Projected energy of a walk with FP32 contractions (tile path) against the same
walk in FP64 (general path): equal within statistical error and within 1e-4,
for the L = 16 chain with its DMRG trial (exact diagonalisation: -6.911737145575105)."""
import numpy as np
import pytest

from mpsqmc_amd import mps as M
from mpsqmc_amd import systems
from mpsqmc_amd.analysis import blocking
from mpsqmc_amd.engine import Engine

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]


def _walk(precision, vd, data, n, steps, warm):
    s = systems.chain(16)
    e = Engine(s, vd, data, n_desired=n, capacity=3 * n, max_bond=16, precision=precision, seed=3)
    e.load_copies(vd, data, n)
    e.prepare()
    e.run(1, warm + steps)
    tr = e.read_trace(warm + 1, steps)
    path = e.path
    e.close()
    return blocking(tr[:, 2]), path


def test_fp32_walk_energy_matches_fp64():
    vd, data, E_dmrg = M.dmrg(systems.chain(16), 16, max_sweeps=6, tol=1e-10, seed=1)
    b32, p32 = _walk("fp32", vd, data, 256, 400, 50)
    b64, p64 = _walk("fp64", vd, data, 256, 400, 50)
    assert p32 == "tile32" and p64 == "general"
    diff = abs(b32["mean"] - b64["mean"])
    err = np.hypot(b32["error"], b64["error"])
    assert diff < 4 * err + 1e-7
    assert diff < 1e-4
    # both below the variational DMRG energy and close to exact diagonalisation
    E_ed = -6.911737145575105
    for b in (b32, b64):
        assert abs(b["mean"] - E_ed) < 1e-3
