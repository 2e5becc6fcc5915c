"""The roofline numerators quoted by bench.py against the environment-sweep
flop count: the independent per-workload figures of the work analysis (61.5
MFLOP for the L = 32, D = 32 chain; 84.5 / 200 / 325 / 474 MFLOP for the 4x4
snake at walker bonds 16 / 32 / 48 / 64; 25.7 GFLOP for the 6x6 snake at
D = 128), which bench.py quotes, must follow from the formula."""
import importlib.util
import os

from mpsqmc_amd import systems
from mpsqmc_amd.flops import bond_profile, env_flops_fp32

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_chain_count():
    s = systems.chain(32)
    first, second, _ = s.couplings()
    f = env_flops_fp32(first, second, bond_profile(32, 32), bond_profile(32, 32))
    assert f == 61_469_232.0
    assert abs(_bench().F32_PER_WS - f) / f < 1e-3


def test_snake_tables():
    b = _bench()
    for name, wl in b.SNAKES.items():
        s = systems.snake_lattice(wl["rows"], wl["cols"])
        first, second, _ = s.couplings()
        for D in wl["DW"]:
            f = env_flops_fp32(first, second, bond_profile(s.L, wl["DT"]), bond_profile(s.L, D))
            assert abs(wl["F32"][D] - f) / f < 2e-3, (name, D, f)
