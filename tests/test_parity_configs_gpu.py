"""Oracle parity at the sizes of the benchmark configurations.

- L = 16 chain, DMRG trial D = 16, 128 walkers (trial copies), FP64: three Walk
  steps on the host random stream, every sampled grid point identical and
  weights, overlaps, local energies, E_P and the population to 1e-10.
- L = 32 chain, D = 32 (the benchmark shape), FP32 tile path: three steps,
  tolerance 2e-4 on weights, overlaps and local energies.
- L = 32, D = 32: a whole walk in FP32 against the same walk in FP64 (same
  device random numbers): projected energies within 1e-4.
- 6x6 open lattice on the snake path (L = 36, 60 couplings up to range 11),
  trial D = 128: one step of a walker against the oracle in FP64 (1e-10) and
  FP32 (2e-4), and FP32 against FP64 over several device-RNG steps.
- L = 32, D = 32 with a learned proposal table: samples and importance weights
  against the oracle's restatement.
"""
import os

import numpy as np
import pytest

from mpsqmc_amd import mps as M
from mpsqmc_amd import systems
from mpsqmc_amd.analysis import blocking
from mpsqmc_amd.engine import Engine
from mpsqmc_amd.rng import Random

from .common import oracle_population, random_trial, rel

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

THREADS = min(16, os.cpu_count() or 1)


def _walk_steps(s, tvd, tdata, vds, datas, O, precision, steps, tol, seed, n_desired=None, index_frac=1.0,
                energy_mode=0):
    """Walk steps on the host random stream, engine against oracle."""
    n = len(vds)
    N = n_desired or n
    orc = oracle_population(s, tvd, tdata, threads=THREADS, energy_mode=energy_mode)
    orc.set_walkers(list(vds), list(datas), np.ones(n), np.asarray(O, float), np.zeros(n))
    eng = Engine(s, tvd, tdata, n_desired=N, capacity=4 * N, max_bond=int(max(v.max() for v in vds)),
                 precision=precision, seed=seed)
    eng.load_walkers(list(vds), list(datas), np.ones(n), np.asarray(O, float), np.zeros(n))
    E_e = eng.prepare()
    E_o = orc.energy(1, False)[0]
    assert abs(E_e - E_o) <= tol * abs(E_o)
    rng = Random(seed)
    for step in range(1, steps + 1):
        nw = orc.nwalkers
        assert eng.num_walkers() == nw
        u = rng.uniforms(nw * orc.nC)
        sw_o, idx_o = orc.propagate(u)
        sw_e = eng.propagate(u)
        same = np.all(eng.indices(nw) == idx_o, axis=1)
        assert same.mean() >= index_frac, f"step {step}: sampled grid points differ ({same.mean():.3f})"
        wo, Oo, _ = orc.scalars()
        se = eng.scalars()
        assert rel(se["w"][same], wo[same]) < tol, f"step {step}: weights"
        assert rel(se["O"][same], Oo[same]) < tol, f"step {step}: overlaps"
        EP_o, om_o, EL_o = orc.energy(step, True)
        EP_e, om_e = eng.energy(step, True)
        assert rel(eng.scalars()["EL"][same], EL_o[same]) < 10 * tol, f"step {step}: local energies"
        if not same.all():
            break  # a walker took another sample: the trajectories part
        assert abs(sw_e - sw_o) <= tol * abs(sw_o)
        assert abs(EP_e - EP_o) <= tol * abs(EP_o)
        scaling = N / sw_o
        upc = rng.peek(nw)
        n_o, used_o, _ = orc.popcontrol(scaling, upc, 4 * N)
        n_e, used_e = eng.population_control(scaling, upc)
        rng.discard(used_o)
        assert n_o == n_e and used_o == used_e
    eng.close()


def test_chain16_d16_n128_fp64():
    s = systems.chain(16)
    tvd, tdata, _ = M.dmrg(s, 16, max_sweeps=6, tol=1e-10, seed=1)
    n = 128
    _walk_steps(s, tvd, tdata, [tvd] * n, [tdata] * n, np.ones(n), "fp64", steps=3, tol=1e-10, seed=20131015)


def test_chain32_d32_fp32_three_steps():
    s = systems.chain(32)
    tvd, tdata = random_trial(32, 32, 19)
    n = 8
    _walk_steps(s, tvd, tdata, [tvd] * n, [tdata] * n, np.ones(n), "fp32", steps=3, tol=2e-4, seed=19,
                index_frac=0.98)


def test_chain32_d32_walk_energy_fp32_vs_fp64():
    s = systems.chain(32)
    vd, data, _ = M.dmrg(s, 32, max_sweeps=6, tol=1e-10, seed=1)
    out = {}
    for prec in ("fp32", "fp64"):
        e = Engine(s, vd, data, n_desired=256, capacity=768, max_bond=32, precision=prec, seed=7)
        e.load_copies(vd, data, 256)
        e.prepare()
        e.run(1, 350)
        out[prec] = (blocking(e.read_trace(51, 300)[:, 2]), e.path)
        e.close()
    (b32, p32), (b64, p64) = out["fp32"], out["fp64"]
    assert p32 == "tile32" and p64 == "general"
    diff = abs(b32["mean"] - b64["mean"])
    assert diff < 1e-4, diff
    assert diff < 4 * np.hypot(b32["error"], b64["error"]) + 1e-7


def _snake6x6_walkers():
    s = systems.snake_lattice(6, 6)
    tvd, tdata = random_trial(s.L, 128, 3)
    # walkers: the trial and a walker close to it (overlaps ~ 1 keep FP32 well conditioned)
    wvd, wdata = random_trial(s.L, 128, 4)
    mix = tdata + 0.05 * wdata * (np.linalg.norm(tdata) / np.linalg.norm(wdata))
    mix, _ = M.left_normalize(tvd, mix)
    return s, tvd, tdata, [tvd, tvd], [tdata, mix]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-4)])
def test_snake6x6_d128_one_step(precision, tol):
    # one walker (the perturbed trial): the oracle's step costs ~0.6 TFLOP per walker
    s, tvd, tdata, vds, datas = _snake6x6_walkers()
    vds, datas = vds[1:], datas[1:]
    import oracle
    O = [abs(oracle.inner(s.L, 2, tvd, tdata, v, d)) for v, d in zip(vds, datas)]
    # the oracle's local energy by operator insertion (H|trial> uncompressed would hold tens of GB)
    _walk_steps(s, tvd, tdata, vds, datas, O, precision, steps=1, tol=tol, seed=5, energy_mode=1)


def test_snake6x6_d128_fp32_vs_fp64_device_steps():
    s, tvd, tdata, vds, datas = _snake6x6_walkers()
    vds, datas = vds * 2, datas * 2
    res = {}
    for prec in ("fp32", "fp64"):
        e = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=128, precision=prec, seed=9)
        e.load_walkers(vds, datas)
        e.prepare()
        e.run(1, 4)
        res[prec] = (e.read_trace(1, 4), e.scalars())
        e.close()
    (t32, s32), (t64, s64) = res["fp32"], res["fp64"]
    assert np.array_equal(t32[:, 1], t64[:, 1])  # same copies
    assert rel(t32[:, 2], t64[:, 2]) < 2e-4      # E_P
    assert rel(t32[:, 5], t64[:, 5]) < 2e-4      # sum of weights
    assert rel(s32["O"], s64["O"]) < 2e-4 and rel(s32["EL"], s64["EL"]) < 2e-3


def test_learned_proposal_table_chain32_d32():
    # the learned proposal at the benchmark shape: predictions = the exact labels
    # perturbed by 10 %, mix 0.2; samples and importance weights against the oracle
    n = 8
    s = systems.chain(32)
    tvd, tdata = random_trial(32, 32, 25)
    vds, datas = [tvd] * n, [tdata] * n
    w, O, eh = np.ones(n), np.ones(n), np.zeros(n)

    def pop():
        orc = oracle_population(s, tvd, tdata, threads=THREADS)
        orc.set_walkers(vds, datas, w, O, eh)
        return orc
    u = np.random.default_rng(26).uniform(size=n * 31)
    _, _, lab = pop().propagate_with_labels(u)
    pred = (lab * (1 + 0.1 * np.random.default_rng(27).normal(size=lab.shape))).astype(np.float32)
    eng = Engine(s, tvd, tdata, n_desired=n, capacity=4 * n, max_bond=32, precision="fp32", seed=25)
    assert eng.path == "tile32"
    eng.load_walkers(vds, datas, w, O, eh)
    eng.prepare()
    eng.set_proposal_table(pred, mix=0.2)
    orc = pop()
    orc.set_proposal(pred.astype(np.float64), mix=0.2)
    sw_e = eng.propagate(u)
    sw_o, idx_o = orc.propagate(u)
    same = np.all(eng.indices(n) == idx_o, axis=1)
    assert same.mean() >= 0.75
    w_e, w_o = eng.scalars()["w"], orc.scalars()[0]
    assert np.all(np.abs(w_e[same] - w_o[same]) <= 2e-4 * np.abs(w_o[same]))
    if same.all():
        assert abs(sw_e - sw_o) <= 2e-4 * abs(sw_o)
    eng.close()
