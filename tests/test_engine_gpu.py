"""GPU parity of the walker engine (C ABI, libmqe.so) against the FP64 oracle on
identical inputs: same trial, tables, grid, walkers and uniforms.

Bars: sampled grid indices identical; weights, overlaps, local energies to
1e-10 relative in FP64 mode; a stated tolerance in FP32 mode."""
import numpy as np
import pytest

from mpsqmc_amd import mps as M
from mpsqmc_amd import systems
from mpsqmc_amd.engine import Engine
from mpsqmc_amd.rng import Random

from .common import dense_to_mps, ground_energy, oracle_population, random_trial, rel

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

FP64_TOL = 1e-10


def _walkers(L, D, n, seed, uneven=False):
    vds, datas = [], []
    rng = np.random.default_rng(seed)
    for i in range(n):
        Di = int(rng.choice([max(1, D // 2), D])) if uneven else D
        vd, data = random_trial(L, Di, seed * 1000 + i)
        vds.append(vd)
        datas.append(data)
    return vds, datas


def _pair(system, D, n, seed, precision="fp64", uneven=False, walkers=None):
    tvd, tdata = random_trial(system.L, D, seed)
    orc = oracle_population(system, tvd, tdata)
    eng = Engine(system, tvd, tdata, n_desired=n, capacity=4 * n, max_bond=D, precision=precision, seed=seed)
    if walkers is None:
        vds, datas = _walkers(system.L, D, n, seed + 1, uneven)
    else:
        vds, datas = walkers
    rng = np.random.default_rng(seed + 2)
    w = rng.uniform(0.5, 1.5, n)
    O = rng.uniform(0.5, 1.5, n)
    eh = rng.uniform(-1, 1, n)
    orc.set_walkers(vds, datas, w, O, eh)
    eng.load_walkers(vds, datas, w, O, eh)
    return orc, eng, (tvd, tdata)


def _same_states(orc, eng, n, atol):
    """States agree up to gauge: compare overlaps with random probe MPS."""
    L = orc.L
    probes = [random_trial(L, 3, 900 + q) for q in range(2)]
    for i in range(n):
        vo, do, *_ = orc.walker(i)
        ve, de = eng.walker(i)
        for pv, pd in probes:
            a = M.to_dense(vo, do) @ M.to_dense(pv, pd) if L <= 12 else None
            b = M.to_dense(ve, de) @ M.to_dense(pv, pd) if L <= 12 else None
            if a is not None:
                assert abs(a - b) <= atol * max(1.0, abs(a))


def test_overlap_kernel_fp64():
    s = systems.chain(10)
    orc, eng, (tvd, tdata) = _pair(s, 6, 12, 3)
    ov = eng.overlaps(tvd, tdata)
    import oracle
    for i in range(12):
        vo, do, *_ = orc.walker(i)
        want = oracle.inner(10, 2, tvd, tdata, vo, do)
        assert abs(ov[i] - want) <= 1e-12 * max(1.0, abs(want))
    pv, pd = random_trial(10, 9, 77)
    ov = eng.overlaps(pv, pd)
    for i in range(12):
        vo, do, *_ = orc.walker(i)
        want = oracle.inner(10, 2, pv, pd, vo, do)
        assert abs(ov[i] - want) <= 1e-12 * max(1.0, abs(want))


def test_prepare_energy_fp64():
    s = systems.chain(8)
    orc, eng, _ = _pair(s, 4, 10, 5)
    E = eng.prepare()
    EP, om, EL = orc.energy(1, dofluct=False)
    assert abs(E - EP) <= FP64_TOL * abs(EP)
    sc = eng.scalars()
    assert rel(sc["EL"], EL) < FP64_TOL


def _run_parity(system, D, n, seed, steps, precision="fp64", tol=FP64_TOL, uneven=False, check_states=True,
                canon=None):
    orc, eng, _ = _pair(system, D, n, seed, precision, uneven)
    if canon is not None:
        eng.set_canonical_interval(canon)
    E_e = eng.prepare()
    E_o, _, _ = orc.energy(1, dofluct=False)
    assert abs(E_e - E_o) <= tol * abs(E_o)
    rng = Random(seed)
    carry = tol  # loosest per-walker tolerance so far: deviations are inherited by later steps and copies
    for step in range(1, steps + 1):
        nw = orc.nwalkers
        assert eng.num_walkers() == nw
        # FP32: the relative error of overlap-derived quantities grows like
        # u_fp32 |trial| |phi| / |O|; random test walkers can be nearly orthogonal
        # to the trial (O ~ 5e-5), so the per-walker tolerance is scaled by
        # max(1, 1e-3 / O) (walkers of a real walk have O ~ 1)
        O_before = np.abs(eng.scalars()["O"])
        wtol = tol * (np.maximum(1.0, 1e-3 / O_before) if precision == "fp32" else np.ones(nw))
        wtol = np.maximum(wtol, carry)
        carry = float(np.max(wtol))
        u = rng.uniforms(nw * orc.nC)
        sw_o, idx_o = orc.propagate(u)
        sw_e = eng.propagate(u)
        idx_e = eng.indices(nw)
        if precision == "fp64":
            assert np.array_equal(idx_o, idx_e), f"step {step}: sampled grid points differ"
        else:
            assert np.mean(idx_o == idx_e) > 0.98
        wo, Oo, eho = orc.scalars()
        se = eng.scalars()
        same = np.all(idx_o == idx_e, axis=1)
        dw = np.abs(se["w"] - wo) / np.abs(wo)
        dO = np.abs(se["O"] - Oo) / np.abs(Oo)
        if not same.all():
            # FP32: a uniform within rounding of a CDF boundary picked a neighbouring
            # grid point, after which that walker's trajectory differs; compare the
            # walkers with identical samples and stop
            assert np.all(dw[same] < wtol[same]) and np.all(dO[same] < wtol[same])
            return
        assert abs(sw_e - sw_o) <= np.sum(wtol * np.abs(wo))
        assert np.all(dw < wtol), f"step {step}: weights"
        assert np.all(dO < wtol), f"step {step}: overlaps"
        if check_states and system.L <= 12 and precision == "fp64":
            _same_states(orc, eng, nw, 1e-9)
        EP_o, om_o, EL_o = orc.energy(step, True)
        EP_e, om_e = eng.energy(step, True)
        dEL = np.abs(eng.scalars()["EL"] - EL_o) / np.maximum(np.abs(EL_o), 1e-300)
        assert np.all(dEL < wtol * 10), f"step {step}: local energies"
        assert abs(EP_e - EP_o) <= tol * abs(EP_o) * np.max(wtol / tol)
        assert abs(om_e - om_o) <= max((1e-7 if precision == "fp64" else tol) * abs(om_o), 1e-13)
        scaling = n / sw_o  # N_desired / (N_cur * avg w), MPSQMC2.cpp:480-483
        k = nw
        upc = rng.peek(k)
        n_o, used_o, nc_o = orc.popcontrol(scaling, upc, 4 * n)
        n_e, used_e = eng.population_control(scaling, upc)
        rng.discard(used_o)
        assert n_o == n_e and used_o == used_e
        wo, Oo, eho = orc.scalars()
        se = eng.scalars()
        # copies inherit their parent's weight: the loosest per-walker tolerance applies
        assert rel(se["w"], wo) < np.max(wtol) and rel(se["eh"], eho) < np.max(wtol) * 10


def test_parity_chain_fp64():
    _run_parity(systems.chain(8), 4, 12, 11, steps=3)


def test_parity_chain_field_fp64():
    _run_parity(systems.chain(7, h=0.7), 4, 8, 12, steps=2)


def test_parity_snake_fp64():
    # 3x3 open lattice on a snake path: couplings of range up to 5
    _run_parity(systems.snake_lattice(3, 3), 4, 8, 13, steps=2)


def test_parity_uneven_bonds_fp64():
    _run_parity(systems.chain(8), 8, 10, 14, steps=2, uneven=True)


def test_parity_chain_fp32():
    _run_parity(systems.chain(8), 4, 12, 15, steps=2, precision="fp32", tol=2e-4, check_states=False)


def test_parity_snake_fp32_general():
    # long-range couplings take the general path; bonds of 16 run its products on the matrix cores
    _run_parity(systems.snake_lattice(3, 3), 16, 8, 16, steps=2, precision="fp32", tol=2e-4, check_states=False)


def test_parity_chain_field_fp32_general():
    # a field takes the general path (one-site terms and full re-canonicalisation)
    _run_parity(systems.chain(10, h=0.5), 16, 8, 17, steps=2, precision="fp32", tol=2e-4, check_states=False)


def test_walk_from_trial_copies_fp64():
    # SetupWalkers with copyTrial (MPSQMC2.cpp:400-409): O = 1 at start
    s = systems.chain(8)
    tvd, tdata = random_trial(8, 4, 31)
    orc = oracle_population(s, tvd, tdata)
    orc.init_walkers(6, 4)
    eng = Engine(s, tvd, tdata, n_desired=6, capacity=24, max_bond=4, precision="fp64")
    eng.load_copies(tvd, tdata, 6)
    assert abs(eng.prepare() - orc.energy(1, False)[0]) < 1e-10 * 4
    E0, trace = orc.walk(3, 99, 6, 24)
    rng = Random(99)
    for step in range(1, 4):
        n = eng.num_walkers()
        sw = eng.propagate(rng.uniforms(n * eng.nC))
        scaling = 6 / sw
        EP, om = eng.energy(step, True)
        assert abs(EP - trace[step - 1, 2]) <= 1e-10 * abs(EP)
        assert abs(np.log(scaling) / s.dtau - trace[step - 1, 3]) <= 1e-8
        nn, used = eng.population_control(scaling, rng.peek(n))
        rng.discard(used)


def test_exact_trial_energy_is_e0():
    s = systems.chain(10)
    E0 = ground_energy(s)
    import scipy.sparse.linalg as spla
    from .common import dense_hamiltonian
    w, V = spla.eigsh(dense_hamiltonian(s), k=1, which="SA", tol=1e-13)
    vd, data = dense_to_mps(np.real(V[:, 0]) / np.linalg.norm(V[:, 0]), 10, 32)
    for prec, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
        eng = Engine(s, vd, data, n_desired=16, capacity=64, max_bond=32, precision=prec, seed=5)
        eng.load_copies(vd, data, 16)
        assert abs(eng.prepare() - E0) < tol * abs(E0)
        eng.run(1, 4)
        tr = eng.read_trace(1, 4)
        assert np.allclose(tr[:, 2], E0, rtol=tol)
        eng.close()


def test_device_step_statistics_fp32():
    # device RNG path (mqe_step): the population stays near N and the projected
    # energy of a ground-state-like trial (ED state truncated to D = 4) stays
    # within a small window of the exact energy
    s = systems.chain(10)
    E0 = ground_energy(s)
    import scipy.sparse.linalg as spla
    from .common import dense_hamiltonian
    w, V = spla.eigsh(dense_hamiltonian(s), k=1, which="SA", tol=1e-13)
    vd, data = dense_to_mps(np.real(V[:, 0]) / np.linalg.norm(V[:, 0]), 10, 4)
    data, _ = M.left_normalize(vd, data)
    eng = Engine(s, vd, data, n_desired=256, capacity=1024, max_bond=4, precision="fp32", seed=41)
    eng.load_copies(vd, data, 256)
    eng.prepare()
    eng.run(1, 60)
    tr = eng.read_trace(1, 60)
    assert np.all(np.isfinite(tr))
    assert np.all((tr[:, 1] > 128) & (tr[:, 1] < 768))
    assert abs(tr[30:, 2].mean() - E0) < 0.05
    eng.close()


def test_tile32_path_selected():
    s = systems.chain(32)
    vd, data = random_trial(32, 32, 3)
    eng = Engine(s, vd, data, n_desired=4, capacity=16, max_bond=32, precision="fp32")
    assert eng.path == "tile32"
    eng.close()
    eng = Engine(s, vd, data, n_desired=4, capacity=16, max_bond=32, precision="fp64")
    assert eng.path == "general"
    eng.close()


@pytest.mark.parametrize("canon", [1, 2, 4])
def test_tile32_parity_chain16_fp32(canon):
    # FP32 MFMA path against the FP64 oracle: tolerance 2e-4 relative; the LQ
    # re-canonicalisation every `canon` steps changes the gauge only
    _run_parity(systems.chain(16), 16, 8, 17, steps=5, precision="fp32", tol=2e-4, check_states=False, canon=canon)


def test_parity_large_bond_fp64():
    # bond 128 in the middle of an L = 14 chain: the 256 x 128 FP64 matrix of the LQ
    # exceeds the LDS workspace, so its leading columns stay in global scratch
    import os
    s = systems.chain(14)
    tvd, tdata = random_trial(14, 128, 5)
    orc = oracle_population(s, tvd, tdata, threads=min(16, os.cpu_count() or 1))
    vds, datas = zip(*[random_trial(14, 128, 50 + i) for i in range(4)])
    w = np.array([1.0, 0.8, 1.2, 0.9])
    orc.set_walkers(list(vds), list(datas), w, np.ones(4), np.zeros(4))
    eng = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=128, precision="fp64")
    assert eng.path == "general"
    eng.load_walkers(list(vds), list(datas), w, np.ones(4), np.zeros(4))
    E_e = eng.prepare()
    E_o = orc.energy(1, False)[0]
    assert abs(E_e - E_o) <= FP64_TOL * abs(E_o)
    u = Random(21).uniforms(4 * orc.nC)
    sw_o, idx_o = orc.propagate(u)
    sw_e = eng.propagate(u)
    assert np.array_equal(idx_o, eng.indices(4))
    assert abs(sw_e - sw_o) <= FP64_TOL * abs(sw_o)
    wo, Oo, _ = orc.scalars()
    sc = eng.scalars()
    assert rel(sc["w"], wo) < FP64_TOL and rel(sc["O"], Oo) < FP64_TOL
    eng.close()


def test_light_sweep_long_chain_d128_fp32():
    # steps between re-orthonormalisations rescale each site to the magnitude of a
    # right-canonical one; 36 sites at D = 128 with walkers nearly orthogonal to
    # the trial (O ~ 1e-6) are where a wrong target scale underflows FP32.  The
    # light steps must give the weights, overlaps and local energies of
    # re-orthonormalising every step (same walkers, same uniforms).
    s = systems.chain(36)
    tvd, tdata = random_trial(36, 128, 1)
    vds, datas = zip(*[random_trial(36, 128, 40 + i) for i in range(4)])
    runs = []
    for canon in (64, 1):
        eng = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=128, precision="fp32", seed=3)
        assert eng.path == "general"
        eng.set_canonical_interval(canon)
        eng.load_walkers(list(vds), list(datas), np.ones(4), np.ones(4), np.zeros(4))
        eng.prepare()
        rng = Random(9)
        out = []
        for step in range(3):
            eng.propagate(rng.uniforms(4 * eng.nC))
            sc = eng.scalars()
            out.append((eng.indices(4).copy(), sc["w"].copy(), sc["O"].copy(), sc["EL"].copy()))
        eng.close()
        runs.append(out)
    for (i1, w1, O1, E1), (i2, w2, O2, E2) in zip(*runs):
        assert np.all(np.isfinite(w1)) and np.all(w1 > 0) and np.all(np.isfinite(E1))
        same = np.all(i1 == i2, axis=1)
        if not same.any():
            break
        # FP32 conditioning of overlap-derived quantities, as in the parity tests
        tol = 2e-4 * np.maximum(1.0, 1e-3 / np.abs(O2[same]))
        assert np.all(np.abs(w1[same] - w2[same]) <= tol * np.abs(w2[same]))
        assert np.all(np.abs(O1[same] - O2[same]) <= tol * np.abs(O2[same]))
        assert np.all(np.abs(E1[same] - E2[same]) <= 10 * tol * np.maximum(np.abs(E2[same]), 1.0))
        if not same.all():
            break
