"""Device recompression of walkers (mqe_compress_walkers) against the oracle's
CompressState + setOverlap (MPSstate.cpp:379-404, TwoSiteObject.cpp:71-184,
Walker.cpp:54-66), and the uneven-bond snake workload (4x4 open lattice, trial
D_T = 64, walkers compressed to D_W in {16, 32, 48, 64}) set up on the device
and propagated against the oracle.

Bars: bond dimensions identical; overlaps with the trial to 1e-10 relative in
FP64 and 1e-5 in FP32 (walker tensors stored in FP32); states equal to 1e-10
(dense, small L); energies and propagation as in the parity tests.

The state compressed is the trial in a random gauge: in canonical form the
first sweep meets exactly degenerate two-site singular values near the chain
ends (sqrt(2), several times), and which subspace a truncation there keeps
depends on the SVD implementation (the reference's is LAPACK's dgesdd); in a
generic gauge the truncation is well defined and two independent
implementations agree (tests/test_abi_host.py)."""
import os

import numpy as np
import pytest

import oracle
from mpsqmc_amd import mps as M
from mpsqmc_amd import systems
from mpsqmc_amd.engine import Engine
from mpsqmc_amd.rng import Random

from .common import oracle_population, random_gauge, random_trial, rel

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

THREADS = min(16, os.cpu_count() or 1)


def _oracle_compressed(L, vd0, data0, D, tvd, tdata):
    """SetupWalkers for DW < DT on the state (vd0, data0): CompressState(D),
    LeftNormalize, overlap with the trial, phase."""
    vd, data, _ = oracle.compress(L, 2, vd0, data0, D)
    _, data = oracle.left_normalize(L, 2, vd, data)
    O = oracle.inner(L, 2, tvd, tdata, vd, data)
    if O < 0:
        data = data.copy()
        n0 = int(vd[0]) * int(vd[1]) * 2
        data[:n0] *= -1.0
        O = -O
    return vd, data, O


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-5)])
def test_compress_matches_oracle(precision, tol):
    L = 10
    s = systems.chain(L)
    tvd, tdata = random_trial(L, 16, 3)
    gvd, gdata = random_gauge(tvd, tdata, 4)
    targets = [2, 4, 8, 3, 16, 0]  # 16: only the 1e-13 Schmidt cut; 0: left unchanged
    n = len(targets)
    eng = Engine(s, tvd, tdata, n_desired=n, capacity=4 * n, max_bond=16, precision=precision, seed=1)
    eng.load_copies(gvd, gdata, n)
    eng.compress_walkers(targets)
    O = eng.scalars()["O"]
    for i, D in enumerate(targets):
        vd, data = eng.walker(i)
        if D < 1:
            assert np.array_equal(vd, gvd) and O[i] == 1.0
            continue
        ovd, odata, oO = _oracle_compressed(L, gvd, gdata, D, tvd, tdata)
        assert np.array_equal(vd, ovd), (D, vd, ovd)
        assert abs(O[i] - oO) <= tol * oO
        # same state: dense vectors (the stored gauge differs)
        a, b = M.to_dense(vd, data), M.to_dense(ovd, odata)
        assert np.allclose(a, b, atol=(1e-10 if precision == "fp64" else 1e-5))
    # the compressed population runs: energies at the start equal the oracle's
    E = eng.prepare()
    orc = oracle_population(s, tvd, tdata)
    vds, datas, Os = zip(*[_oracle_compressed(L, gvd, gdata, D, tvd, tdata) if D >= 1 else (gvd, gdata, 1.0)
                           for D in targets])
    orc.set_walkers(list(vds), list(datas), np.ones(n), np.array(Os), np.zeros(n))
    EP, _, EL = orc.energy(1, False)
    assert abs(E - EP) <= tol * 10 * abs(EP)
    assert rel(eng.scalars()["EL"], EL) < tol * 100
    eng.close()


def test_compress_tile_layout_fp32():
    # the FP32 tile path keeps 32 x 32 zero-padded blocks: a smaller bond must clear the padding
    L = 16
    s = systems.chain(L)
    tvd, tdata = random_trial(L, 32, 5)
    gvd, gdata = random_gauge(tvd, tdata, 6)
    targets = [8, 16, 12, 0]
    eng = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=32, precision="fp32", seed=2)
    assert eng.path == "tile32"
    eng.load_copies(gvd, gdata, 4)
    eng.compress_walkers(targets)
    O = eng.scalars()["O"]
    for i, D in enumerate(targets[:3]):
        vd, _ = eng.walker(i)
        ovd, _, oO = _oracle_compressed(L, gvd, gdata, D, tvd, tdata)
        assert np.array_equal(vd, ovd)
        assert abs(O[i] - oO) <= 1e-5 * oO
    E = eng.prepare()
    orc = oracle_population(s, tvd, tdata)
    vds, datas, Os = zip(*[_oracle_compressed(L, gvd, gdata, D, tvd, tdata) for D in targets[:3]])
    orc.set_walkers(list(vds) + [gvd], list(datas) + [gdata], np.ones(4), np.array(list(Os) + [1.0]), np.zeros(4))
    assert abs(E - orc.energy(1, False)[0]) <= 2e-4 * abs(E)
    eng.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-4)])
def test_snake4x4_uneven_walkers_from_compression(precision, tol):
    # the uneven snake workload at its stated bonds: trial D_T = 64 on the 4x4
    # open lattice, walkers = the trial compressed on the device to D_W in
    # {16, 32, 48} and uncompressed (D_W = 64 = D_T, SetupWalkers does not
    # compress), then two propagation steps against the oracle on the same uniforms
    s = systems.snake_lattice(4, 4)
    L = s.L
    tvd, tdata = random_trial(L, 64, 7)
    gvd, gdata = random_gauge(tvd, tdata, 8)
    dws = [16, 32, 48, 64, 16, 32, 48, 64]
    targets = [D if D < 64 else 0 for D in dws]
    n = len(dws)
    eng = Engine(s, tvd, tdata, n_desired=n, capacity=4 * n, max_bond=64, precision=precision, seed=3)
    assert eng.path == "general"
    eng.load_copies(gvd, gdata, n)
    eng.compress_walkers(targets)
    ref = {D: _oracle_compressed(L, gvd, gdata, D, tvd, tdata) for D in (16, 32, 48)}
    ref[64] = (gvd, gdata, 1.0)
    O = eng.scalars()["O"]
    for i, D in enumerate(dws):
        vd, _ = eng.walker(i)
        assert np.array_equal(vd, ref[D][0]), D
        assert abs(O[i] - ref[D][2]) <= (tol if precision == "fp64" else 1e-5) * ref[D][2]
    orc = oracle_population(s, tvd, tdata, threads=THREADS)
    orc.set_walkers([ref[D][0] for D in dws], [ref[D][1] for D in dws], np.ones(n), np.array([ref[D][2] for D in dws]),
                    np.zeros(n))
    E_e = eng.prepare()
    E_o = orc.energy(1, False)[0]
    assert abs(E_e - E_o) <= tol * abs(E_o)
    rng = Random(11)
    for step in (1, 2):
        nw = orc.nwalkers
        u = rng.uniforms(nw * orc.nC)
        sw_o, idx_o = orc.propagate(u)
        sw_e = eng.propagate(u)
        same = np.all(eng.indices(nw) == idx_o, axis=1)
        if precision == "fp64":
            assert same.all()
        else:
            assert np.mean(same) >= 0.75
        wo, Oo, _ = orc.scalars()
        se = eng.scalars()
        assert rel(se["w"][same], wo[same]) < tol and rel(se["O"][same], Oo[same]) < tol
        EP_o, _, EL_o = orc.energy(step, True)
        EP_e, _ = eng.energy(step, True)
        assert rel(se["EL"][same], EL_o[same]) < 10 * tol
        if not same.all():
            break
        assert abs(EP_e - EP_o) <= 10 * tol * abs(EP_o)
    eng.close()
