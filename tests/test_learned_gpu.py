"""The learned proposal on the GPU: the network's forward, loss and gradient
kernels against the FP64 restatement, training, the engine's training rows
(features and exact labels) and the importance-sampled propagation against the
FP64 oracle on the same uniforms."""
import numpy as np
import pytest

from mpsqmc_amd import systems
from mpsqmc_amd.engine import Engine
from mpsqmc_amd.mlp import MLP, NFEAT, NLABEL
from oracle import mlp as omlp

from .common import oracle_population, random_trial

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

H = 128


def _net_with_random_params(seed):
    net = MLP(NFEAT, H, NLABEL, seed=seed)
    rng = np.random.default_rng(seed)
    p = (rng.normal(size=net.num_params) * 0.15).astype(np.float32)
    net.set_params(p)
    sc = (rng.normal(size=NFEAT), rng.uniform(0.5, 2, NFEAT), rng.normal(size=NLABEL), rng.uniform(0.5, 2, NLABEL))
    sc = tuple(np.asarray(a, np.float32) for a in sc)
    net.set_scaling(*sc)
    return net, p.astype(np.float64), tuple(a.astype(np.float64) for a in sc)


def test_mlp_forward_matches_fp64_restatement():
    net, p, sc = _net_with_random_params(3)
    X = np.random.default_rng(4).normal(size=(300, NFEAT)).astype(np.float32)  # not a multiple of the 64-row tile
    got = net.forward(X)
    want = omlp.forward(p, X.astype(np.float64), NFEAT, H, NLABEL, sc)
    # FP32 products of length <= 128: relative 1e-4 of the output scale
    assert np.max(np.abs(got - want)) <= 1e-4 * np.max(np.abs(want))
    net.close()


def test_mlp_gradient_matches_fp64_restatement():
    net, p, sc = _net_with_random_params(5)
    rng = np.random.default_rng(6)
    X = rng.normal(size=(200, NFEAT)).astype(np.float32)
    Y = rng.normal(size=(200, NLABEL)).astype(np.float32)
    loss, g = net.loss_grad(X, Y)
    loss_o, g_o = omlp.loss_grad(p, X.astype(np.float64), Y.astype(np.float64), NFEAT, H, NLABEL, sc)
    assert loss == pytest.approx(loss_o, rel=1e-5)
    assert np.max(np.abs(g - g_o)) <= 1e-4 * np.max(np.abs(g_o))
    net.close()


def test_mlp_fit_learns_a_smooth_map():
    rng = np.random.default_rng(7)
    nin = 7
    A = rng.normal(size=(nin, 24))
    B = rng.normal(size=(24, NLABEL)) / 5
    X = rng.normal(size=(20000, nin)).astype(np.float32)
    Y = (np.tanh(X @ A) @ B).astype(np.float32)
    net = MLP(nin, H, NLABEL, seed=8)
    loss = net.fit(X, Y, epochs=30, batch=512, lr=2e-3, standardize=True)
    # standardised targets have unit variance: the fit explains > 95 % of it
    assert loss < 0.05
    Xt = rng.normal(size=(500, nin)).astype(np.float32)
    Yt = np.tanh(Xt @ A) @ B
    assert np.mean((net.forward(Xt) - Yt) ** 2) < 0.1 * np.var(Yt)
    net.close()


def _tile_pair(L, D, n, seed):
    s = systems.chain(L)
    tvd, tdata = random_trial(L, D, seed)
    vds, datas = zip(*[random_trial(L, D, seed * 100 + i) for i in range(n)])
    rng = np.random.default_rng(seed)
    w, O, eh = rng.uniform(0.5, 1.5, n), rng.uniform(0.5, 1.5, n), np.zeros(n)

    def oracle_pop():
        orc = oracle_population(s, tvd, tdata)
        orc.set_walkers(list(vds), list(datas), w, O, eh)
        return orc
    eng = Engine(s, tvd, tdata, n_desired=n, capacity=4 * n, max_bond=D, precision="fp32", seed=seed)
    assert eng.path == "tile32"
    eng.load_walkers(list(vds), list(datas), w, O, eh)
    return eng, oracle_pop


def _agree_prefix(idx_a, idx_b):
    """mask[b, c]: walker b sampled the same points at couplings 0..c-1"""
    same = idx_a == idx_b
    out = np.ones_like(same)
    out[:, 1:] = np.cumprod(same[:, :-1], axis=1).astype(bool)
    return out


def test_engine_training_rows_match_oracle():
    eng, oracle_pop = _tile_pair(10, 8, 16, 21)
    eng.prepare()
    sc = eng.scalars()
    be = eng.bond_energies()
    eng.collect_samples(10000)
    orc = oracle_pop()
    u = np.random.default_rng(22).uniform(size=16 * eng.nC)
    eng.propagate(u)
    _, idx_o, lab_o = orc.propagate_with_labels(u)
    idx_e = eng.indices(16)
    assert eng.num_samples() == 16 * eng.nC
    X, Y = eng.read_samples()
    # labels: combos/O of every coupling reached with the same samples (FP32 contractions)
    m = _agree_prefix(idx_e, idx_o).reshape(-1)
    assert m.mean() > 0.9
    lab = lab_o.reshape(-1, 16)
    assert np.max(np.abs(Y[m] - lab[m])) <= 2e-4 * np.max(np.abs(lab))
    # features: start-of-step scalars and the bond energies of the preparation sweep
    nC = eng.nC
    F = X.reshape(16, nC, NFEAT)
    assert np.allclose(F[:, 0, 0], np.log(sc["O"]), rtol=1e-6, atol=1e-6)
    assert np.allclose(F[:, 0, 1], np.log(sc["w"]), rtol=1e-6, atol=1e-6)
    assert np.allclose(F[:, 0, 2], sc["EL"], rtol=1e-6, atol=1e-6)
    assert np.allclose(F[:, :, 4], be, rtol=1e-5, atol=1e-6)
    assert np.allclose(F[:, 1:, 3], be[:, :-1], rtol=1e-5, atol=1e-6) and np.all(F[:, 0, 3] == 0)
    assert np.allclose(F[:, :-1, 5], be[:, 1:], rtol=1e-5, atol=1e-6) and np.all(F[:, -1, 5] == 0)
    assert np.allclose(F[0, :, 6], np.arange(nC) / (nC - 1))
    assert np.all(F[:, :, 7:] == 0)  # no recorded labels before the first collected step
    # a second step: its features carry the labels each walker recorded in the first
    eng.propagate(np.random.default_rng(23).uniform(size=16 * nC))
    X2, _ = eng.read_samples(16 * nC, 16 * nC)
    assert np.array_equal(X2.reshape(16, nC, NFEAT)[:, :, 7:], Y.reshape(16, nC, 16))
    eng.close()


def test_bond_energies_sum_to_the_local_energy():
    eng, _ = _tile_pair(12, 8, 8, 23)
    eng.prepare()
    eng.propagate()
    eng.energy(1, True)
    be = eng.bond_energies()
    assert np.allclose(be.sum(1), eng.scalars()["EL"], rtol=1e-10, atol=1e-12)
    eng.close()


def test_engine_learned_proposal_matches_oracle():
    n = 16
    eng, oracle_pop = _tile_pair(10, 8, n, 25)
    eng.prepare()
    rng = np.random.default_rng(26)
    u = rng.uniform(size=n * eng.nC)
    _, _, lab = oracle_pop().propagate_with_labels(u)
    pred = (lab * (1 + 0.1 * rng.normal(size=lab.shape))).astype(np.float32)
    eng.set_proposal_table(pred, mix=0.2)
    orc = oracle_pop()
    orc.set_proposal(pred.astype(np.float64), mix=0.2)
    sw_e = eng.propagate(u)
    sw_o, idx_o = orc.propagate(u)
    idx_e = eng.indices(n)
    same = np.all(idx_e == idx_o, axis=1)
    assert same.mean() > 0.8
    w_e, w_o = eng.scalars()["w"], orc.scalars()[0]
    # FP32 contractions: 2e-3 relative on the weights of walkers with identical
    # samples (a sample where the exact PDF vanishes gives weight 0 on both sides)
    assert np.all(np.abs(w_e[same] - w_o[same]) <= 2e-3 * np.abs(w_o[same]))
    if same.all():
        assert sw_e == pytest.approx(sw_o, rel=2e-3)
    eng.close()


def test_collect_train_and_run_learned():
    L, D, n = 16, 16, 256
    s = systems.chain(L)
    tvd, tdata = random_trial(L, D, 31)
    eng = Engine(s, tvd, tdata, n_desired=n, capacity=3 * n, max_bond=D, precision="fp32", seed=31)
    eng.load_copies(tvd, tdata, n)
    eng.prepare()
    eng.collect_samples(8 * n * eng.nC)
    eng.run(1, 8)
    assert eng.num_samples() > 0
    net = MLP(NFEAT, H, NLABEL, seed=32)
    loss = eng.train_proposal(net, epochs=4, batch=1024, lr=1e-3)
    assert np.isfinite(loss) and loss < 1.0  # better than predicting the mean
    eng.collect_samples(0)
    eng.set_proposal(net, mix=0.1)
    eng.run(9, 8)
    tr = eng.read_trace(9, 8)
    assert np.all(np.isfinite(tr)) and tr[-1, 1] > 0
    eng.set_proposal(None)
    eng.close()
    net.close()


def test_learned_api_errors():
    from mpsqmc_amd._lib import MqeError
    s = systems.chain(8)
    tvd, tdata = random_trial(8, 4, 41)
    gen = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=4, precision="fp64")
    assert gen.path == "general"
    with pytest.raises(MqeError):
        gen.collect_samples(100)
    gen.close()
    eng = Engine(s, tvd, tdata, n_desired=4, capacity=16, max_bond=4, precision="fp32")
    eng.load_copies(tvd, tdata, 4)
    eng.prepare()
    wrong = MLP(5, 16, NLABEL, seed=1)
    with pytest.raises(MqeError):
        eng.set_proposal(wrong, 0.1)          # input width must be MQE_NFEAT
    net = MLP(NFEAT, 16, NLABEL, seed=1)
    with pytest.raises(MqeError):
        eng.set_proposal(net, 1.5)            # mix outside [0, 1]
    with pytest.raises(MqeError):
        eng.train_proposal(net)               # nothing collected
    eng.close()
    wrong.close()
    net.close()
