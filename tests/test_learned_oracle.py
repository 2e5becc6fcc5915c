"""The learned proposal's checkers on the CPU: the FP64 network restatement
against finite differences, and the oracle's importance-sampled propagation
against the reference proposal (exactness for exact labels, unbiasedness of
the weights for arbitrary ones)."""
import numpy as np
import pytest

from mpsqmc_amd import systems
from oracle import mlp as omlp

from .common import oracle_population, random_trial


def test_oracle_mlp_gradient_matches_finite_differences():
    nin, H, nout = 7, 12, 16
    rng = np.random.default_rng(0)
    p = rng.normal(size=omlp.num_params(nin, H, nout)) * 0.4
    X = rng.normal(size=(24, nin))
    Y = rng.normal(size=(24, nout))
    sc = (rng.normal(size=nin), rng.uniform(0.5, 2, nin), rng.normal(size=nout), rng.uniform(0.5, 2, nout))
    loss, g = omlp.loss_grad(p, X, Y, nin, H, nout, sc)
    assert loss == pytest.approx(omlp.loss_only(p, X, Y, nin, H, nout, sc), rel=1e-13)
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (omlp.loss_only(p + e, X, Y, nin, H, nout, sc) - omlp.loss_only(p - e, X, Y, nin, H, nout, sc)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_oracle_mlp_pack_roundtrip_and_layout():
    nin, H, nout = 3, 4, 2
    p = np.arange(omlp.num_params(nin, H, nout), dtype=float)
    W1, b1, W2, b2, W3, b3 = omlp.unpack(p, nin, H, nout)
    assert W1[1, 0] == 1.0 and W1[0, 1] == H          # column-major W1 [H x n_in]
    assert b1[0] == H * nin
    assert np.array_equal(omlp.pack(W1, b1, W2, b2, W3, b3), p)


def test_proposal_is_a_distribution_with_floor():
    grid = systems.chain(4).grid()
    rng = np.random.default_rng(1)
    q = omlp.proposal(rng.normal(size=16), grid, 0.2)
    assert q.sum() == pytest.approx(1.0)
    assert q.min() >= 0.2 / grid.shape[0] - 1e-15
    assert np.allclose(omlp.proposal(-np.eye(4).T.reshape(-1), grid, 0.2), 1.0 / grid.shape[0])


def _population(L, D, n, seed):
    s = systems.chain(L)
    tvd, tdata = random_trial(L, D, seed)
    vds, datas = zip(*[random_trial(L, D, seed * 100 + i) for i in range(n)])
    rng = np.random.default_rng(seed)
    w, O, eh = rng.uniform(0.5, 1.5, n), np.ones(n), np.zeros(n)

    def make():
        orc = oracle_population(s, tvd, tdata)
        orc.set_walkers(list(vds), list(datas), w, O, eh)
        return orc
    return make


def test_oracle_exact_labels_reproduce_the_reference_proposal():
    # with the exact labels and mix = 0 the proposal is the reference's PDF, so
    # the samples and weights are the reference's
    make = _population(8, 4, 12, 3)
    a = make()
    u = np.random.default_rng(4).uniform(size=a.nwalkers * a.nC)
    sw_a, idx_a, lab = a.propagate_with_labels(u)
    b = make()
    b.set_proposal(lab, mix=0.0)
    sw_b, idx_b = b.propagate(u)
    assert np.array_equal(idx_a, idx_b)
    assert sw_b == pytest.approx(sw_a, rel=1e-12)
    assert np.allclose(b.scalars()[0], a.scalars()[0], rtol=1e-12)


@pytest.mark.parametrize("mix", [0.5, 1.0])
def test_oracle_learned_weight_has_the_reference_expectation(mix):
    # one coupling (L = 2): over u the sampled point has law q, so the mean of
    # pdf(x)/q(x) is the reference's PDF norm; stratified u makes it exact to
    # the stratification error
    s = systems.chain(2)
    tvd, tdata = random_trial(2, 2, 5)
    ref = oracle_population(s, tvd, tdata)
    ref.init_walkers(1, 2)
    ref.propagate(np.array([0.5]))
    pdfnorm = ref.scalars()[0][0]
    K = 50000
    orc = oracle_population(s, tvd, tdata)
    orc.init_walkers(K, 2)
    orc.set_proposal(np.broadcast_to(np.random.default_rng(1).normal(size=16), (K, 1, 16)), mix=mix)
    orc.propagate((np.arange(K) + 0.5) / K)
    assert orc.scalars()[0].mean() == pytest.approx(pdfnorm, rel=1e-3)


def test_oracle_learned_weights_are_unbiased_over_a_step():
    # a whole step (5 couplings) under a perturbed proposal: same mean weight
    # as the reference's proposal within statistical error; larger variance
    L, D, n = 6, 4, 4000
    s = systems.chain(L)
    tvd, tdata = random_trial(L, D, 5)
    rng = np.random.default_rng(2)
    a = oracle_population(s, tvd, tdata)
    a.init_walkers(8, D)
    _, _, lab = a.propagate_with_labels(rng.uniform(size=8 * a.nC))
    pred = lab[0] * (1 + 0.2 * rng.normal(size=lab[0].shape))

    def run(p):
        orc = oracle_population(s, tvd, tdata)
        orc.init_walkers(n, D)
        if p is not None:
            orc.set_proposal(np.broadcast_to(p, (n, orc.nC, 16)), mix=0.2)
        orc.propagate(rng.uniform(size=n * orc.nC))
        return orc.scalars()[0]

    w_ref, w_l = run(None), run(pred)
    se = np.sqrt(w_ref.var() / n + w_l.var() / n)
    assert abs(w_l.mean() - w_ref.mean()) < 4 * se
    assert w_l.var() > w_ref.var()  # the reference proposal is the zero-variance one per coupling
