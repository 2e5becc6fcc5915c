"""Pins of the FP64 CPU oracle against known answers, closed forms, brute force
and the reference's own golden numbers (main.cpp:111,153,199; the .mps
fixtures).  Each test names the oracle function it pins."""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.special

import oracle
from mpsqmc_amd import mps as M
from mpsqmc_amd import systems

from .common import dense_hamiltonian, dense_to_mps, ground_energy, oracle_population, random_trial, spin_ops

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# --- mt_next / mt_uniform -----------------------------------------------------
def test_mt19937_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937
    assert int(oracle.mt_raw(5489, 10000)[-1]) == 4123659995


def test_mt19937_matches_numpy_legacy_seeding():
    # numpy's MT19937 with init_genrand seeding is an independent implementation
    bg = np.random.MT19937(0)
    bg._legacy_seeding(20131015)
    assert np.array_equal(bg.random_raw(2000).astype(np.uint32), oracle.mt_raw(20131015, 2000))
    u = oracle.mt_uniforms(20131015, 3)
    assert np.array_equal(u, oracle.mt_raw(20131015, 3) / 2.0 ** 32)


# --- op_elem ------------------------------------------------------------------
@pytest.mark.parametrize("d", [2, 3, 4])
def test_operators_spin_algebra(d):
    sz = oracle.operator(2, d)
    sx = oracle.operator(3, d)
    isy = oracle.operator(4, d)
    sy = -1j * isy
    s = (d - 1) / 2
    assert np.allclose(sx @ sy - sy @ sx, 1j * sz)
    assert np.allclose(sy @ sz - sz @ sy, 1j * sx)
    cas = sx @ sx + sy @ sy + sz @ sz
    assert np.allclose(cas, s * (s + 1) * np.eye(d))
    assert np.allclose(np.diag(sz), [-s + i for i in range(d)])  # smallest m first (OpSz.cpp:9)
    ex, ey, ez = spin_ops(d)
    assert np.allclose(sx, ex) and np.allclose(sy, ey) and np.allclose(sz, ez)


# --- orc_trotter_tables -----------------------------------------------------------
@pytest.mark.parametrize("J,dtau", [(1.0, 0.01), (0.5, 0.02), (-1.0, 0.01), (2.0, 0.1)])
def test_trotter_tables_closed_form(J, dtau):
    S, U, VT, P = oracle.trotter_tables(2, J, dtau)
    t = J * dtau
    a = (np.exp(3 * t / 4) + 3 * np.exp(-t / 4)) / 4
    b = np.exp(-t / 4) - np.exp(3 * t / 4)
    want = sorted([2 * abs(a), abs(b) / 2, abs(b) / 2, abs(b) / 2], reverse=True)
    assert np.allclose(S, want, rtol=1e-13, atol=1e-15)
    # P is exp(-dtau J S.S) reshuffled to [(r1 c1), (r2 c2)]: build it independently
    sx, sy, sz = spin_ops(2)
    SS = np.real(np.kron(sx, sx) + np.kron(sy, sy) + np.kron(sz, sz))  # index r1*2 + r2 (kron order)
    Pd = sla.expm(-t * SS)
    Pr = np.zeros((4, 4))
    for r1 in range(2):
        for r2 in range(2):
            for c1 in range(2):
                for c2 in range(2):
                    Pr[r1 + 2 * c1, r2 + 2 * c2] = Pd[r1 * 2 + r2, c1 * 2 + c2]
    P = P.reshape(4, 4, order="F")
    assert np.allclose(P, Pr, atol=1e-14)
    # sum_k s_k U_k (x) V_k reconstructs P
    Ur = U.reshape(4, 4, order="F")
    Vr = VT.reshape(4, 4, order="F")
    assert np.allclose(Ur @ np.diag(S) @ Vr, P, atol=1e-14)
    assert np.allclose(Ur.T @ Ur, np.eye(4), atol=1e-13)
    assert np.allclose(Vr @ Vr.T, np.eye(4), atol=1e-13)
    # eigenvalues of S.S (TrotterHeisenberg.cpp:145-150)
    assert np.allclose(np.linalg.eigvalsh(SS), [-0.75, 0.25, 0.25, 0.25])


# --- erfinv_own / orc_marsaglia --------------------------------------------------
def test_erfinv_against_scipy():
    # the grid uses y = +-0.25, +-0.75 (GridGenerator.cpp:175-179): last-ulp agreement there
    for y in [-0.75, -0.25, 0.25, 0.75, 0.1, 1e-9]:
        assert abs(oracle.erfinv(y) - scipy.special.erfinv(y)) <= 2.5e-16 * max(1.0, abs(scipy.special.erfinv(y)))
    # near |y| = 1 erfinv is ill-conditioned; agreement to the conditioning of erf
    for y in [0.9, -0.999]:
        assert abs(oracle.erfinv(y) - scipy.special.erfinv(y)) <= 1e-13 * abs(scipy.special.erfinv(y))


def test_marsaglia_grid_pins():
    g, N = oracle.marsaglia(4, 4)
    G = g.reshape(N, 4)
    assert N == 240
    assert np.allclose(np.linalg.norm(G, axis=1), 1.0, atol=1e-15)
    assert np.allclose(G.T @ G, 60.0 * np.eye(4), atol=1e-13)
    # first point: all coordinates equal to erfinv(-0.75), normalised
    assert np.allclose(G[0], -0.5)
    vals = np.unique(np.round(np.abs(G[G != 0]), 12))
    assert vals.size > 0
    # no two points closer than the de-duplication threshold
    d2 = ((G[:, None, :] - G[None, :, :]) ** 2).sum(-1) + np.eye(N)
    assert d2.min() >= 1e-16


def test_marsaglia_p5_count():
    g, N = oracle.marsaglia(4, 5)
    G = g.reshape(N, 4)
    assert np.allclose(np.linalg.norm(G, axis=1), 1.0)
    # the middle value erfinv(0) = 0 enters; (0,0,0,0) has zero norm and is skipped
    assert N < 5 ** 4


# --- inner ----------------------------------------------------------------------
@pytest.mark.parametrize("L,Da,Db", [(6, 3, 5), (8, 4, 4), (9, 8, 2)])
def test_inner_vs_dense(L, Da, Db):
    va, a = M.random_mps(L, Da, 2, 1)
    vb, b = M.random_mps(L, Db, 2, 2)
    want = M.to_dense(va, a) @ M.to_dense(vb, b)
    got = oracle.inner(L, 2, va, a, vb, b)
    assert abs(got - want) <= 1e-12 * abs(want)
    assert abs(oracle.inner(L, 2, vb, b, va, a) - want) <= 1e-12 * abs(want)


# --- householder_qr / left_normalize --------------------------------------------
def test_qr():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((12, 7))
    Q, R = oracle.qr(A)
    assert np.allclose(Q.T @ Q, np.eye(7), atol=1e-14)
    assert np.allclose(np.triu(R), R)
    assert np.allclose(Q @ R, A, atol=1e-13)
    # rank-deficient input
    A[:, 3] = A[:, 1]
    Q, R = oracle.qr(A)
    assert np.allclose(Q.T @ Q, np.eye(7), atol=1e-13)
    assert np.allclose(Q @ R, A, atol=1e-13)


@pytest.mark.parametrize("L,D", [(6, 4), (10, 8)])
def test_left_normalize(L, D):
    vd, data = M.random_mps(L, D, 2, 7)
    psi = M.to_dense(vd, data)
    r, out = oracle.left_normalize(L, 2, vd, data)
    assert abs(abs(r) - np.linalg.norm(psi)) <= 1e-12 * np.linalg.norm(psi)
    phi = M.to_dense(vd, out)
    assert np.allclose(phi * r, psi, atol=1e-12 * np.abs(psi).max())
    for k in range(L):
        A = M.site_tensor(vd, out, k)
        G = sum(A[s].T @ A[s] for s in range(2))
        assert np.allclose(G, np.eye(G.shape[0]), atol=1e-13)


# --- jacobi_svd / compress_state --------------------------------------------------
@pytest.mark.parametrize("m,n", [(8, 5), (5, 8), (6, 6)])
def test_svd(m, n):
    rng = np.random.default_rng(m * 10 + n)
    A = rng.standard_normal((m, n))
    U, S, VT = oracle.svd(A)
    assert np.allclose(U @ np.diag(S) @ VT, A, atol=1e-13)
    assert np.allclose(S, np.linalg.svd(A, compute_uv=False), atol=1e-13)
    assert np.all(np.diff(S) <= 0)


def test_compress_lossless_and_truncating():
    L = 8
    vd, data = M.random_mps(L, 4, 2, 11)
    psi = M.to_dense(vd, data)
    vo, out, D = oracle.compress(L, 2, vd, data, 16)
    assert np.allclose(M.to_dense(vo, out), psi, atol=1e-12 * np.abs(psi).max())
    # truncation to 2: fidelity below 1, state right-canonical with norm in site 0
    vo2, out2, D2 = oracle.compress(L, 2, vd, data, 2)
    assert D2 == 2 and vo2.max() == 2
    for k in range(1, L):
        A = M.site_tensor(vo2, out2, k)
        G = sum(A[s] @ A[s].T for s in range(2))
        assert np.allclose(G, np.eye(G.shape[0]), atol=1e-12)
    phi = M.to_dense(vo2, out2)
    f = (phi @ psi) / (np.linalg.norm(phi) * np.linalg.norm(psi))
    assert 0.5 < f < 1.0


# --- apply_mpo / energies against the reference's golden numbers -------------------
def test_mpo_variational_energy_vs_dense():
    for L in (5, 6, 7):
        sysm = systems.chain(L)
        sysm.J[systems.packed_index(0, L - 1)] = 0.3   # long-range term too
        sysm.h = 0.2
        vd, data = M.random_mps(L, 4, 2, L)
        vo, H = oracle.apply_mpo(L, 2, sysm.J, sysm.h, vd, data)
        got = oracle.inner(L, 2, vd, data, vo, H)
        psi = M.to_dense(vd, data)
        want = psi @ (dense_hamiltonian(sysm) @ psi)
        assert abs(got - np.real(want)) <= 1e-11 * abs(want)


def test_fci_l10_chain_main_cpp():
    # main.cpp:153, E("FCI") = -4.25803520728288 for the L = 10 spin-1/2 chain
    sysm = systems.chain(10)
    assert abs(ground_energy(sysm) - (-4.25803520728288)) < 1e-12
    vd, data = dense_to_mps(_ground_vector(sysm), 10, 32)
    pop = oracle_population(sysm, vd, data, build_terms=False)
    pop.init_walkers(2, 32)
    EP, om, EL = pop.energy(1, dofluct=False)
    assert abs(EP - (-4.25803520728288)) < 1e-11


def test_fci_square_pbc_main_cpp():
    # main.cpp:111, E("FCI") = -11.2284832084289 for the 4x4 periodic J1 = 1 lattice
    assert abs(ground_energy(systems.square_periodic(4)) - (-11.2284832084289)) < 1e-10


def test_fci_spin1_l10_main_cpp():
    # main.cpp:199, E(FCI) = -12.8945601322109 for the L = 10 spin-1 chain
    sysm = systems.chain(10)
    assert abs(ground_energy(sysm, d=3) - (-12.8945601322109)) < 1e-10
    # the oracle's spin-1 MPO reproduces <psi|H|psi> of the dense Hamiltonian
    vd, data = M.random_mps(6, 3, 3, 5)
    s6 = systems.chain(6)
    vo, H = oracle.apply_mpo(6, 3, s6.J, 0.0, vd, data)
    got = oracle.inner(6, 3, vd, data, vo, H)
    psi = M.to_dense(vd, data, d=3)
    want = np.real(psi @ (dense_hamiltonian(s6, d=3) @ psi))
    assert abs(got - want) <= 1e-11 * abs(want)


def _ground_vector(sysm):
    import scipy.sparse.linalg as spla
    H = dense_hamiltonian(sysm)
    w, V = spla.eigsh(H, k=1, which="SA", tol=1e-13)
    v = np.real(V[:, 0])
    return v / np.linalg.norm(v)


def test_debug_mps_fixture():
    # the reference's own trial file (debug.mps, L = 20, D = 4) round-trips and
    # its variational energy is reproduced by ApplyMPO + InnerProduct
    path = os.path.join(GOLDEN, "debug.mps")
    L, Dt, d, vd, data = M.read_mps(path)
    assert (L, Dt, d) == (20, 4, 2)
    r, nd = oracle.left_normalize(L, 2, vd, data)
    assert abs(abs(r) - 1.0) < 1e-12   # stored normalised
    sysm = systems.chain(20)
    vo, H = oracle.apply_mpo(L, 2, sysm.J, 0.0, vd, data)
    E = oracle.inner(L, 2, vd, data, vo, H)
    assert abs(E - (-8.654625014038682)) < 1e-10
    assert E > -8.682473334398976  # variational: above the L = 20 ground state


@pytest.mark.parametrize("name,E_var", [("L50D4.mps", -21.863425804367605), ("L50D8.mps", -21.966322890532627)])
def test_l50_trial_fixtures(name, E_var):
    # the reference's L = 50 trials (input/Heisenberg1D/L50D{4,8}.mps): stored
    # normalised; ApplyMPO + InnerProduct give their variational energies with
    # the reference's MPO and with the minimal one
    L, Dt, d, vd, data = M.read_mps(os.path.join(GOLDEN, name))
    assert L == 50 and d == 2
    r, nd = oracle.left_normalize(L, 2, vd, data)
    assert abs(abs(r) - 1.0) < 1e-10
    sysm = systems.chain(L)
    vo, H = oracle.apply_mpo(L, 2, sysm.J, 0.0, vd, data)
    E = oracle.inner(L, 2, vd, data, vo, H)
    assert abs(E - E_var) < 1e-10
    v1, h1 = oracle.apply_mpo_min(L, 2, sysm.J, 0.0, vd, data)
    assert abs(oracle.inner(L, 2, vd, data, v1, h1) - E_var) < 1e-10


# --- population control ---------------------------------------------------------
def test_population_control_counts():
    sysm = systems.chain(6)
    vd, data = random_trial(6, 2, 1)
    pop = oracle_population(sysm, vd, data, build_terms=False)
    w = np.array([0.1, 1.0, 2.6, 0.2, 1.49, 3.0])
    pop.set_walkers([vd] * 6, [data] * 6, w, np.ones(6), np.arange(6.0))
    u = np.array([0.95, 0.3, 0.5, 0.7])
    n, used, nc = pop.popcontrol(1.0, u, 100)
    # s<0.25: floor(s+u) copies, weight 1; s>1.5: floor(s+u) copies, weight s/n
    assert list(nc) == [int(0.1 + 0.95), 1, int(2.6 + 0.3), int(0.2 + 0.5), 1, int(3.0 + 0.7)]
    assert used == 4 and n == sum(nc)
    wn, On, ehn = pop.scalars()
    exp_w, exp_eh = [], []
    for wi, c, e in zip(w, nc, range(6)):
        nw = 1.0 if wi < 0.25 else (wi / c if wi > 1.5 else wi)
        exp_w += [nw] * c
        exp_eh += [e] * c
    assert np.allclose(wn, exp_w) and np.allclose(ehn, exp_eh)


def test_population_control_unbiased():
    sysm = systems.chain(4)
    vd, data = random_trial(4, 2, 1)
    pop = oracle_population(sysm, vd, data, build_terms=False)
    rng = np.random.default_rng(5)
    w = rng.uniform(0.05, 3.0, 40)
    tot = 0.0
    reps = 400
    for r in range(reps):
        pop.set_walkers([vd] * 40, [data] * 40, w, np.ones(40), np.zeros(40))
        n, used, nc = pop.popcontrol(1.0, rng.uniform(size=40), 1000)
        wn, _, _ = pop.scalars()
        tot += wn.sum()
    assert abs(tot / reps - w.sum()) < 0.03 * w.sum()


# --- the propagation step against a brute-force dense restatement ------------------
def _dense_ops(tabs_S, tabs_U, tabs_VT, x):
    Ux = np.zeros((2, 2))
    Vx = np.zeros((2, 2))
    U = tabs_U.reshape(2, 2, 4, order="F")     # U[i, j, k]
    V = tabs_VT.reshape(4, 2, 2, order="F")    # VT[k, i, j]
    for k in range(4):
        pf = np.sqrt(tabs_S[k] * 4) * x[k]
        Ux = Ux + pf * U[:, :, k]
        Vx = Vx + pf * V[k]
    return Ux, Vx


def _op_on(L, i, O):
    out = np.array([[1.0]])
    for q in range(L):
        out = np.kron(out, O if q == i else np.eye(2))
    return out


def test_propagate_matches_dense_brute_force():
    L = 6
    sysm = systems.chain(L)
    tvd, tdata = random_trial(L, 4, 21)
    trial = M.to_dense(tvd, tdata)
    pop = oracle_population(sysm, tvd, tdata)
    wvd, wdata = random_trial(L, 4, 22)
    pop.set_walkers([wvd], [wdata], np.array([1.0]), np.array([1.0]), np.array([0.0]))
    u = np.random.default_rng(3).uniform(size=pop.nC)
    sw, idx = pop.propagate(u)
    # dense restatement
    grid = sysm.grid()
    S, U, VT = sysm.tables_per_coupling()
    f, s, _ = sysm.couplings()
    phi = M.to_dense(wvd, wdata)
    phi /= np.linalg.norm(phi)
    O = trial @ phi
    if O < 0:
        phi, O = -phi, -O
    w = 1.0
    for c in range(pop.nC):
        num = np.zeros(len(grid))
        for g, x in enumerate(grid):
            Ux, Vx = _dense_ops(S[c], U[c], VT[c], x)
            B = _op_on(L, f[c], Ux) @ _op_on(L, s[c], Vx)
            num[g] = trial @ (B @ phi)
        pdf = np.maximum(0.0, num) / (O * len(grid))
        norm = pdf.sum()
        w *= norm
        cum = np.cumsum(pdf / norm)
        gsel = min(int(np.sum(cum[:-1] <= u[c])), len(grid) - 1)
        assert gsel == idx[0, c]
        Ux, Vx = _dense_ops(S[c], U[c], VT[c], grid[gsel])
        phi = _op_on(L, f[c], Ux) @ (_op_on(L, s[c], Vx) @ phi)
        phi /= np.linalg.norm(phi)
        O = trial @ phi
        if O < 0:
            phi, O = -phi, -O
    wo, Oo, _ = pop.scalars()
    assert abs(wo[0] - w) <= 1e-11 * w
    assert abs(Oo[0] - O) <= 1e-11 * O
    vdo, do, *_ = pop.walker(0)
    assert np.allclose(M.to_dense(vdo, do), phi, atol=1e-11)


def test_propagate_with_field_matches_dense_brute_force():
    # h != 0: PropagateSeparately applies exp(dtau h S^z / 2) on every site before
    # and after the couplings, each followed by setOverlap (MPSQMC2.cpp:532-534,
    # 596-598; TrotterHeisenberg.cpp:80-88 for H = ... - h sum S^z).  The dense
    # restatement builds that factor with scipy's expm of the field term of the
    # dense Hamiltonian, and E_L = <trial|H|phi>/<trial|phi> with the dense H.
    import scipy.linalg as sl
    from .common import dense_hamiltonian
    L, h = 6, 0.7
    sysm = systems.chain(L, h=h)
    tvd, tdata = random_trial(L, 4, 23)
    trial = M.to_dense(tvd, tdata)
    pop = oracle_population(sysm, tvd, tdata)
    wvd, wdata = random_trial(L, 4, 24)
    pop.set_walkers([wvd], [wdata], np.array([1.0]), np.array([1.0]), np.array([0.0]))
    u = np.random.default_rng(5).uniform(size=pop.nC)
    sw, idx = pop.propagate(u)
    H = dense_hamiltonian(sysm, sparse=False)
    H0 = dense_hamiltonian(systems.chain(L), sparse=False)
    half = sl.expm(-0.5 * sysm.dtau * (H - H0))  # exp(dtau h/2 sum S^z)
    assert np.allclose(np.diag(np.diag(half)), half)
    grid = sysm.grid()
    S, U, VT = sysm.tables_per_coupling()
    f, s, _ = sysm.couplings()

    def normalise(v):
        v = v / np.linalg.norm(v)
        O = trial @ v
        return (-v, -O) if O < 0 else (v, O)

    phi, O = normalise(half @ M.to_dense(wvd, wdata))
    w = 1.0
    for c in range(pop.nC):
        num = np.zeros(len(grid))
        for g, x in enumerate(grid):
            Ux, Vx = _dense_ops(S[c], U[c], VT[c], x)
            num[g] = trial @ (_op_on(L, f[c], Ux) @ (_op_on(L, s[c], Vx) @ phi))
        pdf = np.maximum(0.0, num) / (O * len(grid))
        norm = pdf.sum()
        w *= norm
        cum = np.cumsum(pdf / norm)
        gsel = min(int(np.sum(cum[:-1] <= u[c])), len(grid) - 1)
        assert gsel == idx[0, c]
        Ux, Vx = _dense_ops(S[c], U[c], VT[c], grid[gsel])
        phi, O = normalise(_op_on(L, f[c], Ux) @ (_op_on(L, s[c], Vx) @ phi))
    phi, O = normalise(half @ phi)
    wo, Oo, _ = pop.scalars()
    assert abs(wo[0] - w) <= 1e-11 * w
    assert abs(Oo[0] - O) <= 1e-11 * O
    vdo, do, *_ = pop.walker(0)
    assert np.allclose(M.to_dense(vdo, do), phi, atol=1e-11)
    EP, _, EL = pop.energy(1, False)
    assert abs(EL[0] - (trial @ (H @ phi)) / O) <= 1e-10 * max(1.0, abs(EL[0]))


def test_energy_exact_trial_is_eigenvalue():
    # E_L == E0 for every walker when the trial is an exact eigenstate
    L = 8
    sysm = systems.chain(L)
    E0 = ground_energy(sysm)
    vd, data = dense_to_mps(_ground_vector(sysm), L, 16)
    pop = oracle_population(sysm, vd, data)
    pop.init_walkers(4, 16)
    E, trace = pop.walk(3, 7, 4, 100)
    assert abs(E - E0) < 1e-11
    assert np.allclose(trace[:, 2], E0, atol=1e-10)


def test_minimal_mpo_matches_reference_mpo():
    # same vector as ApplyMPO (HeisenbergMPO W tensors), smaller bonds
    for sysm in (systems.chain(7, h=0.3), systems.snake_lattice(3, 3)):
        L = sysm.L
        vd, data = M.random_mps(L, 3, 2, 4)
        v0, h0 = oracle.apply_mpo(L, 2, sysm.J, sysm.h, vd, data)
        v1, h1 = oracle.apply_mpo_min(L, 2, sysm.J, sysm.h, vd, data)
        assert np.allclose(M.to_dense(v0, h0), M.to_dense(v1, h1), atol=1e-12)
        assert v1.max() <= v0.max()
    v, _ = oracle.apply_mpo_min(32, 2, systems.chain(32).J, 0.0, *M.random_mps(32, 4, 2, 1))
    assert v.max() == 5 * 4   # rank of H|psi> across a chain cut: 5 D


def test_energy_modes_agree():
    L = 7
    sysm = systems.chain(L)
    sysm.J[systems.packed_index(1, 5)] = 0.4
    sysm.h = 0.3
    vd, data = random_trial(L, 4, 9)
    wv, wd = random_trial(L, 3, 10)
    pops = [oracle_population(sysm, vd, data, build_terms=False, energy_mode=m) for m in (0, 1, 2)]
    for p in pops:
        p.set_walkers([wv, vd], [wd, data], np.ones(2), np.array([0.7, 1.0]), np.zeros(2))
    ea, eb, ec = (p.energy(1, False)[2] for p in pops)
    assert np.allclose(ea, eb, rtol=1e-12)
    assert np.allclose(ea, ec, rtol=1e-12)
