"""Shared helpers of the test-suite: dense exact diagonalisation, trial
construction, and paired oracle / engine set-ups on identical inputs."""
import numpy as np

import oracle
from mpsqmc_amd import mps as M
from mpsqmc_amd import systems

def spin_ops(d):
    """Sx, Sy (complex), Sz for spin (d-1)/2, basis ordered smallest m first."""
    s = (d - 1) / 2.0
    m = np.array([-s + i for i in range(d)])
    sp = np.zeros((d, d))
    for i in range(d - 1):  # S+ |m> = sqrt(s(s+1) - m(m+1)) |m+1>
        sp[i + 1, i] = np.sqrt(s * (s + 1) - m[i] * (m[i] + 1))
    sm = sp.T
    return (sp + sm) / 2.0, (sp - sm) / 2.0j, np.diag(m)


def dense_hamiltonian(system, d=2, sparse=True):
    """H = sum_{i<j} J_ij S_i.S_j - h sum_i S^z_i on d^L states (site 0 most
    significant, the order of numpy.kron)."""
    import scipy.sparse as sp
    L = system.L
    sx, sy, sz = spin_ops(d)
    eye = sp.identity(d, format="csr")

    def op_at(o, k):
        out = None
        for q in range(L):
            f = sp.csr_matrix(o) if q == k else eye
            out = f if out is None else sp.kron(out, f, format="csr")
        return out

    H = sp.csr_matrix((d ** L, d ** L), dtype=complex)
    ops = [[op_at(o, k) for k in range(L)] for o in (sx, sy, sz)]
    f, s, Jc = system.couplings()
    for i, j, v in zip(f, s, Jc):
        for a in range(3):
            H = H + v * (ops[a][i] @ ops[a][j])
    if system.h != 0.0:
        for k in range(L):
            H = H - system.h * ops[2][k]
    H = H.real.tocsr() if np.abs(H.imag).max() < 1e-14 else H
    return H if sparse else H.toarray()


def ground_energy(system, d=2):
    import scipy.sparse.linalg as sla
    H = dense_hamiltonian(system, d)
    if H.shape[0] <= 64:
        return float(np.linalg.eigvalsh(H.toarray()).min())
    return float(sla.eigsh(H, k=1, which="SA", tol=1e-12)[0][0])


def random_trial(L, D, seed):
    vd, data = M.random_mps(L, D, 2, seed)
    data, _ = M.left_normalize(vd, data)
    return vd, data


def oracle_population(system, trial_vd, trial_data, threads=1, energy_mode=0, build_terms=True):
    t = system.tables_distinct()
    return oracle.Population(system.L, 2, trial_vd, trial_data, system.J, system.h, system.dtau, t,
                             system.grid().reshape(-1), 4, build_terms=build_terms, threads=threads,
                             energy_mode=energy_mode)


def rel(a, b):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def exact_trial(system, D):
    """Ground state of the system by dense diagonalisation, as a left-canonical
    MPS truncated to bond dimension D (test set-up for small L)."""
    import scipy.sparse.linalg as sla
    H = dense_hamiltonian(system)
    if H.shape[0] <= 64:
        w, V = np.linalg.eigh(H.toarray())
        psi = V[:, 0]
    else:
        w, V = sla.eigsh(H, k=1, which="SA", tol=1e-13)
        psi = V[:, 0]
    psi = np.real(psi)
    return dense_to_mps(psi / np.linalg.norm(psi), system.L, D)


def dense_to_mps(psi, L, D, d=2):
    """Exact left-canonical MPS of a dense vector (site 0 most significant),
    truncated to D by SVD."""
    tensors = []
    rest = psi.reshape(1, -1)
    dl = 1
    for k in range(L - 1):
        rest = rest.reshape(dl * d, -1)
        U, S, Vt = np.linalg.svd(rest, full_matrices=False)
        keep = min(D, int(np.sum(S > 1e-14)) or 1)
        U, S, Vt = U[:, :keep], S[:keep], Vt[:keep]
        A = U.reshape(dl, d, keep).transpose(1, 0, 2)  # (s, l, r)
        tensors.append(A)
        rest = (S[:, None] * Vt)
        dl = keep
    A = rest.reshape(dl, d, 1).transpose(1, 0, 2)
    tensors.append(A)
    return M.from_site_tensors(tensors)


def random_gauge(vd, data, seed, eps=0.3):
    """The same state in a random gauge: A_k <- A_k G_k, A_{k+1} <- G_k^-1 A_{k+1}
    on every inner bond (G_k = 1 + eps * gaussian).  Canonical MPS have exactly
    degenerate two-site singular values near the chain ends, where the
    reference's CompressState truncation picks an SVD-implementation-dependent
    subspace; in a generic gauge the truncation is well defined."""
    rng = np.random.default_rng(seed)
    L = len(vd) - 1
    T = [M.site_tensor(vd, data, k) for k in range(L)]  # A[s, l, r]
    for k in range(1, L):
        D = int(vd[k])
        G = np.eye(D) + eps * rng.normal(size=(D, D)) / np.sqrt(D)
        Gi = np.linalg.inv(G)
        T[k - 1] = np.einsum("slr,rq->slq", T[k - 1], G)
        T[k] = np.einsum("ql,slr->sqr", Gi, T[k])
    return M.from_site_tensors(T)
