"""Flat MPS containers in the reference's layout.

An MPS is held as ``(vd, data)``: ``vd`` the L+1 bond dimensions and ``data``
the site tensors concatenated in site order, each stored as
``A[row + dimL*(col + dimR*s)]`` (MPStensor.cpp:91-95).

Text I/O follows MPSstate.cpp:92-128 (reader) and :661-686 (writer):
``L Dtrunc d``, then ``i VirtualD[i]`` for i = 0..L, then per site ``i size``
followed by ``i j value`` lines.
"""
import ctypes

import numpy as np

from ._lib import check, lib


def bond_dims(L, D, d=2):
    """VirtualD of MPSstate(L, D, d) (MPSstate.cpp:20-28)."""
    vd = np.zeros(L + 1, dtype=np.int32)
    lib().mqe_bond_dims(L, D, d, vd.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return vd


def bond_dims_py(L, D, d=2):
    """Pure-Python copy of the bond rule (used where the library is not loaded)."""
    vd = [1] * (L + 1)
    for k in range(1, L):
        vd[k] = min(vd[k - 1] * d, D)
    vd[L] = 1
    for k in range(L - 1, 0, -1):
        vd[k] = min(min(vd[k + 1] * d, D), vd[k])
    return np.array(vd, dtype=np.int32)


def flat_size(vd, d=2):
    vd = np.asarray(vd, dtype=np.int64)
    return int(np.sum(vd[:-1] * vd[1:] * d))


def site_offsets(vd, d=2):
    vd = np.asarray(vd, dtype=np.int64)
    sizes = vd[:-1] * vd[1:] * d
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def site_tensor(vd, data, k, d=2):
    """Site k as an array A[s, row, col] (a view-independent copy)."""
    off = site_offsets(vd, d)
    dl, dr = int(vd[k]), int(vd[k + 1])
    blk = np.asarray(data[off[k]:off[k + 1]])
    return blk.reshape(d, dr, dl).transpose(0, 2, 1).copy()


def from_site_tensors(tensors):
    """Inverse of site_tensor: list of A[s, row, col] -> (vd, data)."""
    vd = [tensors[0].shape[1]] + [t.shape[2] for t in tensors]
    parts = [np.ascontiguousarray(t.transpose(0, 2, 1)).reshape(-1) for t in tensors]
    return np.array(vd, dtype=np.int32), np.concatenate(parts)


def random_mps(L, D, d=2, rng=None):
    """Random MPS with the MPSstate bond rule; entries +-U(0,1) like MPStensor::random
    (MPStensor.cpp:75-79), drawn from numpy's generator."""
    rng = np.random.default_rng(rng)
    vd = bond_dims_py(L, D, d)
    n = flat_size(vd, d)
    data = rng.uniform(0.0, 1.0, n) * np.where(rng.uniform(0.0, 1.0, n) < 0.5, -1.0, 1.0)
    return vd, data


def left_normalize(vd, data, d=2):
    """MPSstate::LeftNormalize (MPSstate.cpp:216-229); returns (data, final R)."""
    vd = np.ascontiguousarray(vd, dtype=np.int32)
    out = np.array(data, dtype=np.float64)
    nrm = ctypes.c_double()
    check(lib().mqe_left_normalize(len(vd) - 1, d, vd.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(nrm)))
    return out, nrm.value


def compress(vd, data, Dt, d=2):
    """MPSstate::CompressState(Dt) (MPSstate.cpp:379-404)."""
    vd = np.ascontiguousarray(vd, dtype=np.int32)
    data = np.ascontiguousarray(data, dtype=np.float64)
    L = len(vd) - 1
    vo = np.zeros(L + 1, dtype=np.int32)
    out = np.zeros(data.size)
    check(lib().mqe_compress(L, d, vd.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             data.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), Dt,
                             vo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return vo, out[: flat_size(vo, d)].copy()


def read_mps(path):
    """Read the reference's .mps text format; returns (L, Dtrunc, d, vd, data)."""
    with open(path) as f:
        tok = f.read().split()
    it = iter(tok)
    L, Dt, d = int(next(it)), int(next(it)), int(next(it))
    vd = np.zeros(L + 1, dtype=np.int32)
    for i in range(L + 1):
        idx = int(next(it))
        vd[i] = int(next(it))
        if idx != i:
            raise ValueError(f"{path}: bond index {idx} != {i}")
    parts = []
    for i in range(L):
        idx = int(next(it))
        n = int(next(it))
        if idx != i or n != vd[i] * vd[i + 1] * d:
            raise ValueError(f"{path}: bad site header at site {i}")
        vals = np.empty(n)
        for j in range(n):
            next(it)
            next(it)
            vals[j] = float(next(it))
        parts.append(vals)
    return L, Dt, d, vd, np.concatenate(parts)


def write_mps(path, vd, data, d=2, Dtrunc=None):
    """Write the reference's .mps text format (precision as an ostream default is
    lower; 17 significant digits are written so the file round-trips)."""
    vd = np.asarray(vd)
    L = len(vd) - 1
    Dt = int(Dtrunc if Dtrunc is not None else vd.max())
    off = site_offsets(vd, d)
    with open(path, "w") as f:
        f.write(f"{L}\t{Dt}\t{d}\n")
        for i in range(L + 1):
            f.write(f"{i}\t{int(vd[i])}\n")
        for i in range(L):
            n = int(off[i + 1] - off[i])
            f.write(f"{i}\t{n}\n")
            for j in range(n):
                f.write(f"{i}\t{j}\t{data[off[i] + j]:.17g}\n")


def to_dense(vd, data, d=2):
    """Full state vector with site 0 as the most significant index (the order of
    numpy.kron(op_0, op_1, ...)).  Small L only."""
    L = len(vd) - 1
    psi = np.ones((1, 1))  # (configurations of sites < k, right bond)
    for k in range(L):
        A = site_tensor(vd, data, k, d)  # (s, l, r)
        psi = np.einsum("pl,slr->psr", psi, A).reshape(-1, A.shape[2])
    return psi[:, 0].copy()


def dmrg(system, D, max_sweeps=20, tol=1e-10, seed=20131015):
    """Trial state by two-site DMRG (DMRG.cpp:50-100, as MPSQMC2::SetupTrial
    uses it): ground state of the system's Hamiltonian in the MPSstate bond
    profile of D, left-normalized.  Returns (vd, data, energy)."""
    f, s, Jc = system.couplings()
    L = system.L
    f = np.ascontiguousarray(f, dtype=np.int32)
    s = np.ascontiguousarray(s, dtype=np.int32)
    Jc = np.ascontiguousarray(Jc, dtype=np.float64)
    vd = np.asarray(bond_dims(L, D, system.d), dtype=np.int32)
    out = np.zeros(flat_size(vd, system.d))
    vo = np.zeros(L + 1, dtype=np.int32)
    E = ctypes.c_double()
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().mqe_dmrg(L, system.d, len(f), f.ctypes.data_as(ip), s.ctypes.data_as(ip), Jc.ctypes.data_as(dp),
                         float(system.h), int(D), int(max_sweeps), float(tol), int(seed), vo.ctypes.data_as(ip),
                         out.ctypes.data_as(dp), ctypes.byref(E)))
    return vo, out, E.value
