"""ctypes front-end of the FP64 CPU oracle (``oracle/mpsqmc_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``.  The product library never loads it.

The oracle restates the reference's walker algorithm (MPSQMC2.cpp:518-747,
Walker.cpp:54-77, MPSstate.cpp:216-284,505-659, MPStensor.cpp:100-175,
TwoSiteObject.cpp:71-184, TrotterHeisenberg.cpp:102-285,
GridGenerator.cpp:167-225, HeisenbergMPO.cpp:99-325) in plain C, FP64.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint32)


def build(force=False):
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                              stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_inner.restype = ctypes.c_double
        L.orc_left_normalize.restype = ctypes.c_double
        L.orc_erfinv.restype = ctypes.c_double
        L.orc_marsaglia.restype = ctypes.c_int
        L.orc_compress.restype = ctypes.c_int
        L.orc_apply_mpo.restype = ctypes.c_long
        L.orc_mpo_entry.restype = ctypes.c_double
        L.orc_apply_mpo_min.restype = ctypes.c_long
        L.orc_pop_create.restype = ctypes.c_void_p
        L.orc_pop_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_pop_destroy.argtypes = [ctypes.c_void_p]
        L.orc_pop_set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_pop_set_energy_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_pop_ncouplings.argtypes = [ctypes.c_void_p]
        L.orc_pop_couplings.argtypes = [ctypes.c_void_p, _ip, _ip, _dp]
        L.orc_pop_nwalkers.argtypes = [ctypes.c_void_p]
        L.orc_pop_init_walkers.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_pop_set_walkers.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.orc_pop_walker_size.restype = ctypes.c_long
        L.orc_pop_walker_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_pop_get_walker.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp, _dp, _dp]
        L.orc_pop_get_scalars.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_pop_get_trial.restype = ctypes.c_long
        L.orc_pop_get_trial.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_pop_get_htrial.restype = ctypes.c_long
        L.orc_pop_get_htrial.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_pop_set_proposal.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.orc_pop_set_labels.argtypes = [ctypes.c_void_p, _dp]
        L.orc_pop_propagate.restype = ctypes.c_double
        L.orc_pop_propagate.argtypes = [ctypes.c_void_p, _dp, _ip]
        L.orc_pop_energy.restype = ctypes.c_double
        L.orc_pop_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_pop_popcontrol.restype = ctypes.c_int
        L.orc_pop_popcontrol.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_int,
                                         _ip, _ip]
        L.orc_pop_walk.restype = ctypes.c_int
        L.orc_pop_walk.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                   _dp, _dp]
        L.orc_nthreads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


# --------------------------------------------------------------------------
# primitives
# --------------------------------------------------------------------------
def mt_raw(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_mt_raw(ctypes.c_uint32(seed), n, out.ctypes.data_as(_up))
    return out


def mt_uniforms(seed, n, skip=0):
    out = np.zeros(n)
    lib().orc_mt_uniforms(ctypes.c_uint32(seed), ctypes.c_long(skip), n, out.ctypes.data_as(_dp))
    return out


def operator(op, d):
    """op: 0 zero, 1 I, 2 Sz, 3 Sx, 4 iSy; returns d x d (row, col)."""
    out = np.zeros(d * d)
    lib().orc_operator(op, d, out.ctypes.data_as(_dp))
    return out.reshape(d, d, order="F")


def bond_dims(L, D, d):
    vd = np.zeros(L + 1, dtype=np.int32)
    lib().orc_bond_dims(L, D, d, vd.ctypes.data_as(_ip))
    return vd


def inner(L, d, vda, a, vdb, b):
    vda, pa = _i(vda)
    a, xa = _d(a)
    vdb, pb = _i(vdb)
    b, xb = _d(b)
    return lib().orc_inner(L, d, pa, xa, pb, xb)


def left_normalize(L, d, vd, data):
    vd, pv = _i(vd)
    data = np.array(data, dtype=np.float64)
    r = lib().orc_left_normalize(L, d, pv, data.ctypes.data_as(_dp))
    return r, data


def qr(A):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    m, n = A.shape
    R = np.zeros((n, n), order="F")
    lib().orc_qr(m, n, A.ctypes.data_as(_dp), R.ctypes.data_as(_dp))
    return A, R


def svd(A):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    m, n = A.shape
    k = min(m, n)
    S = np.zeros(k)
    U = np.zeros((m, k), order="F")
    VT = np.zeros((k, n), order="F")
    lib().orc_svd(m, n, A.ctypes.data_as(_dp), S.ctypes.data_as(_dp), U.ctypes.data_as(_dp),
                  VT.ctypes.data_as(_dp))
    return U, S, VT


def trotter_tables(d, J, dtau):
    n = d * d
    S = np.zeros(n)
    U = np.zeros(n * n)
    VT = np.zeros(n * n)
    P = np.zeros(n * n)
    lib().orc_trotter_tables(d, ctypes.c_double(J), ctypes.c_double(dtau), S.ctypes.data_as(_dp),
                             U.ctypes.data_as(_dp), VT.ctypes.data_as(_dp), P.ctypes.data_as(_dp))
    return S, U, VT, P


def erfinv(y):
    return lib().orc_erfinv(ctypes.c_double(y))


def marsaglia(dim, p):
    cap = dim * p ** dim
    g = np.zeros(cap)
    N = lib().orc_marsaglia(dim, p, g.ctypes.data_as(_dp))
    return g[: dim * N].copy(), N


def compress(L, d, vd, data, Dt):
    vd, pv = _i(vd)
    data, pd = _d(data)
    vo = np.zeros(L + 1, dtype=np.int32)
    out = np.zeros(data.size + 16)
    D = lib().orc_compress(L, d, pv, pd, Dt, vo.ctypes.data_as(_ip), out.ctypes.data_as(_dp))
    n = sum(int(vo[k]) * int(vo[k + 1]) * d for k in range(L))
    return vo, out[:n].copy(), D


def apply_mpo(L, d, J, h, vd, data):
    J, pJ = _d(J)
    vd, pv = _i(vd)
    data, pd = _d(data)
    vo = np.zeros(L + 1, dtype=np.int32)
    n = lib().orc_apply_mpo(L, d, pJ, ctypes.c_double(h), pv, pd, vo.ctypes.data_as(_ip), None)
    out = np.zeros(n)
    lib().orc_apply_mpo(L, d, pJ, ctypes.c_double(h), pv, pd, vo.ctypes.data_as(_ip), out.ctypes.data_as(_dp))
    return vo, out


def apply_mpo_min(L, d, J, h, vd, data):
    J, pJ = _d(J)
    vd, pv = _i(vd)
    data, pd = _d(data)
    vo = np.zeros(L + 1, dtype=np.int32)
    n = lib().orc_apply_mpo_min(L, d, pJ, ctypes.c_double(h), pv, pd, vo.ctypes.data_as(_ip), None)
    out = np.zeros(n)
    lib().orc_apply_mpo_min(L, d, pJ, ctypes.c_double(h), pv, pd, vo.ctypes.data_as(_ip), out.ctypes.data_as(_dp))
    return vo, out


def mpo_dims(L):
    dims = np.zeros(L + 1, dtype=np.int32)
    lib().orc_mpo_dims(L, dims.ctypes.data_as(_ip))
    return dims


# --------------------------------------------------------------------------
# population (MPSQMC2 restated)
# --------------------------------------------------------------------------
class Population:
    """Walker population driven by the reference algorithm.

    tables: dict with keys dJ (ndiff,), S (ndiff, n2), U (ndiff, n2*n2), VT (ndiff, n2*n2).
    """

    def __init__(self, L, d, trial_vd, trial_data, J, h, dtau, tables, grid, gdim, build_terms=True,
                 threads=None, energy_mode=0):
        self.L, self.d = L, d
        self._keep = []
        tvd, ptv = _i(trial_vd)
        td, ptd = _d(trial_data)
        Jv, pJ = _d(J)
        dJ, pdJ = _d(tables["dJ"])
        S, pS = _d(tables["S"])
        U, pU = _d(tables["U"])
        VT, pVT = _d(tables["VT"])
        g, pg = _d(grid)
        self._keep += [tvd, td, Jv, dJ, S, U, VT, g]
        Ng = g.size // gdim
        self.h = lib().orc_pop_create(L, d, ptv, ptd, pJ, h, dtau, len(dJ), pdJ, pS, pU, pVT, pg, Ng, gdim,
                                      1 if build_terms else 0)
        self.nC = lib().orc_pop_ncouplings(self.h)
        lib().orc_pop_set_threads(self.h, threads or 1)
        lib().orc_pop_set_energy_mode(self.h, energy_mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pop_destroy(self.h)
            self.h = None

    def couplings(self):
        f = np.zeros(self.nC, dtype=np.int32)
        s = np.zeros(self.nC, dtype=np.int32)
        j = np.zeros(self.nC)
        lib().orc_pop_couplings(self.h, f.ctypes.data_as(_ip), s.ctypes.data_as(_ip), j.ctypes.data_as(_dp))
        return f, s, j

    @property
    def nwalkers(self):
        return lib().orc_pop_nwalkers(self.h)

    def init_walkers(self, N, DW):
        lib().orc_pop_init_walkers(self.h, N, DW)

    def set_walkers(self, vds, datas, w, O, eh):
        n = len(vds)
        vd = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.int32) for v in vds]))
        data = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in datas]))
        w, pw = _d(w)
        O, pO = _d(O)
        eh, pe = _d(eh)
        lib().orc_pop_set_walkers(self.h, n, vd.ctypes.data_as(_ip), data.ctypes.data_as(_dp), pw, pO, pe)

    def walker(self, i):
        n = lib().orc_pop_walker_size(self.h, i)
        vd = np.zeros(self.L + 1, dtype=np.int32)
        data = np.zeros(n)
        w = ctypes.c_double()
        O = ctypes.c_double()
        e = ctypes.c_double()
        lib().orc_pop_get_walker(self.h, i, vd.ctypes.data_as(_ip), data.ctypes.data_as(_dp), ctypes.byref(w),
                                 ctypes.byref(O), ctypes.byref(e))
        return vd, data, w.value, O.value, e.value

    def scalars(self):
        n = self.nwalkers
        w = np.zeros(n)
        O = np.zeros(n)
        e = np.zeros(n)
        lib().orc_pop_get_scalars(self.h, w.ctypes.data_as(_dp), O.ctypes.data_as(_dp), e.ctypes.data_as(_dp))
        return w, O, e

    def trial(self):
        vd = np.zeros(self.L + 1, dtype=np.int32)
        n = lib().orc_pop_get_trial(self.h, vd.ctypes.data_as(_ip), None)
        data = np.zeros(n)
        lib().orc_pop_get_trial(self.h, vd.ctypes.data_as(_ip), data.ctypes.data_as(_dp))
        return vd, data

    def htrial(self):
        vd = np.zeros(self.L + 1, dtype=np.int32)
        n = lib().orc_pop_get_htrial(self.h, vd.ctypes.data_as(_ip), None)
        data = np.zeros(n)
        lib().orc_pop_get_htrial(self.h, vd.ctypes.data_as(_ip), data.ctypes.data_as(_dp))
        return vd, data

    def propagate(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        idx = np.zeros(self.nwalkers * self.nC, dtype=np.int32)
        s = lib().orc_pop_propagate(self.h, u.ctypes.data_as(_dp), idx.ctypes.data_as(_ip))
        return s, idx.reshape(self.nwalkers, self.nC)

    def set_proposal(self, pred, mix=0.1):
        """Sample from the proposal built from predicted labels (nwalkers x nC x 16,
        walker order) with the importance weight; None restores the reference's."""
        if pred is None:
            self._pred = None
            lib().orc_pop_set_proposal(self.h, None, 0.0)
            return
        self._pred = np.ascontiguousarray(pred, dtype=np.float64).reshape(-1)
        if self._pred.size != self.nwalkers * self.nC * 16:
            raise ValueError("pred must hold nwalkers * nC * 16 values")
        lib().orc_pop_set_proposal(self.h, self._pred.ctypes.data_as(_dp), float(mix))

    def propagate_with_labels(self, u):
        """propagate(u) that also returns combos/O per (walker, coupling)."""
        lab = np.zeros((self.nwalkers, self.nC, 16))
        lib().orc_pop_set_labels(self.h, lab.ctypes.data_as(_dp))
        try:
            s, idx = self.propagate(u)
        finally:
            lib().orc_pop_set_labels(self.h, None)
        return s, idx, lab

    def energy(self, step, dofluct=True):
        EP = ctypes.c_double()
        EL = np.zeros(self.nwalkers)
        om = lib().orc_pop_energy(self.h, step, 1 if dofluct else 0, ctypes.byref(EP), EL.ctypes.data_as(_dp))
        return EP.value, om, EL

    def popcontrol(self, scaling, u, cap):
        u = np.ascontiguousarray(u, dtype=np.float64)
        used = ctypes.c_int()
        nc = np.zeros(self.nwalkers, dtype=np.int32)
        r = lib().orc_pop_popcontrol(self.h, ctypes.c_double(scaling), u.ctypes.data_as(_dp), u.size, cap,
                                     ctypes.byref(used), nc.ctypes.data_as(_ip))
        return r, used.value, nc

    def walk(self, nsteps, seed, Ndesired, cap):
        trace = np.zeros((nsteps, 6))
        E0 = ctypes.c_double()
        r = lib().orc_pop_walk(self.h, nsteps, ctypes.c_uint32(seed), Ndesired, cap, ctypes.byref(E0),
                               trace.ctypes.data_as(_dp))
        if r != 0:
            raise RuntimeError("oracle walk: walker capacity exceeded")
        return E0.value, trace
