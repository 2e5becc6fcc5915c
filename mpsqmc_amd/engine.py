"""Python front-end of the walker engine (thin wrapper over the C ABI).

The method names follow the reference driver that the engine replaces:
``propagate`` = MPSQMC2::PropagateSeparately (MPSQMC2.cpp:518-611),
``energy`` = MPSQMC2::EnergyFunctionAndHistory (:688-747),
``population_control`` = MPSQMC2::SeparatePopulationControl (:613-686).
"""
import ctypes
import os

import numpy as np

from . import mps as _mps
from ._lib import MQE_PREC_FP32, MQE_PREC_FP64, MqeMps, MqeParams, check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class Engine:
    """One GPU-resident walker population.

    system: mpsqmc_amd.systems.System; trial: (vd, data), already left-normalized
    (MPSQMC2.cpp:239).  precision: "fp64" (parity) or "fp32" (FP32 contractions).
    """

    def __init__(self, system, trial_vd, trial_data, n_desired, capacity=None, max_bond=None, precision="fp64",
                 seed=20131015):
        self.system = system
        self.L = system.L
        self.n_desired = int(n_desired)
        self.capacity = int(capacity if capacity is not None else max(1000, 3 * n_desired))
        tvd = np.ascontiguousarray(trial_vd, dtype=np.int32)
        tdata = np.ascontiguousarray(trial_data, dtype=np.float64)
        self.max_bond = int(max_bond if max_bond is not None else tvd.max())
        prm = MqeParams(self.L, system.d, MQE_PREC_FP32 if precision == "fp32" else MQE_PREC_FP64, self.n_desired,
                        self.capacity, self.max_bond, system.dtau, system.h, seed)
        self._keep = [tvd, tdata]
        m = MqeMps(self.L, system.d, _iptr(tvd), _dptr(tdata))
        f, s, Jc = system.couplings()
        S, U, VT = system.tables_per_coupling()
        grid = np.ascontiguousarray(system.grid())
        f = np.ascontiguousarray(f, dtype=np.int32)
        s = np.ascontiguousarray(s, dtype=np.int32)
        Jc = np.ascontiguousarray(Jc)
        S = np.ascontiguousarray(S).reshape(-1)
        U = np.ascontiguousarray(U).reshape(-1)
        VT = np.ascontiguousarray(VT).reshape(-1)
        self.nC = len(f)
        h = ctypes.c_void_p()
        check(lib().mqe_create(ctypes.byref(prm), ctypes.byref(m), self.nC, _iptr(f), _iptr(s), _dptr(Jc), _dptr(S),
                               _dptr(U), _dptr(VT), _dptr(grid), grid.shape[0], grid.shape[1], ctypes.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().mqe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- walkers ------------------------------------------------------------
    def load_walkers(self, vds, datas, w=None, O=None, eh=None):
        n = len(vds)
        vd = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.int32) for v in vds]))
        data = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in datas]))
        w = np.ascontiguousarray(np.ones(n) if w is None else w, dtype=np.float64)
        O = np.ascontiguousarray(np.ones(n) if O is None else O, dtype=np.float64)
        eh = np.ascontiguousarray(np.zeros(n) if eh is None else eh, dtype=np.float64)
        check(lib().mqe_load_walkers(self.h, n, _iptr(vd), _dptr(data), _dptr(w), _dptr(O), _dptr(eh)))
        self.n = n

    def load_copies(self, vd, data, n, O=1.0):
        """n copies of one MPS (SetupWalkers with copyTrial = true)."""
        self.load_walkers([vd] * n, [data] * n, np.ones(n), np.full(n, O), np.zeros(n))

    def compress_walkers(self, targets):
        """CompressState(targets[i]) + setOverlap of every walker on the device
        (targets[i] < 1: unchanged); call prepare() afterwards."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        assert t.size == self.num_walkers()
        check(lib().mqe_compress_walkers(self.h, _iptr(t)))

    # -- tile shapes of the general path ------------------------------------
    @property
    def tiles(self):
        """FP32 product tile shape (16, 32 or 64) per walker bond bucket (8 entries)."""
        t = np.zeros(8, dtype=np.int32)
        check(lib().mqe_get_tiles(self.h, _iptr(t)))
        return t

    def set_tiles(self, tiles):
        t = np.ascontiguousarray(tiles, dtype=np.int32)
        assert t.size == 8
        check(lib().mqe_set_tiles(self.h, _iptr(t)))

    def autotune(self, reps=2):
        """Tune the tile table on the current population; returns (tiles, ms_default, ms_tuned)."""
        t = np.zeros(8, dtype=np.int32)
        a, b = ctypes.c_double(), ctypes.c_double()
        check(lib().mqe_autotune(self.h, int(reps), _iptr(t), ctypes.byref(a), ctypes.byref(b)))
        return t, a.value, b.value

    # -- the step -----------------------------------------------------------
    def prepare(self):
        E = ctypes.c_double()
        check(lib().mqe_prepare(self.h, ctypes.byref(E)))
        return E.value

    def propagate(self, uniforms=None):
        sw = ctypes.c_double()
        if uniforms is not None:
            u = np.ascontiguousarray(uniforms, dtype=np.float64)
            assert u.size == self.n * self.nC
            check(lib().mqe_propagate(self.h, _dptr(u), ctypes.byref(sw)))
        else:
            check(lib().mqe_propagate(self.h, None, ctypes.byref(sw)))
        return sw.value

    def energy(self, step, do_fluct=True):
        EP = ctypes.c_double()
        om = ctypes.c_double()
        check(lib().mqe_energy(self.h, step, 1 if do_fluct else 0, ctypes.byref(EP), ctypes.byref(om)))
        return EP.value, om.value

    def population_control(self, scaling, uniforms=None):
        nn = ctypes.c_int()
        nu = ctypes.c_int()
        if uniforms is not None:
            u = np.ascontiguousarray(uniforms, dtype=np.float64)
            check(lib().mqe_population_control(self.h, scaling, _dptr(u), u.size, ctypes.byref(nn), ctypes.byref(nu)))
        else:
            check(lib().mqe_population_control(self.h, scaling, None, 0, ctypes.byref(nn), ctypes.byref(nu)))
        self.n = nn.value
        return nn.value, nu.value

    def step(self, step):
        check(lib().mqe_step(self.h, step))

    def run(self, first_step, n_steps, use_graph=False):
        check(lib().mqe_run(self.h, first_step, n_steps, 1 if use_graph else 0))

    def set_canonical_interval(self, steps):
        check(lib().mqe_set_canonical_interval(self.h, steps))

    # -- learned proposal (tile path) ------------------------------------------
    def collect_samples(self, max_samples):
        """Store (features, combos/O) rows of the following propagations, up to max_samples."""
        check(lib().mqe_collect_samples(self.h, int(max_samples)))

    def num_samples(self):
        n = ctypes.c_long()
        check(lib().mqe_num_samples(self.h, ctypes.byref(n)))
        return n.value

    def read_samples(self, first=0, n=None):
        from .mlp import NFEAT, NLABEL
        if n is None:
            n = self.num_samples() - first
        X = np.zeros((n, NFEAT), np.float32)
        Y = np.zeros((n, NLABEL), np.float32)
        fp = ctypes.POINTER(ctypes.c_float)
        check(lib().mqe_read_samples(self.h, int(first), int(n), X.ctypes.data_as(fp), Y.ctypes.data_as(fp)))
        return X, Y

    def train_proposal(self, net, epochs=5, batch=1024, lr=1e-3):
        """Fit the network on the collected rows (on the device); returns the last epoch's loss."""
        loss = ctypes.c_double()
        check(lib().mqe_train_proposal(self.h, net.h, int(epochs), int(batch), float(lr), ctypes.byref(loss)))
        return loss.value

    def set_proposal(self, net, mix=0.1):
        """Sample from the network's proposal with the importance weight (None: the reference's)."""
        self._net = net
        check(lib().mqe_set_proposal(self.h, net.h if net is not None else None, float(mix)))

    def set_proposal_table(self, pred, mix=0.1):
        """Fixed predicted labels (n x nC x 16, walker order) instead of a network; None clears."""
        if pred is None:
            check(lib().mqe_set_proposal_table(self.h, None, 0, 0.0))
            return
        pred = np.ascontiguousarray(pred, dtype=np.float32).reshape(-1, self.nC, 16)
        fp = ctypes.POINTER(ctypes.c_float)
        check(lib().mqe_set_proposal_table(self.h, pred.ctypes.data_as(fp), pred.shape[0], float(mix)))

    def bond_energies(self):
        """Bond local energies of the last energy sweep (n x nC); rows sum to E_L."""
        n = self.num_walkers()
        out = np.zeros((n, self.nC))
        check(lib().mqe_read_bond_energies(self.h, _dptr(out)))
        return out

    # -- checkpoint -----------------------------------------------------------
    def save_state(self, path, step=0):
        """Write the population snapshot (mqe_save_state) with a step number."""
        check(lib().mqe_save_state(self.h, os.fsencode(path), int(step)))

    def load_state(self, path):
        """Restore a snapshot written by an engine with the same parameters; returns its step."""
        st = ctypes.c_long()
        check(lib().mqe_load_state(self.h, os.fsencode(path), ctypes.byref(st)))
        self.n = self.num_walkers()
        return st.value

    @property
    def kernel_name(self):
        """Name of the walker kernel this engine launches."""
        return "walker32_kernel" if self.path == "tile32" else "walker_kernel"

    @property
    def path(self):
        """'tile32' (FP32 MFMA path) or 'general'."""
        return "tile32" if lib().mqe_kernel_path(self.h) == 1 else "general"

    def sync(self):
        check(lib().mqe_sync(self.h))
        self.n = self.num_walkers()

    def read_trace(self, first_step, n_steps):
        out = np.zeros((n_steps, 6))
        check(lib().mqe_read_trace(self.h, first_step, n_steps, _dptr(out)))
        return out

    # -- inspection ---------------------------------------------------------
    def num_walkers(self):
        n = ctypes.c_int()
        check(lib().mqe_num_walkers(self.h, ctypes.byref(n)))
        return n.value

    def scalars(self):
        n = self.num_walkers()
        self.n = n
        w, O, eh, EL = (np.zeros(n) for _ in range(4))
        check(lib().mqe_read_scalars(self.h, _dptr(w), _dptr(O), _dptr(eh), _dptr(EL)))
        return dict(w=w, O=O, eh=eh, EL=EL)

    def indices(self, n=None):
        """Sampled grid points of the last propagation (walkers x couplings); n: the first n walkers."""
        last = self.last_kernel_ms()[1]
        idx = np.zeros(max(last, 1) * self.nC, dtype=np.int32)
        check(lib().mqe_read_indices(self.h, _iptr(idx)))
        n = last if n is None else n
        assert n <= last, (n, last)
        return idx[: n * self.nC].reshape(n, self.nC)

    def walker(self, i):
        vd = np.zeros(self.L + 1, dtype=np.int32)
        size = ctypes.c_long()
        check(lib().mqe_read_walker(self.h, i, _iptr(vd), None, ctypes.byref(size)))
        data = np.zeros(size.value)
        check(lib().mqe_read_walker(self.h, i, _iptr(vd), _dptr(data), ctypes.byref(size)))
        return vd, data

    def overlaps(self, vd, data):
        vd = np.ascontiguousarray(vd, dtype=np.int32)
        data = np.ascontiguousarray(data, dtype=np.float64)
        m = MqeMps(self.L, self.system.d, _iptr(vd), _dptr(data))
        out = np.zeros(self.num_walkers())
        check(lib().mqe_overlaps(self.h, ctypes.byref(m), _dptr(out)))
        return out

    def set_timing(self, on=True):
        check(lib().mqe_set_timing(self.h, 1 if on else 0))

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        check(lib().mqe_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        nw = ctypes.c_int()
        check(lib().mqe_last_kernel_ms(self.h, ctypes.byref(ms), ctypes.byref(nw)))
        return ms.value, nw.value


__all__ = ["Engine", "MQE_PREC_FP32", "MQE_PREC_FP64", "_mps"]
