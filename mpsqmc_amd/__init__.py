"""MI355X walker engine for MPS projector QMC of Heisenberg models.

The hot path (per-walker Trotter propagation, normalisation, trial overlaps,
local energies and population control of MPSQMC2) runs as hand-written HIP
kernels for gfx950 in ``libmqe.so`` behind the C ABI of ``include/mqe.h``.
"""
from ._lib import LIB_PATH, MqeError, lib  # noqa: F401

__all__ = ["LIB_PATH", "MqeError", "lib"]
