"""Heisenberg models, propagator tables and the auxiliary-field grid.

A ``System`` carries what HeisenbergMPO + TrotterHeisenberg + GridGenerator
carry in the reference: the couplings J_ij in the packed storage
``J[(j*(j-1))/2 + i]`` (HeisenbergMPO.cpp:266-272), the field h of
``-h sum S^z`` (HeisenbergMPO.h:98-103), dtau, the per-coupling Trotter tables
and the Marsaglia grid.
"""
import ctypes

import numpy as np

from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def trotter_tables(J, dtau, d=2):
    """Operator-Schmidt tables (S[4], U[16], VT[16]) of exp(-dtau J S.S)
    (TrotterHeisenberg.cpp:102-285 storage)."""
    S = np.zeros(d * d)
    U = np.zeros(d ** 4)
    VT = np.zeros(d ** 4)
    check(lib().mqe_trotter_tables(d, float(J), float(dtau), S.ctypes.data_as(_dp), U.ctypes.data_as(_dp),
                                   VT.ctypes.data_as(_dp)))
    return S, U, VT


def marsaglia_grid(dim=4, p=4):
    """GridGenerator(dim).FillMarsaglia(p) (GridGenerator.cpp:167-225): (N, dim) array."""
    g = np.zeros(dim * p ** dim)
    n = ctypes.c_int()
    check(lib().mqe_marsaglia_grid(dim, p, g.ctypes.data_as(_dp), ctypes.byref(n)))
    return g[: dim * n.value].reshape(n.value, dim).copy()


def packed_index(i, j):
    if i > j:
        i, j = j, i
    return (j * (j - 1)) // 2 + i


class System:
    def __init__(self, L, J_packed, h=0.0, dtau=0.01, d=2, grid_points_per_coordinate=4, name="custom"):
        self.L = int(L)
        self.d = int(d)
        self.J = np.asarray(J_packed, dtype=np.float64).copy()
        assert self.J.size == L * (L - 1) // 2
        self.h = float(h)
        self.dtau = float(dtau)
        self.name = name
        self.p = grid_points_per_coordinate
        self._grid = None
        self._tabs = None

    def coupling(self, i, j):
        return self.J[packed_index(i, j)]

    def couplings(self):
        """Nonzero couplings in lexicographic order (MPSQMC2.cpp:286-299)."""
        f, s, v = [], [], []
        for i in range(self.L - 1):
            for j in range(i + 1, self.L):
                x = self.J[packed_index(i, j)]
                if x != 0.0:
                    f.append(i)
                    s.append(j)
                    v.append(x)
        return np.array(f, dtype=np.int32), np.array(s, dtype=np.int32), np.array(v)

    def distinct_couplings(self):
        """Distinct nonzero J in first-seen order (TrotterHeisenberg.cpp:38-65)."""
        out = []
        for i in range(self.L - 1):
            for j in range(i + 1, self.L):
                x = self.J[packed_index(i, j)]
                if x != 0.0 and x not in out:
                    out.append(x)
        return np.array(out)

    def grid(self):
        if self._grid is None:
            self._grid = marsaglia_grid(self.d * self.d, self.p)
        return self._grid

    def tables_distinct(self):
        """dict(dJ, S, U, VT) per distinct coupling."""
        dJ = self.distinct_couplings()
        S, U, VT = [], [], []
        for x in dJ:
            s, u, vt = trotter_tables(x, self.dtau, self.d)
            S.append(s)
            U.append(u)
            VT.append(vt)
        return dict(dJ=dJ, S=np.array(S), U=np.array(U), VT=np.array(VT))

    def tables_per_coupling(self):
        """(S[nC,4], U[nC,16], VT[nC,16]) expanded over the coupling list."""
        if self._tabs is None:
            t = self.tables_distinct()
            _, _, Jc = self.couplings()
            idx = [int(np.nonzero(t["dJ"] == x)[0][0]) for x in Jc]
            self._tabs = (t["S"][idx].copy(), t["U"][idx].copy(), t["VT"][idx].copy())
        return self._tabs


def chain(L, J=1.0, h=0.0, dtau=0.01, periodic=False):
    """1D Heisenberg chain, open boundary unless periodic (main.cpp:38-41)."""
    Jp = np.zeros(L * (L - 1) // 2)
    for i in range(L - 1):
        Jp[packed_index(i, i + 1)] = J
    if periodic and L > 2:
        Jp[packed_index(0, L - 1)] = J
    return System(L, Jp, h=h, dtau=dtau, name=f"chain{L}")


def snake_site(r, c, cols):
    """Snake (boustrophedon) path: even rows left->right, odd rows right->left."""
    return r * cols + (c if r % 2 == 0 else cols - 1 - c)


def snake_lattice(rows, cols, J=1.0, h=0.0, dtau=0.01):
    """2D open square lattice with nearest-neighbour J, mapped to a chain by the
    snake path (long-range MPO bonds)."""
    L = rows * cols
    Jp = np.zeros(L * (L - 1) // 2)
    for r in range(rows):
        for c in range(cols):
            a = snake_site(r, c, cols)
            if c + 1 < cols:
                Jp[packed_index(a, snake_site(r, c + 1, cols))] = J
            if r + 1 < rows:
                Jp[packed_index(a, snake_site(r + 1, c, cols))] = J
    return System(L, Jp, h=h, dtau=dtau, name=f"snake{rows}x{cols}")


def square_periodic(base, J1=1.0, J2=0.0, dtau=0.01):
    """J1-J2 square lattice with periodic boundaries, column-major site index
    ix + base*iy (the lattice of main.cpp:67-134)."""
    L = base * base
    Jp = np.zeros(L * (L - 1) // 2)
    for ix in range(base):
        for iy in range(base):
            a = ix + base * iy
            for dx, dy, v in ((1, 0, J1), (0, 1, J1), (1, 1, J2), (1, -1, J2)):
                b = (ix + dx) % base + base * ((iy + dy) % base)
                if a != b and v != 0.0:
                    Jp[packed_index(a, b)] = v
    return System(L, Jp, dtau=dtau, name=f"square{base}pbc")
