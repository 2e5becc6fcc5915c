"""CPU tests of the C ABI library (no GPU needed): every symbol declared in
include/mqe.h is exported, the host-side set-up helpers agree with the oracle,
and calls that need a device fail with an error code instead of crashing."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from mpsqmc_amd import _lib, mps as M, systems
from mpsqmc_amd.rng import Random

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mqe.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mqe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("mqe_create", "mqe_load_walkers", "mqe_prepare", "mqe_propagate", "mqe_energy",
              "mqe_population_control", "mqe_step", "mqe_run", "mqe_read_walker", "mqe_destroy", "mqe_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes binding covers all of them
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def test_bond_dims_match_oracle():
    for L, D in [(10, 4), (16, 16), (32, 32), (50, 8), (7, 3)]:
        assert list(M.bond_dims(L, D)) == list(oracle.bond_dims(L, D, 2))
        assert list(M.bond_dims(L, D)) == list(M.bond_dims_py(L, D))


@pytest.mark.parametrize("J,dtau", [(1.0, 0.01), (0.5, 0.05), (-1.0, 0.02)])
def test_trotter_tables_match_oracle(J, dtau):
    S, U, VT = systems.trotter_tables(J, dtau)
    So, Uo, VTo, _ = oracle.trotter_tables(2, J, dtau)
    np.testing.assert_allclose(S, So, rtol=0, atol=1e-15)

    # the basis inside the degenerate triplet is a convention (closed form here,
    # an SVD in the oracle): compare the two-site operators they represent
    def op(S, U, VT):
        out = np.zeros((4, 4))
        for k in range(4):
            Uk = np.array([[U[i + 2 * (j + 2 * k)] for j in range(2)] for i in range(2)])
            Vk = np.array([[VT[k + 4 * (i + 2 * j)] for j in range(2)] for i in range(2)])
            out += S[k] * np.kron(Uk, Vk)
        return out

    np.testing.assert_allclose(op(S, U, VT), op(So, Uo, VTo), rtol=0, atol=1e-14)


def test_marsaglia_grid_matches_oracle():
    g = systems.marsaglia_grid(4, 4)
    go, n = oracle.marsaglia(4, 4)
    go = go.reshape(n, 4)
    assert g.shape == go.shape == (240, 4)
    np.testing.assert_allclose(g, go, rtol=0, atol=1e-15)


def test_left_normalize_matches_oracle():
    vd, data = M.random_mps(12, 8, 2, 3)
    a, na = M.left_normalize(vd, data)
    nb, b = oracle.left_normalize(12, 2, vd, data)
    assert abs(abs(na) - abs(nb)) <= 1e-12 * abs(nb)
    # same state (gauge may differ by the QR sign convention): compare overlaps
    ab = oracle.inner(12, 2, vd, a, vd, b)
    assert abs(abs(ab) - 1.0) < 1e-12


def test_compress_matches_oracle():
    # a generic (not left-normalised) state: a left-normalised one has
    # degenerate singular values at the first cuts, where the kept subspace is
    # a convention of the SVD
    vd, data = M.random_mps(10, 16, 2, 5)
    v1, c1 = M.compress(vd, data, 6)
    v2, c2, _ = oracle.compress(10, 2, vd, data, 6)
    assert list(v1) == list(v2)
    n1 = oracle.inner(10, 2, v1, c1, v1, c1)
    n2 = oracle.inner(10, 2, v2, c2, v2, c2)
    ov = oracle.inner(10, 2, v1, c1, v2, c2)
    assert abs(ov / np.sqrt(n1 * n2) - 1.0) < 1e-10


def test_compress_gauged_trial_matches_oracle():
    # a left-normalised trial in a random gauge (the device recompression tests
    # compress such states): the same state, generic two-site singular values,
    # so the host restatement and the oracle keep the same states
    from .common import random_trial, random_gauge
    tvd, tdata = random_trial(10, 16, 3)
    # canonical form: the first sweep's two-site matrices near the left end have
    # equal singular values (sqrt(2) four times at the second bond), so a
    # truncation to fewer states there is a convention of the SVD
    A = [M.site_tensor(tvd, tdata, k) for k in range(3)]  # A[s, l, r]

    def two_site_svd(a, b):
        dl, dr = a.shape[1], b.shape[2]
        th = np.einsum("slm,tmr->lsrt", a, b).reshape(dl * 2, dr * 2)
        return np.linalg.svd(th, full_matrices=False)
    U, S, VT = two_site_svd(A[0], A[1])  # bond 1, moving right: S VT into site 1
    A1 = (S[:, None] * VT).reshape(len(S), A[1].shape[2], 2).transpose(2, 0, 1)
    _, S2, _ = two_site_svd(A1, A[2])
    assert np.allclose(S2, np.sqrt(2.0), atol=1e-12) and len(S2) == 4
    gvd, gdata = random_gauge(tvd, tdata, 4)
    assert np.allclose(M.to_dense(gvd, gdata), M.to_dense(tvd, tdata))
    for D in (3, 4, 8):
        v1, c1 = M.compress(gvd, gdata, D)
        v2, c2, _ = oracle.compress(10, 2, gvd, gdata, D)
        assert list(v1) == list(v2)
        n1 = oracle.inner(10, 2, v1, c1, v1, c1)
        n2 = oracle.inner(10, 2, v2, c2, v2, c2)
        assert abs(oracle.inner(10, 2, v1, c1, v2, c2) / np.sqrt(n1 * n2) - 1.0) < 1e-10


def test_host_rng_contract():
    r = Random(20131015)
    u = r.uniforms(1000)
    np.testing.assert_array_equal(u, oracle.mt_uniforms(20131015, 1000))
    p = r.peek(5)
    np.testing.assert_array_equal(p, r.uniforms(5))
    r.discard(7)
    np.testing.assert_array_equal(r.uniforms(3), oracle.mt_uniforms(20131015, 3, skip=1000 + 5 + 7))


def test_engine_calls_fail_cleanly_without_a_device():
    """mqe_create either succeeds (a device is visible) or returns a negative
    code with a message; argument errors are reported before any device work."""
    lib = _lib.lib()
    h = ctypes.c_void_p()
    rc = lib.mqe_create(None, None, 0, None, None, None, None, None, None, None, 0, 0, ctypes.byref(h))
    assert rc == -1
    assert lib.mqe_last_error()
    d = ctypes.c_double()
    assert lib.mqe_prepare(None, ctypes.byref(d)) < 0
    assert lib.mqe_propagate(None, None, ctypes.byref(d)) < 0
    assert lib.mqe_trotter_tables(3, 1.0, 0.01, None, None, None) == -1
    assert lib.mqe_save_state(None, b"/tmp/x", 0) == -1
    assert lib.mqe_load_state(None, b"/tmp/x", None) == -1


def test_snake_lattice_bonds():
    s = systems.snake_lattice(4, 4)
    f, sec, J = s.couplings()
    assert len(f) == 24
    assert max(b - a for a, b in zip(f, sec)) == 7
    s6 = systems.snake_lattice(6, 6)
    f6, s6b, _ = s6.couplings()
    assert len(f6) == 60 and max(b - a for a, b in zip(f6, s6b)) == 11


def test_mps_file_round_trip(tmp_path):
    L, Dt, d, vd, data = M.read_mps(os.path.join(ROOT, "tests", "golden", "debug.mps"))
    p = tmp_path / "x.mps"
    M.write_mps(str(p), vd, data, d=d, Dtrunc=Dt)
    L2, Dt2, d2, vd2, data2 = M.read_mps(str(p))
    assert (L2, Dt2, d2) == (L, Dt, d)
    assert list(vd) == list(vd2)
    np.testing.assert_array_equal(data, data2)
