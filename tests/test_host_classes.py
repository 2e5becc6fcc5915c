"""The MPSstate / Walker methods of the C++ host classes (include/mpsqmc.hpp,
mirroring include/MPSstate.h:19-102 and include/Walker.h:13-52 of the
reference) against dense state vectors, one method per run of mpsqmc_tool.
Host code only (no GPU)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

from mpsqmc_amd import mps as M
from mpsqmc_amd import systems

from .common import dense_hamiltonian, random_trial

TOOL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mpsqmc_amd", "mpsqmc_tool")


def _tool(tmp_path, method, vd, data, *args, trial=None):
    src = str(tmp_path / "in.mps")
    dst = str(tmp_path / "out.mps")
    M.write_mps(src, vd, data)
    cmd = [TOOL, method, "--in", src, "--out", dst] + [str(a) for a in args]
    if trial is not None:
        tp = str(tmp_path / "trial.mps")
        M.write_mps(tp, *trial)
        cmd += ["--trial", tp]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    vals = dict(line.split() for line in out.stdout.strip().splitlines() if line)
    _, _, _, ovd, odata = M.read_mps(dst)
    return {k: float(v) for k, v in vals.items()}, M.to_dense(ovd, odata), (ovd, odata)


def _on(L, i, O):
    out = np.array([[1.0]])
    for q in range(L):
        out = np.kron(out, O if q == i else np.eye(2))
    return out


@pytest.mark.parametrize("hc", [False, True])
def test_apply_svd_term(tmp_path, hc):
    # MPSstate.cpp:505-573: sqrt(s_a) U_a on site i, sqrt(s_b) V_b on site j (transposed for HC)
    L = 6
    s = systems.chain(L)
    vd, data = M.random_mps(L, 4, 2, 3)
    S, U, VT = s.tables_per_coupling()
    for (i, j, a, b) in ((1, 2, 0, 3), (3, 4, 2, 1)):
        _, out, _ = _tool(tmp_path, "svd-term", vd, data, "--first", i, "--second", j, "--a", a, "--b", b,
                          *(["--hc"] if hc else []))
        c = i  # chain: coupling (i, i+1) is number i
        Ua = U[c].reshape(2, 2, 4, order="F")[:, :, a] * np.sqrt(S[c][a])
        Vb = VT[c].reshape(4, 2, 2, order="F")[b] * np.sqrt(S[c][b])
        if hc:
            Ua, Vb = Ua.T, Vb.T
        want = _on(L, i, Ua) @ (_on(L, j, Vb) @ M.to_dense(vd, data))
        assert np.allclose(out, want, atol=1e-13)


def test_apply_grid_term(tmp_path):
    # MPSstate.cpp:575-633: sum_k sqrt(s_k dim) x_k U_k on site i and the same with V_k on site j
    L = 6
    s = systems.snake_lattice(2, 3)
    vd, data = M.random_mps(L, 4, 2, 5)
    S, U, VT = s.tables_per_coupling()
    f, sec, _ = s.couplings()
    grid = s.grid()
    for c, g in ((0, 7), (len(f) - 1, 200)):
        _, out, _ = _tool(tmp_path, "grid-term", vd, data, "--snake", "2x3", "--first", f[c], "--second", sec[c],
                          "--point", g)
        Ux = sum(np.sqrt(S[c][k] * 4) * grid[g][k] * U[c].reshape(2, 2, 4, order="F")[:, :, k] for k in range(4))
        Vx = sum(np.sqrt(S[c][k] * 4) * grid[g][k] * VT[c].reshape(4, 2, 2, order="F")[k] for k in range(4))
        want = _on(L, f[c], Ux) @ (_on(L, sec[c], Vx) @ M.to_dense(vd, data))
        assert np.allclose(out, want, atol=1e-13)


def test_apply_one_site_term_everywhere(tmp_path):
    # MPSstate.cpp:635-659 with TrotterHeisenberg.cpp:80-88: exp(dtau h S^z / 2) on every site
    L, h, dt = 5, 0.8, 0.05
    vd, data = M.random_mps(L, 4, 2, 6)
    _, out, _ = _tool(tmp_path, "field", vd, data, "--field", h, "--dtau", dt)
    H = dense_hamiltonian(systems.chain(L, h=h), sparse=False)
    H0 = dense_hamiltonian(systems.chain(L), sparse=False)
    want = sl.expm(-0.5 * dt * (H - H0)) @ M.to_dense(vd, data)
    assert np.allclose(out, want, atol=1e-13)
    _, out0, _ = _tool(tmp_path, "field", vd, data, "--field", 0.0)  # no field: untouched
    assert np.array_equal(out0, M.to_dense(vd, data))


@pytest.mark.parametrize("shape", ["chain", "snake"])
def test_apply_mpo(tmp_path, shape):
    # MPSstate.cpp:413-471: H|psi> as an MPS; <psi|H psi> is the energy
    if shape == "chain":
        s, extra = systems.chain(7, h=0.3), ["--field", 0.3]
    else:
        s, extra = systems.snake_lattice(2, 4), ["--snake", "2x4"]
    vd, data = M.random_mps(s.L, 4, 2, 8)
    vals, out, (ovd, _) = _tool(tmp_path, "mpo", vd, data, *extra)
    psi = M.to_dense(vd, data)
    H = dense_hamiltonian(s, sparse=False)
    assert np.allclose(out, H @ psi, atol=1e-12 * np.abs(H @ psi).max())
    assert abs(vals["energy"] - psi @ H @ psi) <= 1e-12 * abs(psi @ H @ psi)
    assert ovd[0] == 1 and ovd[-1] == 1


def test_right_normalize_and_scalar(tmp_path):
    L = 7
    vd, data = M.random_mps(L, 5, 2, 9)
    psi = M.to_dense(vd, data)
    vals, out, (ovd, odata) = _tool(tmp_path, "right-normalize", vd, data)
    assert np.allclose(vals["norm"] * out, psi, atol=1e-12 * np.abs(psi).max())
    assert abs(np.linalg.norm(out) - 1.0) < 1e-12
    for k in range(1, L):  # sum_s A[s] A[s]^T = 1 right of site 0
        A = M.site_tensor(ovd, odata, k)
        G = sum(A[q] @ A[q].T for q in range(2))
        assert np.allclose(G, np.eye(G.shape[0]), atol=1e-12)
    _, out, _ = _tool(tmp_path, "scalar", vd, data, "--factor", -2.5)
    assert np.allclose(out, -2.5 * psi, atol=1e-12 * np.abs(psi).max())


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_walker_overlap_and_local_energy(tmp_path, sign):
    # Walker.cpp:54-77: setOverlap left-normalises and fixes the phase (O > 0);
    # calcIndividualProjectedEnergy(H trial) = <trial|H|phi> / O
    s = systems.snake_lattice(2, 3)
    tvd, tdata = random_trial(6, 4, 12)
    wvd, wdata = M.random_mps(6, 3, 2, 13)
    wdata = wdata * sign
    vals, out, _ = _tool(tmp_path, "walker", wvd, wdata, "--snake", "2x3", trial=(tvd, tdata))
    t = M.to_dense(tvd, tdata)
    phi = M.to_dense(wvd, wdata)
    phi = phi / np.linalg.norm(phi)
    O = t @ phi
    if O < 0:
        phi, O = -phi, -O
    H = dense_hamiltonian(s, sparse=False)
    assert abs(vals["overlap"] - O) <= 1e-12 * max(O, 1e-3)
    assert abs(vals["energy"] - (t @ H @ phi) / O) <= 1e-11 * max(1.0, abs(vals["energy"]))
    assert np.allclose(out, phi, atol=1e-12)
