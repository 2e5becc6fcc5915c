"""Tile shapes of the general path's FP32 products (16 x 16, 32 x 32, 32 x 64
per wave, chosen per walker bond bucket) and their autotuner.

Every table gives the FP64 engine's weights, overlaps and local energies to the
FP32 tolerance on the same uniforms (uneven walker bonds 16 / 32 / 48 / 64 on
the 4x4 snake with a D = 64 trial); FP64 results do not depend on the table
at all; mqe_autotune leaves the population as it was, stores its table in the
cache file and a new engine with the same key starts from it."""
import os

import numpy as np
import pytest

from mpsqmc_amd import systems
from mpsqmc_amd.engine import Engine
from mpsqmc_amd.rng import Random

from .common import random_gauge, random_trial, rel

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("require_gpu")]

TABLES = {"all16": [16] * 8, "all32": [32] * 8, "wide": [64] * 8, "mixed": [16, 32, 64, 64, 32, 32, 32, 32]}


def _setup(precision, seed=5):
    s = systems.snake_lattice(4, 4)
    tvd, tdata = random_trial(s.L, 64, seed)
    dws = [16, 32, 48, 64, 16, 32, 48, 64]
    e = Engine(s, tvd, tdata, n_desired=8, capacity=32, max_bond=64, precision=precision, seed=seed)
    gvd, gdata = random_gauge(tvd, tdata, seed + 1)  # generic gauge: a well-defined truncation
    e.load_copies(gvd, gdata, 8)
    e.compress_walkers([D if D < 64 else 0 for D in dws])
    return e


def _steps(e, tiles, steps=2):
    if tiles is not None:
        e.set_tiles(tiles)
    e.prepare()
    rng = Random(13)
    out = []
    for _ in range(steps):
        e.propagate(rng.uniforms(e.num_walkers() * e.nC))
        sc = e.scalars()
        out.append((e.indices().copy(), sc["w"].copy(), sc["O"].copy(), sc["EL"].copy()))
    return out


def test_every_table_matches_fp64():
    ref = _setup("fp64")
    want = _steps(ref, None)
    ref.close()
    for name, t in TABLES.items():
        e = _setup("fp32")
        got = _steps(e, t)
        assert list(e.tiles) == t
        e.close()
        for (i1, w1, O1, E1), (i2, w2, O2, E2) in zip(got, want):
            same = np.all(i1 == i2, axis=1)
            assert same.mean() >= 0.75, name
            assert rel(w1[same], w2[same]) < 2e-4, name
            assert rel(O1[same], O2[same]) < 2e-4, name
            assert rel(E1[same], E2[same]) < 2e-3, name
            if not same.all():
                break


def test_fp64_ignores_the_table():
    runs = []
    for t in (TABLES["all16"], TABLES["wide"]):
        e = _setup("fp64")
        runs.append(_steps(e, t))
        e.close()
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


def test_autotune_restores_the_walk_and_caches(tmp_path, monkeypatch):
    cache = str(tmp_path / "tiles.txt")
    monkeypatch.setenv("MQE_TUNE_CACHE", cache)
    a = _setup("fp32")
    a.prepare()
    tiles, ms0, ms1 = a.autotune(reps=2)
    assert set(tiles) <= {16, 32, 64}
    assert ms0 > 0 and ms1 > 0 and ms1 <= ms0 * 1.02
    assert os.path.exists(cache)
    # the walk continues from the state before tuning: same as an untuned engine with that table
    ta = []
    rng = Random(3)
    for _ in range(2):
        a.propagate(rng.uniforms(a.num_walkers() * a.nC))
        ta.append(a.scalars()["w"].copy())
    a.close()
    monkeypatch.setenv("MQE_TUNE_CACHE", str(tmp_path / "none.txt"))
    b = _setup("fp32")
    b.prepare()          # prepared with the default table, as `a` was before tuning
    b.set_tiles(tiles)
    rng = Random(3)
    for k in range(2):
        b.propagate(rng.uniforms(b.num_walkers() * b.nC))
        np.testing.assert_array_equal(b.scalars()["w"], ta[k])
    b.close()
    # a new engine with the same key starts from the cached table
    monkeypatch.setenv("MQE_TUNE_CACHE", cache)
    c = _setup("fp32")
    assert list(c.tiles) == list(tiles)
    c.close()
