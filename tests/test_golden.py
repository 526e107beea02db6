"""Golden fixtures (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the FP64 oracle): per-step contact counts
and SHA-256 hashes of the packed FP64 state (pos, quat, vel, rot).

The CPU tests hold the oracle to the committed fixtures; the GPU tests hold
the HIP path to them bit for bit, with no oracle in the loop."""
import hashlib
import os

import numpy as np
import pytest

from cubez_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")
STATE = ("pos", "quat", "vel", "rot")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


def state_hash(get, sl=slice(None)):
    h = hashlib.sha256()
    for k in STATE:
        h.update(np.ascontiguousarray(get(k)[sl], "<f8").tobytes())
    return h.hexdigest()


# --- the oracle against the fixtures (CPU) ----------------------------------

def _oracle_run(sc, world, marks):
    from oracle.cbz_oracle import OracleWorld

    o = OracleWorld(sc, world)
    o.calculate_derived()
    counts, hashes, done = [], {}, 0
    for k in marks:
        counts.append(o.step(sc.dt, k - done, sc.iter_mult))
        done = k
        hashes[k] = state_hash(o.get)
    return np.concatenate(counts), hashes, o


@pytest.mark.parametrize("seed", [0, 3, 7])
def test_oracle_reproduces_cubedrop_fixture(golden, seed):
    counts, hashes, o = _oracle_run(scenes.cubedrop(seed), 0, (1, 60, 600))
    assert np.array_equal(counts, golden["cubedrop_counts"][seed])
    assert [hashes[k] for k in (1, 60, 600)] == list(golden["cubedrop_hash"][seed])


def test_oracle_reproduces_ballistic_and_batched_fixtures(golden):
    counts, hashes, _ = _oracle_run(scenes.ballistic(0, 48, 48), 0, (1, 60, 240))
    assert np.array_equal(counts, golden["ballistic_counts"])
    assert [hashes[k] for k in (1, 60, 240)] == list(golden["ballistic_hash"])
    sc = scenes.batched(3, 12, 8, 64)
    for w in range(sc.n_worlds):
        counts, hashes, _ = _oracle_run(sc, w, (240,))
        assert np.array_equal(counts, golden["batched_counts"][:, w])
        assert hashes[240] == golden["batched_hash"][w]


def test_fixture_states_are_consistent_with_hashes(golden):
    """The stored seed-0 states hash to the stored hashes (the file is
    self-consistent)."""
    for i, k in enumerate((1, 60)):
        h = state_hash(lambda f: golden[f"cubedrop0_step{k}_{f}"])
        assert h == golden["cubedrop_hash"][0][i]


# --- the HIP path against the fixtures (GPU) ---------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(8))
def test_gpu_reproduces_cubedrop_fixture(golden, seed):
    from cubez_amd import World

    sc = scenes.cubedrop(seed)
    with World(sc) as w:
        got, done = [], 0
        for i, k in enumerate((1, 60, 600)):
            got.append(w.step(sc.dt, k - done, sc.iter_mult, counts=True)[:, 0])
            done = k
            assert state_hash(w.get) == golden["cubedrop_hash"][seed][i], f"state after step {k}"
            if seed == 0 and k in (1, 60):
                for f in STATE:
                    assert np.array_equal(w.get(f), golden[f"cubedrop0_step{k}_{f}"]), (k, f)
        assert np.array_equal(np.concatenate(got), golden["cubedrop_counts"][seed])


@pytest.mark.gpu
def test_gpu_reproduces_ballistic_and_batched_fixtures(golden):
    from cubez_amd import World

    sc = scenes.ballistic(0, 48, 48)
    with World(sc) as w:
        got, done = [], 0
        for i, k in enumerate((1, 60, 240)):
            got.append(w.step(sc.dt, k - done, sc.iter_mult, counts=True)[:, 0])
            done = k
            assert state_hash(w.get) == golden["ballistic_hash"][i], f"state after step {k}"
        assert np.array_equal(np.concatenate(got), golden["ballistic_counts"])
    sc = scenes.batched(3, 12, 8, 64)
    with World(sc) as w:
        counts = w.step(sc.dt, 240, sc.iter_mult, counts=True)
        assert np.array_equal(counts, golden["batched_counts"])
        for wi in range(sc.n_worlds):
            assert state_hash(w.get, sc.world_slice(wi)) == golden["batched_hash"][wi], wi


# --- configurations 1, 2 and 4 at their sizes ---------------------------------

def test_oracle_reproduces_config_fixtures_in_part(golden):
    """Slices of the config fixtures, recomputed by the oracle: config1's
    first 60 steps, config2's smallest world over all 600 steps and its
    largest over 60, and four tower worlds over 240 steps."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(GOLDEN), "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    counts, hashes, _ = _oracle_run(scenes.ballistic(0, 512, 512), 0, (1, 60))
    assert np.array_equal(counts, golden["config1_counts"][:60])
    assert [hashes[1], hashes[60]] == list(golden["config1_hash"][:2])
    sc = scenes.batched(0, 4096)
    ws = list(golden["config2_worlds"])
    assert ws == mg.config2_worlds(sc)
    n = np.diff(sc.body_off)
    small, large = ws.index(int(np.argmin(n))), ws.index(int(np.argmax(n)))
    counts, hashes, _ = _oracle_run(sc, ws[small], (1, 60, 600))
    assert np.array_equal(counts, golden["config2_counts"][:, small])
    assert [hashes[k] for k in (1, 60, 600)] == list(golden["config2_hash"][small])
    counts, hashes, _ = _oracle_run(sc, ws[large], (1, 60))
    assert np.array_equal(counts, golden["config2_counts"][:60, large])
    assert [hashes[1], hashes[60]] == list(golden["config2_hash"][large][:2])
    sc = mg.config4_scene()
    for w in (0, 77, 128, 255):
        counts, hashes, _ = _oracle_run(sc, w, (120, 240))
        assert np.array_equal(counts, golden["config4_counts"][:, w])
        assert [hashes[120], hashes[240]] == list(golden["config4_hash"][w])
