"""The HIP path at the sizes of the benchmark configurations, held to the
oracle's golden fixtures (tests/golden/make_golden.py) bit for bit in FP64,
and the FP32 variant's error against FP64 reported at the steps the
configurations name.

  config1  ballistic mix, 512 spheres + 512 cubes, 300 steps
  config2  4096 batched worlds, 600 steps (8 sampled worlds in the fixture)
  config4  256 of the 16384 stacked-tower worlds, 240 steps at 1/120

With CBZ_REPORT_DIR set, the FP32 error reports are also written there as
JSON."""
import hashlib
import json
import os

import numpy as np
import pytest

from cubez_amd import scenes

pytestmark = pytest.mark.gpu

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


def report(name, data):
    d = os.environ.get("CBZ_REPORT_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, name + ".json"), "w") as fh:
            json.dump(data, fh, indent=1)
    print(name, json.dumps(data))


def rot_angle(q1, q2):
    """Angle of the rotation between two unit quaternions."""
    return 4.0 * np.arctan2(np.linalg.norm(q1 - q2, axis=1), np.linalg.norm(q1 + q2, axis=1))


def errors(w64, w32):
    dp = np.linalg.norm(w32.get("pos") - w64.get("pos"), axis=1)
    dv = np.linalg.norm(w32.get("vel") - w64.get("vel"), axis=1)
    ang = rot_angle(w64.get("quat"), w32.get("quat"))
    finite = bool(np.isfinite(w32.get("pos")).all() and np.isfinite(w32.get("quat")).all())
    return {"pos_max_m": float(dp.max()), "pos_mean_m": float(dp.mean()), "pos_p50_m": float(np.median(dp)),
            "pos_p99_m": float(np.percentile(dp, 99)), "vel_max": float(dv.max()),
            "angle_max_rad": float(np.nan_to_num(ang).max()), "finite": finite}


def test_config1_fp64_reproduces_the_fixture_for_300_steps(golden):
    from cubez_amd import World

    sc = scenes.ballistic(0, 512, 512)
    marks = (1, 60, 120, 180, 240, 300)
    with World(sc) as w:
        got, done = [], 0
        for i, k in enumerate(marks):
            got.append(w.step(sc.dt, k - done, sc.iter_mult, counts=True)[:, 0])
            done = k
            assert state_hash(w.get) == golden["config1_hash"][i], f"state after step {k}"
        assert np.array_equal(np.concatenate(got), golden["config1_counts"])


def test_config1_fp32_error_per_step_and_at_300():
    """FP32 (FP64 SAT and contact basis) against FP64 in lockstep, every step
    of config1's 300.  At step 1 most bodies differ by rounding only (the
    median is checked); where deep initial overlaps are resolved, FP32 can
    pick a different worst contact than FP64 and that cluster's bodies move
    differently from the first step; later the divergence is that of a
    chaotic system.  The report gives the per-step maxima and the
    distribution at steps 1, 60 and 300."""
    from cubez_amd import World

    sc = scenes.ballistic(0, 512, 512)
    per_step = []
    same_counts = []
    with World(sc, precision="f64") as w64, World(sc, precision="f32") as w32:
        for k in range(1, 301):
            c64 = w64.step(sc.dt, 1, sc.iter_mult, counts=True)[0, 0]
            c32 = w32.step(sc.dt, 1, sc.iter_mult, counts=True)[0, 0]
            same_counts.append(bool(c64 == c32))
            per_step.append(errors(w64, w32))
    first_diff = next((k + 1 for k, s in enumerate(same_counts) if not s), None)
    report("config1_fp32_error", {
        "steps": {str(k): per_step[k - 1] for k in (1, 60, 300)},
        "pos_max_per_step_m": [e["pos_max_m"] for e in per_step],
        "contact_counts_identical_through_step": (first_diff - 1) if first_diff else 300,
    })
    assert all(e["finite"] for e in per_step)
    assert per_step[0]["pos_p50_m"] < 1e-5


def test_config2_full_size_fp64_against_the_fixture_and_fp32_error(golden):
    """All 4096 batched worlds for 600 steps: in FP64 the sampled worlds'
    per-step contact counts and their states after steps 1, 60 and 600 are
    the oracle's bit for bit, and every state stays finite; FP32 against
    FP64 at the same steps is reported."""
    from cubez_amd import World

    sc = scenes.batched(0, 4096)
    ws = [int(x) for x in golden["config2_worlds"]]
    rep = {}
    with World(sc, precision="f64") as w64, World(sc, precision="f32") as w32:
        got, done = [], 0
        for i, k in enumerate((1, 60, 600)):
            got.append(w64.step(sc.dt, k - done, sc.iter_mult, counts=True)[:, ws])
            w32.step(sc.dt, k - done, sc.iter_mult)
            done = k
            for j, wi in enumerate(ws):
                assert state_hash(w64.get, sc.world_slice(wi)) == golden["config2_hash"][j][i], (wi, k)
            assert all(np.isfinite(w64.get(f)).all() for f in STATE)
            rep[str(k)] = errors(w64, w32)
            assert rep[str(k)]["finite"]
        assert np.array_equal(np.concatenate(got), golden["config2_counts"])
    report("config2_fp32_error", rep)
    assert rep["1"]["pos_p50_m"] < 1e-5


def test_config4_towers_fp64_against_the_fixture(golden):
    """256 of config4's 16384 stacked-tower worlds (every 64th), 240 steps
    at 1/120: per-world per-step contact counts and the states after steps
    120 and 240 are the oracle's bit for bit."""
    import importlib.util

    from cubez_amd import World

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(os.path.dirname(GOLDEN), "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    sc = mg.config4_scene()
    with World(sc, precision="f64") as w:
        got, done = [], 0
        for i, k in enumerate((120, 240)):
            got.append(w.step(sc.dt, k - done, sc.iter_mult, counts=True))
            done = k
            bad = [wi for wi in range(sc.n_worlds)
                   if state_hash(w.get, sc.world_slice(wi)) != golden["config4_hash"][wi][i]]
            assert not bad, f"worlds differing after step {k}: {bad[:10]}"
        assert np.array_equal(np.concatenate(got), golden["config4_counts"])
