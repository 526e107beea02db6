"""The FP32 variant (error against FP64) and the graph-captured step."""
import json

import numpy as np
import pytest

from cubez_amd import World, precision, scenes

pytestmark = pytest.mark.gpu


def test_graph_step_matches_eager_every_replay():
    """cbz_step without counts replays a captured HIP graph; every replay,
    with eager work and synchronisation between replays, must give the same
    state as the eager path."""
    sc = scenes.batched(11, 16, 8, 64)
    with World(sc) as g, World(sc) as e:
        for k in range(8):
            g.step(sc.dt, 1 if k % 2 else 7, 4)            # graph path
            e.step(sc.dt, 1 if k % 2 else 7, 4, counts=True)  # eager path
            _ = g.get("flags")                              # eager work in between
            g.synchronize()
            for f in ("pos", "quat", "vel", "rot", "motion"):
                assert np.array_equal(g.get(f), e.get(f)), (k, f)
        assert np.array_equal(g.contact_counts(), e.contact_counts())


def test_graph_recaptures_on_new_dt():
    sc = scenes.cubedrop(2, 32)
    with World(sc) as g, World(sc) as e:
        for dt in (1 / 60, 1 / 120, 1 / 60):
            g.step(dt, 20, 4)
            e.step(dt, 20, 4, counts=True)
        assert np.array_equal(g.get("pos"), e.get("pos"))


def test_fp32_error_cubedrop():
    sc = scenes.cubedrop(0)
    rep = precision.compare(sc, (1, 60, 600))
    print(json.dumps(rep))
    assert rep[0]["pos_err_max"] < 1e-4      # one step: float32 rounding of ~50 m positions
    assert rep[1]["pos_err_max"] < 1e-2      # free fall for the first second
    assert all(np.isfinite(r["pos_err_max"]) for r in rep)


def test_fp32_error_ballistic_per_step():
    sc = scenes.ballistic(0, 96, 96)
    rep, per = precision.compare(sc, (1, 60, 120), per_step=True)
    print(json.dumps(rep), per[:5], per[-5:])
    assert per[0] < 1e-3
    assert len(per) == 120


def test_fp32_batched_runs_and_is_close_early():
    sc = scenes.batched(1, 64)
    rep = precision.compare(sc, (1, 60))
    print(json.dumps(rep))
    assert rep[0]["pos_err_max"] < 1e-4
    assert rep[0]["worlds_count_differs"] == 0
