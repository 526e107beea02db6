"""State checkpoints (cbz_checkpoint_*): resume bit for bit, reject damaged
input without touching the context, and record trajectories on disk."""
import numpy as np
import pytest

from cubez_amd import World, checkpoint, scenes
from cubez_amd._native import CubezError
from tests.test_gpu_parity import STATE, same

pytestmark = pytest.mark.gpu

ALL = STATE + ("half", "shape", "acc", "inv_mass", "iit_body", "offset", "lin_damp", "ang_damp", "friction",
               "restitution")


def _mixed_scene():
    """Cubes and spheres over several worlds, with a static body, non-default
    damping, collider offsets and materials, so every field matters."""
    sc = scenes.batched(5, 6, 8, 80)
    f = sc.fields
    n = sc.n_bodies
    rng = np.random.default_rng(0)
    sph = rng.random(n) < 0.3
    f["shape"] = np.where(sph, 1, 0).astype(np.int32)
    f["lin_damp"] = rng.uniform(0.9, 1.0, n)
    f["ang_damp"] = rng.uniform(0.9, 1.0, n)
    f["friction"] = np.where(rng.random(n) < 0.2, 0.0, rng.uniform(0.2, 1.0, n))
    f["restitution"] = rng.uniform(0.0, 0.4, n)
    off = np.tile(np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]), (n, 1))
    off[::7, 9:12] = rng.uniform(-0.1, 0.1, (len(off[::7]), 3))
    f["offset"] = off
    f["flags"] = f["flags"].copy()
    f["flags"][3] = 4  # static
    return sc


def _world(sc, precision):
    w = World(sc, precision=precision)
    w.set_materials(True)
    w.set_plane_materials(np.linspace(0.0, 1.0, sc.n_worlds), np.linspace(0.05, 0.3, sc.n_worlds))
    return w


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_resume_from_checkpoint_is_bit_identical(precision):
    sc = _mixed_scene()
    with _world(sc, precision) as a, _world(sc, precision) as b:
        ca = a.step(sc.dt, 300, 4, counts=True)
        b.step(sc.dt, 150, 4)
        data = b.checkpoint()
        with World(precision=precision) as r:
            r.restore(data)
            assert r.n_bodies == sc.n_bodies and r.n_worlds == sc.n_worlds
            cr = r.step(sc.dt, 150, 4, counts=True)
            assert np.array_equal(ca[150:], cr)
            for k in ALL:
                assert same(a.get(k), r.get(k)), k
            # a checkpoint of the restored world is the same byte string
            assert r.checkpoint() != data  # it has stepped on
        assert b.checkpoint() == data


def test_checkpoint_file_round_trip(tmp_path):
    sc = _mixed_scene()
    with _world(sc, "f64") as w:
        w.step(sc.dt, 60, 4)
        p = tmp_path / "state.cbz"
        n = checkpoint.save_state(w, p)
        h = checkpoint.header(p.read_bytes())
        assert h["n_bodies"] == sc.n_bodies and h["n_worlds"] == sc.n_worlds and h["version"] == 1
        assert n == p.stat().st_size
        with checkpoint.load_state(p) as r:
            assert r.checkpoint() == w.checkpoint()
        with checkpoint.load_state(p, precision="f32") as r32:
            assert np.array_equal(r32.get("pos"), w.get("pos").astype(np.float32).astype(np.float64))


def test_damaged_checkpoint_is_rejected_and_leaves_the_context_alone():
    sc = _mixed_scene()
    with _world(sc, "f64") as w, World(scenes.cubedrop(0)) as other:
        w.step(sc.dt, 30, 4)
        data = bytearray(w.checkpoint())
        before = other.checkpoint()
        flipped = bytearray(data)
        flipped[len(flipped) // 2] ^= 0x10
        for bad in (bytes(flipped), bytes(data[:-8]), bytes(data[:40]), b"XXXXXXXX" + bytes(data[8:])):
            with pytest.raises(CubezError):
                other.restore(bad)
            assert other.checkpoint() == before
        other.step(scenes.cubedrop(0).dt, 5, 4)  # still usable


def test_recorded_trajectory_reproduces_and_round_trips(tmp_path):
    sc = scenes.batched(2, 4, 8, 64)
    runs = []
    for _ in range(2):
        with World(sc) as w:
            runs.append(checkpoint.record_trajectory(w, sc.dt, 120, every=30))
    d = checkpoint.compare_trajectories(*runs)
    assert all(v == 0.0 for v in d.values()), d
    assert runs[0]["pos"].shape == (4, sc.n_bodies, 3) and runs[0]["counts"].shape == (120, sc.n_worlds)
    p = tmp_path / "traj.npz"
    checkpoint.save_trajectory(p, runs[0])
    back = checkpoint.load_trajectory(p)
    assert all(v == 0.0 for v in checkpoint.compare_trajectories(runs[0], back).values())
