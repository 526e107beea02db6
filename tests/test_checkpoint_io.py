"""Host-side checkpoint and trajectory file handling (no GPU)."""
import struct

import numpy as np
import pytest

from cubez_amd import checkpoint


def test_header_layout_matches_the_c_abi():
    # the C header is 64 bytes: magic, u32 version, u32 header size, i32 worlds,
    # i32 planes, i64 bodies, i32 flags, i32 fields, i32 precision, i32 0,
    # u64 payload bytes, u64 hash
    assert checkpoint.HEADER.size == 64
    raw = struct.pack("<8sIIiiqiiiiQQ", b"CBZSTAT1", 1, 64, 3, 2, 77, 1, 20, 0, 0, 1234, 99)
    h = checkpoint.header(raw + b"\0" * 16)
    assert (h["n_worlds"], h["n_planes"], h["n_bodies"], h["material_flags"], h["payload_bytes"]) == (3, 2, 77, 1,
                                                                                                       1234)
    with pytest.raises(ValueError):
        checkpoint.header(b"CBZSTAT0" + raw[8:])
    with pytest.raises(ValueError):
        checkpoint.header(raw[:20])


def test_trajectory_files_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    t = {"dt": np.float64(1 / 60), "every": np.int64(5), "iter_per_contact": np.int64(4),
         "counts": rng.integers(0, 100, (10, 3)), "pos": rng.normal(size=(2, 7, 3))}
    p = tmp_path / "t.npz"
    checkpoint.save_trajectory(p, t)
    back = checkpoint.load_trajectory(p)
    assert set(back) == set(t)
    assert all(v == 0.0 for v in checkpoint.compare_trajectories(t, back).values())
    u = dict(back)
    u["pos"] = u["pos"].copy()
    u["pos"][1, 2, 0] += 1e-9
    u["counts"] = u["counts"].copy()
    u["counts"][0, 0] += 1
    d = checkpoint.compare_trajectories(t, u)
    assert d["pos"] == pytest.approx(1e-9, rel=1e-3) and d["counts"] == float("inf")
