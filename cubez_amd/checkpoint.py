"""State checkpoints and recorded trajectories on disk.

save_state / load_state write and read the C ABI's state checkpoint
(cbz_checkpoint_save / cbz_checkpoint_load: a self-describing byte string
with a payload hash) as a file.  A World restored from it steps on exactly
as the saved one would have.

record_trajectory steps a World and keeps chosen fields every few steps,
with the per-step contact counts; save_trajectory / load_trajectory store
such a record as an .npz of plain arrays (read back without pickle).  Two
runs of the exact path on any MI355X give identical records, so a record
made on one machine is a parity fixture for another.
"""
import struct

import numpy as np

from .world import World

HEADER = struct.Struct("<8sIIiiqiiiiQQ")  # the 64-byte header of cbz_checkpoint_save
MAGIC = b"CBZSTAT1"


def header(data):
    """The checkpoint header as a dict (no validation beyond the magic)."""
    if len(data) < HEADER.size:
        raise ValueError("not a state checkpoint: too short")
    f = HEADER.unpack_from(data, 0)
    if f[0] != MAGIC:
        raise ValueError("not a state checkpoint: bad magic")
    keys = ("magic", "version", "header_bytes", "n_worlds", "n_planes", "n_bodies", "material_flags",
            "n_fields", "precision", "zero", "payload_bytes", "hash")
    return dict(zip(keys, f))


def save_state(world, path):
    data = world.checkpoint()
    with open(path, "wb") as fh:
        fh.write(data)
    return len(data)


def load_state(path, precision="f64", device=0, contact_capacity=0):
    """A new World holding the checkpointed state (in the given precision;
    an FP64 state loaded as FP32 is rounded once)."""
    with open(path, "rb") as fh:
        data = fh.read()
    w = World(precision=precision, device=device, contact_capacity=contact_capacity)
    try:
        w.restore(data)
    except Exception:
        w.close()
        raise
    return w


def record_trajectory(world, dt, n_steps, iter_per_contact=4, every=1, fields=("pos", "quat", "vel", "rot")):
    """Step n_steps and keep `fields` after every `every` steps: arrays
    [n_frames, n_bodies, width], plus counts [n_steps, n_worlds]."""
    frames = {k: [] for k in fields}
    counts = []
    s = 0
    while s < n_steps:
        k = min(every, n_steps - s)
        counts.append(world.step(dt, k, iter_per_contact, counts=True))
        for f in fields:
            frames[f].append(world.get(f))
        s += k
    out = {"dt": np.float64(dt), "every": np.int64(every), "iter_per_contact": np.int64(iter_per_contact),
           "counts": np.concatenate(counts) if counts else np.zeros((0, world.n_worlds), np.int64)}
    for f in fields:
        out[f] = np.stack(frames[f]) if frames[f] else np.zeros((0,))
    return out


def save_trajectory(path, traj):
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in traj.items()})


def load_trajectory(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def compare_trajectories(a, b):
    """Per shared array: the largest absolute difference (0 for identical,
    NaN at matching positions counting as equal); 'counts' compares exactly."""
    out = {}
    for k in sorted(set(a) & set(b)):
        x, y = np.atleast_1d(np.asarray(a[k])), np.atleast_1d(np.asarray(b[k]))
        if x.shape != y.shape:
            out[k] = float("inf")
        elif x.dtype.kind == "f":
            nx, ny = np.isnan(x), np.isnan(y)
            if np.any(nx != ny):
                out[k] = float("inf")
            else:
                d = np.where(nx, 0.0, np.abs(np.where(nx, 0.0, x) - np.where(ny, 0.0, y)))
                out[k] = float(d.max(initial=0.0))
        else:
            out[k] = 0.0 if np.array_equal(x, y) else float("inf")
    return out
