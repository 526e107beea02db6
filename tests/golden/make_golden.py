"""Writes tests/golden/golden.npz from the FP64 oracle (test infrastructure).

Fixtures (all from seeded scenes of cubez_amd.scenes):
  cubedrop  seeds 0-7, 64 cubes on a plane, 600 steps at 1/60,
            maxIterations = 4 x contacts: per-step contact counts, SHA-256 of
            the packed FP64 state (pos, quat, vel, rot) after steps 1, 60
            and 600, and seed 0's full state after steps 1 and 60
  ballistic seed 0 at 48 spheres + 48 cubes, 240 steps: counts and hashes
            after steps 1, 60 and 240
  batched   seed 3, 12 worlds of 8-64 cubes, 240 steps: per-world per-step
            counts and per-world hashes after step 240
  config1   the ballistic mix at full size (512 spheres + 512 cubes, seed
            0), 300 steps: counts and hashes after steps 1, 60, 120, 180,
            240 and 300
  config2   8 of the 4096 batched worlds (seed 0: the smallest, the largest
            and six evenly spaced), 600 steps: per-world per-step counts and
            hashes after steps 1, 60 and 600
  config4   every 64th of the 16384 stacked-tower worlds (seed 0), 240 steps
            at 1/120: per-world per-step counts and hashes after steps 120
            and 240

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cubez_amd import scenes  # noqa: E402
from oracle.cbz_oracle import OracleWorld  # noqa: E402

STATE = ("pos", "quat", "vel", "rot")


def state_hash(get):
    h = hashlib.sha256()
    for k in STATE:
        h.update(np.ascontiguousarray(get(k), "<f8").tobytes())
    return h.hexdigest()


def run(sc, world, steps, marks, keep=()):
    o = OracleWorld(sc, world)
    o.calculate_derived()
    counts, hashes, states, done = [], {}, {}, 0
    for k in sorted(set(marks) | set(keep) | {steps}):
        counts.append(o.step(sc.dt, k - done, sc.iter_mult))
        done = k
        if k in marks:
            hashes[k] = state_hash(o.get)
        if k in keep:
            states[k] = {f: o.get(f).copy() for f in STATE}
    return np.concatenate(counts).astype(np.int32), hashes, states


CONFIG1_MARKS = (1, 60, 120, 180, 240, 300)
CONFIG2_MARKS = (1, 60, 600)
CONFIG4_MARKS = (120, 240)


def config2_worlds(sc):
    """The smallest, the largest and six evenly spaced worlds."""
    n = np.diff(sc.body_off)
    pick = [int(np.argmin(n)), int(np.argmax(n))] + [int(w) for w in np.linspace(0, sc.n_worlds - 1, 8)[1:7]]
    return sorted(set(pick))


def config4_scene():
    return scenes.towers(0, 16384).subset(range(0, 16384, 64))


_POOL_SCENE = None


def _pool_run(args):
    w, steps, marks = args
    c, h, _ = run(_POOL_SCENE, w, steps, marks)
    return c, [h[k] for k in marks]


def run_worlds(sc, worlds, steps, marks):
    """run() over several worlds of one scene in worker processes."""
    import multiprocessing as mp

    global _POOL_SCENE
    _POOL_SCENE = sc
    with mp.get_context("fork").Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(_pool_run, [(w, steps, marks) for w in worlds])
    return np.stack([r[0] for r in res], axis=1), np.array([r[1] for r in res])


def main():
    out = {}
    cd_counts, cd_hash = [], []
    for seed in range(8):
        sc = scenes.cubedrop(seed)
        c, h, st = run(sc, 0, 600, (1, 60, 600), keep=(1, 60) if seed == 0 else ())
        cd_counts.append(c)
        cd_hash.append([h[1], h[60], h[600]])
        for k, d in st.items():
            for f, a in d.items():
                out[f"cubedrop0_step{k}_{f}"] = a
    out["cubedrop_counts"] = np.stack(cd_counts)
    out["cubedrop_hash"] = np.array(cd_hash)
    out["hash_steps"] = np.array([1, 60, 600])

    sc = scenes.ballistic(0, 48, 48)
    c, h, _ = run(sc, 0, 240, (1, 60, 240))
    out["ballistic_counts"] = c
    out["ballistic_hash"] = np.array([h[1], h[60], h[240]])

    sc = scenes.batched(3, 12, 8, 64)
    bc, bh = [], []
    for w in range(sc.n_worlds):
        c, h, _ = run(sc, w, 240, (240,))
        bc.append(c)
        bh.append(h[240])
    out["batched_counts"] = np.stack(bc, axis=1)  # [steps, worlds]
    out["batched_hash"] = np.array(bh)

    sc = scenes.ballistic(0, 512, 512)
    c, h, _ = run(sc, 0, 300, CONFIG1_MARKS)
    out["config1_counts"] = c
    out["config1_hash"] = np.array([h[k] for k in CONFIG1_MARKS])

    sc = scenes.batched(0, 4096)
    ws = config2_worlds(sc)
    out["config2_worlds"] = np.array(ws, np.int32)
    out["config2_counts"], out["config2_hash"] = run_worlds(sc, ws, 600, CONFIG2_MARKS)

    sc = config4_scene()
    counts, hashes = run_worlds(sc, range(sc.n_worlds), 240, CONFIG4_MARKS)
    out["config4_counts"] = counts.astype(np.int16)
    out["config4_hash"] = hashes
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
