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
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
