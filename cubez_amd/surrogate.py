"""The learned resolver surrogate: data collection, training and use.

The network (40 -> 256 -> 256 -> 256 -> 13, SiLU, BF16 weights on the
matrix cores) is trained on the GPU by the library's own forward and
backward kernels (cbz_sur_train) on contacts the exact resolver generates
(cbz_sur_collect).  Per contact it predicts the velocity phase's changes of
the contact's two bodies and the penetration the position phase resolves at
the contact.  The exact resolver remains the judge at run time: in each
phase, from the predicted state its worst-first loop runs until every
contact of the world is within the phase's epsilon plus a tolerance
(positions: DEFAULT_POS_TOL; velocities: DEFAULT_TOL), for at most
DEFAULT_POLISH * C iterations; a world that does not get there is restored
and re-solved by that plain phase.

    python -m cubez_amd.surrogate train --out cubez_amd/models/surrogate.npz
"""
import argparse
import os
import time

import numpy as np

from . import scenes
from .world import World

LAYERS = ((40, 256), (256, 256), (256, 256), (256, 13))
N_OUT = LAYERS[-1][1]
NORM_OUT = slice(80, 80 + 2 * N_OUT)  # out_mean, out_std in the normalisation vector
DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models", "surrogate.npz")
# residual tolerance: 1 mm of extra penetration over one step of 1/60 s
DEFAULT_TOL = 1e-3 * 60.0
# the position phase's, above positionEpsilon: none, the plain phase's own
# stopping rule.  The plain phase drives every contact it selects to zero
# penetration and leaves the rest below positionEpsilon; a polish that stops
# at positionEpsilon + 1 mm leaves partly resolved contacts up to there,
# 2.5 mm more penetration on average in the settled regime of the
# 4096-world workload, 0.05 mm at 0 (with the damping below)
DEFAULT_POS_TOL = 0.0
# polish budget per contact (iterations of the exact loop after the network):
# the plain phase's own budget.  A world the loop does not bring within the
# tolerance in that many iterations is re-solved by the plain phase, so a
# smaller budget gives up on worlds the loop would still have finished: over
# the 600 steps of the 4096-world workload, 1 * C leaves 1.5 % of world-steps
# to the plain phase (3.87 ms/step), 4 * C 0.47 % (3.50 ms/step)
DEFAULT_POLISH = 4
# the predictions are a warm start for the exact loop, scaled by this factor:
# a network that overshoots adds kinetic energy, which the acceptance test
# rejects; 0.8 halves the fallbacks on the batched-worlds workload
DEFAULT_DAMPING = 0.8
# the same for the position output.  Moves summed over a body's contacts
# overshoot more easily (a lifted cube separates its other contacts), and a
# world whose prediction over-separates a contact is rejected at once; a
# smaller factor rejects fewer worlds but leaves contacts partly resolved,
# below positionEpsilon, where the plain phase would have driven them to
# zero.  At 0.9 about 6 % of world-steps take the plain position phase over
# the 600 steps of the 4096-world workload, and the penetration left is that
# of the exact phase (5.72 against 5.82 mm mean)
DEFAULT_POS_DAMPING = 0.9


def unpack(params):
    """Flat parameters -> [(W [out x in], b [out]) per layer]."""
    out, off = [], 0
    for n_in, n_out in LAYERS:
        W = params[off:off + n_in * n_out].reshape(n_out, n_in)
        off += n_in * n_out
        b = params[off:off + n_out]
        off += n_out
        out.append((W, b))
    assert off == len(params)
    return out


def collect(w, seeds, n_worlds=32, at=(300, 600), iter_per_contact=4, reserve=3_200_000, make=None, log=None):
    """Run the worlds of every seed together as one batched scene (64 seeds
    x 32 worlds at steps 300 and 600 give about 2 M rows) and log the exact
    resolver's contacts at the given steps.  A snapshot that would overflow
    the reserved rows is skipped.  Returns the dataset size."""
    make = make or (lambda s: scenes.batched(s, n_worlds))
    sc = scenes.concat([make(s) for s in seeds])
    w.sur_reserve(reserve)
    w.load(sc)
    done = 0
    for k in sorted(at):
        if k - 1 > done:
            w.step(sc.dt, k - 1 - done, iter_per_contact)
        w.integrate(sc.dt)
        w.narrowphase()
        rows = w.sur_dataset(download=False)
        if int(w.contact_counts().sum()) + rows <= reserve:
            w.sur_collect(sc.dt, iter_per_contact)
        done = k
        if log:
            log(f"collect: step {k} of {sc.n_worlds} worlds, {w.sur_dataset(download=False)} rows")
    return w.sur_dataset(download=False)


def train(w, steps=3000, batch=4096, lr=1e-3, seed=0, init=True, log=None):
    """Adam on the dataset: a short warm-up at 0.3 lr, the main rate, then
    the last quarter at 0.3 and 0.1 lr (run in chunks of at most 1000 steps)."""
    if init:
        w.sur_init(seed)
    warm = min(200, steps // 10)
    tail = (steps - warm) // 4
    plan = ((warm, 0.3), (steps - warm - 2 * tail, 1.0), (tail, 0.3), (tail, 0.1))
    losses, chunk = [], 0
    for s, f in plan:
        while s > 0:
            n = min(s, 1000)
            losses.append(w.sur_train(n, batch, lr * f, seed + 7919 * chunk))
            chunk += 1
            s -= n
            if log:
                log(f"train: {sum(len(x) for x in losses)} steps, lr {lr * f:.2e}, loss {losses[-1][-50:].mean():.4f}")
    return np.concatenate(losses) if losses else np.zeros(0)


def save(path, params, norm, meta=None, damping=DEFAULT_DAMPING, pos_damping=DEFAULT_POS_DAMPING):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), norm=np.asarray(norm, np.float32),
             damping=np.float32(damping), pos_damping=np.float32(pos_damping), meta=np.array(repr(meta or {})))


def damp(norm, damping, pos_damping):
    """The normalisation with the output denormalisation (mean and std)
    scaled: the velocity outputs by `damping`, the position output by
    `pos_damping`."""
    norm = np.asarray(norm, np.float32).copy()
    vel = np.arange(N_OUT) != N_OUT - 1
    for base in (80, 80 + N_OUT):  # out_mean, out_std
        norm[base:base + N_OUT][vel] *= np.float32(damping)
        norm[base + N_OUT - 1] *= np.float32(pos_damping)
    return norm


def load(path=DEFAULT_MODEL, damping=None, pos_damping=None):
    """(params, norm) of a saved model, the output denormalisation scaled by
    the model's damping factors (or the ones given)."""
    d = np.load(path, allow_pickle=False)
    if damping is None:
        damping = float(d["damping"]) if "damping" in d.files else 1.0
    if pos_damping is None:
        pos_damping = float(d["pos_damping"]) if "pos_damping" in d.files else damping
    return d["params"], damp(d["norm"], damping, pos_damping)


def validate(candidates, seeds=(2000, 2001, 2002), worlds=512, steps=450, log=None):
    """Index of the candidate (params, norm) whose surrogate steps are
    fastest on validation worlds that no training run uses: steps 0..steps-1
    from the initial state (the landing, the piles forming and settling);
    with each candidate's ms/step."""
    sc = scenes.concat([scenes.batched(s, worlds) for s in seeds])
    ms = []
    for params, norm in candidates:
        with World(sc, precision="f32") as w:
            w.sur_set_norm(norm)
            w.sur_set_params(params)
            w.synchronize()
            t0 = time.perf_counter()
            w.step_sur(sc.dt, steps, sc.iter_mult, DEFAULT_TOL, DEFAULT_POLISH)
            w.synchronize()
            ms.append(1000.0 * (time.perf_counter() - t0) / steps)
        if log:
            log(f"validate: candidate {len(ms) - 1}: {ms[-1]:.3f} ms/step")
    return int(np.argmin(ms)), ms


def install(w, path=DEFAULT_MODEL):
    params, norm = load(path)
    w.sur_set_norm(norm)
    w.sur_set_params(params)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="collect exact-resolver contacts and train on the GPU")
    t.add_argument("--out", default=DEFAULT_MODEL)
    t.add_argument("--seeds", type=int, default=64)
    t.add_argument("--worlds", type=int, default=32)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--batch", type=int, default=4096)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--candidates", type=int, default=8,
                   help="training runs (initialisation and batch seeds 0..N-1); the fastest on validation worlds is kept")
    a = ap.parse_args()
    if a.cmd == "train":
        def log(m):
            print(m, flush=True)

        runs = []
        with World(scenes.batched(0, 2)) as w:
            t0 = time.perf_counter()
            n = collect(w, range(a.seeds), n_worlds=a.worlds, log=log)
            t1 = time.perf_counter()
            for k in range(max(a.candidates, 1)):
                losses = train(w, a.steps, a.batch, a.lr, seed=k, log=log)
                runs.append((w.sur_get_params().copy(), w.sur_get_norm().copy(), losses))
            t2 = time.perf_counter()
        best, ms = validate([(p, damp(nm, DEFAULT_DAMPING, DEFAULT_POS_DAMPING)) for p, nm, _ in runs], log=log)
        params, norm, losses = runs[best]
        meta = {"rows": n, "seeds": a.seeds, "worlds_per_seed": a.worlds, "steps": a.steps,
                "batch": a.batch, "lr": a.lr, "loss_first": float(losses[:50].mean()),
                "loss_last": float(losses[-50:].mean()), "collect_s": t1 - t0, "train_s": t2 - t1,
                "candidates": len(runs), "kept": best, "validation_ms_per_step": ms}
        save(a.out, params, norm, meta)
        print(meta, flush=True)


if __name__ == "__main__":
    main()
