"""Evaluations of the learned resolver beyond the benchmark line.

    python -m cubez_amd.evaluate surrogate   # fresh seeds 1000-1015, batched worlds
    python -m cubez_amd.evaluate towers      # stacked resting contacts, 16384 worlds
    python -m cubez_amd.evaluate broadphase  # one large world: all pairs vs sort-and-sweep

surrogate: 16 fresh seeds x 256 batched worlds (4096 worlds), 600 steps at
  1/60.  Held-out accuracy of the network on the exact resolver's contacts of
  those seeds (logged at steps 300 and 600), then the exact FP64 and FP32
  resolvers, the surrogate with exact fallback and the polish from a zero
  prediction (the method without the network) run side by side: time,
  fallback rate, velocity-phase iterations and the penetration left after the
  last step.
towers: 4-16 towers of 4-12 cubes per world, 16384 worlds, 1200 steps at
  1/120: the same paths; fallback rate, penetration left, and the drift of
  every tower's top cube, from its start and against the FP64 exact run.
broadphase: one world of n cubes (n = 4000, 16000, 40000) spread over
  x, z in +-2 sqrt(n): the all-pairs narrowphase against the broadphase
  (cbz_narrowphase_broad) on the same state, wall time per call (mean of 5,
  host round trips included), the pairs each checks, and that both give the
  same contact count.

Each prints one JSON object."""
import argparse
import json
import sys
import time

import numpy as np

from . import scenes, surrogate
from .world import World


def _penetration(w, n_worlds):
    """Contacts and their penetration in the current state (a narrowphase
    without integration): count, mean, 99th percentile and max."""
    w.narrowphase()
    counts = w.contact_counts()
    pens = [w.contacts(k, internals=False)["pen"] for k in range(n_worlds) if counts[k] > 0]
    p = np.concatenate(pens) if pens else np.zeros(0)
    if not len(p):
        return {"contacts": 0}
    return {"contacts": int(len(p)), "mean_m": float(p.mean()), "p99_m": float(np.percentile(p, 99)),
            "max_m": float(p.max())}


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _run(sc, mode, model, steps, tol, polish):
    """mode: exact_f64, exact_f32, surrogate, zero_prediction."""
    _log(f"evaluate: {mode} on {sc.name or 'scene'}, {sc.n_worlds} worlds, {steps} steps")
    prec = "f64" if mode == "exact_f64" else "f32"
    with World(sc, precision=prec) as w:
        if mode in ("surrogate", "zero_prediction"):
            w.sur_set_norm(model[1])
            w.sur_set_params(model[0])
        w.synchronize()
        t0 = time.perf_counter()
        fb = None
        if mode.startswith("exact"):
            w.step(sc.dt, steps, sc.iter_mult)
        else:
            w.set_profiling(True)
            fb = w.step_sur(sc.dt, steps, sc.iter_mult, tol, polish, mode == "zero_prediction")
            w.set_profiling(False)
        w.synchronize()
        wall = time.perf_counter() - t0
        out = {"seconds": wall, "body_steps_per_s": sc.n_bodies * steps / wall}
        if fb is not None:
            ms, cnt = w.sur_phase_times()
            out["fallback_rate"] = float(fb.sum()) / (steps * sc.n_worlds)
            out["velocity_iterations_per_step"] = float(cnt[3]) / steps
            out["position_iterations_per_step"] = float(cnt[5]) / steps
            out["position_fallback_rate"] = float(cnt[4]) / (steps * sc.n_worlds)
            out["note"] = "timed with per-step profiling events (the surrogate step waits once per step anyway)"
        pos = w.get("pos").copy()
        out["penetration_after"] = _penetration(w, sc.n_worlds)
    return out, pos


def eval_surrogate(model_path=None, seeds=range(1000, 1016), worlds_per_seed=256, steps=600,
                   tol=surrogate.DEFAULT_TOL, polish=surrogate.DEFAULT_POLISH):
    model = surrogate.load(model_path or surrogate.DEFAULT_MODEL)
    sc = scenes.concat([scenes.batched(s, worlds_per_seed) for s in seeds])
    res = {"workload": f"{len(list(seeds))} fresh seeds x {worlds_per_seed} batched worlds "
                       f"({sc.n_worlds} worlds, {sc.n_bodies} bodies), {steps} steps at 1/60",
           "tolerance_m_per_s": tol, "polish_per_contact": polish}
    # held-out accuracy on the exact resolver's contacts of these seeds
    _log("evaluate: collecting held-out contacts")
    with World(scenes.batched(0, 2)) as w:
        rows = surrogate.collect(w, seeds, n_worlds=worlds_per_seed, at=(300, 600), reserve=4_000_000)
        X, Y = w.sur_dataset()
        w.sur_set_norm(model[1])
        w.sur_set_params(model[0])
        P = np.concatenate([w.sur_predict(X[k:k + 262144]) for k in range(0, len(X), 262144)])
    names = ("dv_body0", "dw_body0", "dv_body1", "dw_body1")
    r2 = {}
    for g, nm in enumerate(names):
        y, p = Y[:, 3 * g:3 * g + 3].astype(np.float64), P[:, 3 * g:3 * g + 3].astype(np.float64)
        r2[nm] = float(1.0 - ((y - p) ** 2).sum() / max(((y - y.mean(axis=0)) ** 2).sum(), 1e-30))
    res["held_out"] = {"contacts": int(rows), "r2": r2}
    for mode in ("exact_f64", "exact_f32", "surrogate", "zero_prediction"):
        res[mode], _ = _run(sc, mode, model, steps, tol, polish)
    res["speedup_vs_exact_f64"] = res["surrogate"]["body_steps_per_s"] / res["exact_f64"]["body_steps_per_s"]
    return res


def eval_towers(model_path=None, seed=0, n_worlds=16384, steps=None, tol=surrogate.DEFAULT_TOL,
                polish=surrogate.DEFAULT_POLISH):
    model = surrogate.load(model_path or surrogate.DEFAULT_MODEL)
    sc = scenes.towers(seed, n_worlds)
    steps = steps or sc.steps
    tops = sc.meta["tower_top"]
    p0 = sc.fields["pos"][tops]
    res = {"workload": f"towers: {n_worlds} worlds, {len(tops)} towers, {sc.n_bodies} bodies, {steps} steps "
                       f"at 1/120", "tolerance_m_per_s": tol, "polish_per_contact": polish}
    final = {}
    for mode in ("exact_f64", "exact_f32", "surrogate"):
        res[mode], pos = _run(sc, mode, model, steps, tol, polish)
        final[mode] = pos[tops]
        d = np.linalg.norm(final[mode] - p0, axis=1)
        res[mode]["top_drift_from_start_m"] = {"mean": float(d.mean()), "p99": float(np.percentile(d, 99)),
                                               "max": float(d.max())}
    for mode in ("exact_f32", "surrogate"):
        d = np.linalg.norm(final[mode] - final["exact_f64"], axis=1)
        res[mode]["top_vs_exact_f64_m"] = {"mean": float(d.mean()), "p99": float(np.percentile(d, 99)),
                                           "max": float(d.max())}
    res["speedup_vs_exact_f64"] = res["surrogate"]["body_steps_per_s"] / res["exact_f64"]["body_steps_per_s"]
    return res


def eval_broadphase(sizes=(4000, 16000, 40000), settle_steps=60, reps=5):
    """All pairs against sort-and-sweep on one large world.  The smallest
    world is first stepped until its cubes land (so it has contacts); the
    larger ones are checked as generated."""
    res = {"worlds": []}
    for n in sizes:
        sc = scenes.batched(4, 1, n, n)
        with World(sc) as w:
            if n <= 4000:
                w.step(sc.dt, settle_steps, sc.iter_mult)
            w.integrate(sc.dt)
            w.narrowphase_broad()  # sizes the broadphase buffers
            w.narrowphase()
            w.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                total_all = w.narrowphase()
            w.synchronize()
            t_all = (time.perf_counter() - t0) / reps
            t0 = time.perf_counter()
            for _ in range(reps):
                total_bp, checked = w.narrowphase_broad()
            w.synchronize()
            t_bp = (time.perf_counter() - t0) / reps
        res["worlds"].append({"bodies": n, "all_pairs_ms": 1000 * t_all, "broadphase_ms": 1000 * t_bp,
                              "pairs_all": n * n, "pairs_checked": int(checked), "contacts": int(total_all),
                              "same_contact_count": bool(total_all == total_bp)})
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("which", choices=("surrogate", "towers", "broadphase"))
    ap.add_argument("--model", default=None)
    ap.add_argument("--worlds", type=int, default=None)
    ap.add_argument("--steps", type=int, default=None)
    a = ap.parse_args()
    if a.which == "surrogate":
        kw = {}
        if a.worlds:
            kw["worlds_per_seed"] = a.worlds
        if a.steps:
            kw["steps"] = a.steps
        out = eval_surrogate(a.model, **kw)
    elif a.which == "towers":
        out = eval_towers(a.model, n_worlds=a.worlds or 16384, steps=a.steps)
    else:
        out = eval_broadphase()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
