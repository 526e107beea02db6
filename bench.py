#!/usr/bin/env python3
"""Benchmark of the cubez simulation step on one MI355X.

Workload: 4096 independent worlds, each a ground plane and n ~ U{8..256}
cubes (the batched-worlds configuration), advanced with the reference's
step -- Integrate + CalculateDerivedData, the all-pairs narrowphase,
ResolveContacts(4 x contacts) -- for K steps at dt = 1/60.  A step is one
pass of that hot path over all worlds.  The metric is body-steps/s (bodies x
steps / seconds) with the state resident in HBM.

Prints one JSON line.  Besides the metric it reports:
  roofline      K1 (integrate + derived data, HBM-bound) from its average
                launch time measured with HIP events on the library's stream
  kernels       per-kernel average ms of the timed steps
  cpu_baseline  the FP64 oracle (a C restatement of the Go reference) on a
                bounded sample of the same worlds, on this host's cores
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s, 4096 worlds: surrogate vs exact GPU resolver at ≤1e-3 m error"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of K1 per body-step: an integrated body reads 29 reals
# (pos 3, quat 4, vel 3, rot 3, acc 3, inv_mass 1, iit_body 9, motion 1, the
# two Pow factors) + its flag byte and writes 50 reals (pos, quat, vel, rot,
# last_acc, transform 12, iit_world 9, collider transform 12, motion) + the
# flag byte; a sleeping body reads its flag byte and transform (12) and
# writes the collider transform (12).  Reals are 8 bytes (FP64) or 4 (FP32).


_SCENE = None


def _cpu_world(args):
    """Oracle run of one world (worker process; the scene is inherited)."""
    from oracle.cbz_oracle import OracleWorld

    w, steps = args
    sc = _SCENE
    o = OracleWorld(sc, w)
    o.calculate_derived()
    t0 = time.perf_counter()
    o.step(sc.dt, steps, sc.iter_mult)
    return o.n * steps, time.perf_counter() - t0


def cpu_baseline(scene, steps, budget_s):
    """Time the oracle on an evenly spaced sample of worlds, all steps, with
    one process per host core (up to 16)."""
    import multiprocessing as mp

    from oracle import cbz_oracle

    global _SCENE
    _SCENE = scene
    n_worlds = scene.n_worlds
    cbz_oracle.load()
    cores = max(1, min(16, os.cpu_count() or 1))
    # a first world gives the per-body-step cost to size the sample
    probe = _cpu_world((0, steps))
    per_bs = probe[1] / max(probe[0], 1)
    mean_n = 132.0
    n_sample = int(max(cores, min(n_worlds, budget_s * cores / max(per_bs * mean_n * steps, 1e-9))))
    stride = max(1, n_worlds // n_sample)
    worlds = list(range(0, n_worlds, stride))[:n_sample]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_world, [(w, steps) for w in worlds])
    wall = time.perf_counter() - t0
    bs = sum(r[0] for r in res)
    return {
        "value": bs / wall,
        "unit": "body-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(worlds)} of {n_worlds} worlds (every {stride}th), all {steps} steps, "
                  f"{bs} body-steps in {wall:.1f} s wall on {cores} processes",
    }


def run_path(World, sc, precision, steps, warmup):
    """Warm up, then time `steps` steps from the initial state (graph
    replay), then the same steps again with HIP events between kernels."""
    w = World(sc, precision=precision)
    if warmup > 0:
        w.step(sc.dt, warmup, sc.iter_mult)
    w.load(sc)
    w.synchronize()
    t0 = time.perf_counter()
    w.step(sc.dt, steps, sc.iter_mult)
    w.synchronize()
    wall = time.perf_counter() - t0
    w.load(sc)
    w.set_profiling(True)
    w.step(sc.dt, steps, sc.iter_mult)
    w.set_profiling(False)
    ms, cnt = w.phase_times()
    w.close()
    sp, integrated, contacts, iters = (int(x) for x in cnt)
    names = ("integrate", "np_count", "np_scans", "np_emit", "resolve", "step_profiled")
    return {
        "value": sc.n_bodies * steps / wall,
        "ms_per_step": 1000.0 * wall / steps,
        "kernels": {k: round(ms[i] / sp, 5) for i, k in enumerate(names)},
        "steps_profiled": sp,
        "integrated": integrated,
        "contacts_per_step": contacts / sp,
        "contacts_per_s": contacts / (ms[5] / 1000.0) if ms[5] > 0 else None,
        "iters_per_step": iters / sp,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", choices=("f64", "f32"), default="f32",
                    help="state precision of the timed path (the configuration names FP32)")
    ap.add_argument("--single", action="store_true", help="skip the other precision's companion run")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle wall time")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: worlds are independent, so a multi-GPU run would shard "
              "them with no exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cubez_amd import World, scenes

    sc = scenes.batched(args.seed, args.worlds)
    nb = sc.n_bodies
    main_path = run_path(World, sc, args.precision, args.steps, args.warmup)
    other = "f64" if args.precision == "f32" else "f32"
    side = run_path(World, sc, other, args.steps, args.warmup) if not args.single else None

    kern = main_path["kernels"]
    rs = 8 if args.precision == "f64" else 4
    # K1 roofline: algorithmic bytes per launch / average launch time
    k1_awake = 29 * rs + 1 + 50 * rs + 1
    k1_sleep = 1 + 12 * rs + 12 * rs
    steps_p = main_path["steps_profiled"]
    integrated = main_path["integrated"]
    k1_bytes = (integrated * k1_awake + (nb * steps_p - integrated) * k1_sleep) / steps_p
    achieved = k1_bytes / (kern["integrate"] / 1000.0) / 1e9
    out = {
        "metric": METRIC,
        "value": main_path["value"],
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_path["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": f"batched worlds: {args.worlds} worlds, plane + n~U{{8..256}} cubes each "
                        f"({nb} bodies), first {args.steps} steps at dt=1/60, all-pairs narrowphase, "
                        f"exact resolver with maxIterations=4*C, "
                        + ("FP32 state with FP64 SAT and contact basis" if args.precision == "f32"
                           else "FP64 throughout"),
            "worlds": args.worlds, "bodies": nb, "seed": args.seed, "resolver": "exact",
            "precision": args.precision, "step": "hipGraph replay of one captured step",
        },
        "contacts_per_s": main_path["contacts_per_s"],
        "contacts_per_step": main_path["contacts_per_step"],
        "resolver_iterations_per_step": main_path["iters_per_step"],
        "kernels_ms_per_step": kern,
        "dominant_kernel": max(("integrate", "np_count", "np_scans", "np_emit", "resolve"), key=lambda k: kern[k]),
        "roofline": {
            "kernel": "k_integrate (K1: Integrate + CalculateDerivedData)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "bytes_per_launch": k1_bytes,
        },
    }
    if side is not None:
        out["plain_" + other] = {k: side[k] for k in ("value", "ms_per_step", "kernels")}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, args.steps, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
