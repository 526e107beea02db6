#!/usr/bin/env python3
"""Benchmark of the cubez simulation step on one MI355X.

Workload: 4096 independent worlds, each a ground plane and n ~ U{8..256}
cubes (the batched-worlds configuration), advanced with the reference's
step -- Integrate + CalculateDerivedData, the all-pairs narrowphase,
ResolveContacts(4 x contacts) -- at dt = 1/60.  A step is one pass of that
hot path over all worlds.  The metric is body-steps/s (bodies x steps /
seconds) with the state resident in HBM, for the learned-surrogate resolver
with exact fallback (FP32 state; both phases polished by the exact loop,
positions to positionEpsilon, velocities to 0.06 m/s of residual closing
speed = 1 mm of penetration per step), next to the exact FP64 GPU resolver
it is compared with.

The timed window: with --steps 600 (the default) the whole workload, steps
0..599 from the seeded initial state.  With fewer steps, a window of the
settled regime: the exact FP64 step runs to step --window-start (300) in the
untimed region, its state is checkpointed, and every leg (exact FP64, exact
FP32, surrogate, zero prediction) restores that same state, runs --warmup
untimed steps, restores it again and times --steps steps.  The first steps
from the initial state are free fall with 2 % of the settled contact count,
so a short window from step 0 would time a regime the workload barely spends
time in.

Prints one JSON line.  Besides the metric it reports:
  exact_f64, exact_f32   the exact resolver (hipGraph-replayed step)
  zero_prediction        the surrogate's polish started from a zero
                         prediction: the same method without the network
  fallback_rate          worlds re-solved by a plain phase / world-steps
  error                  penetration left after the window (narrowphase of
                         the final state: mean, p99, max) for the surrogate
                         and both exact paths, and the surrogate's final
                         positions against each exact path's
  kernels                HIP-event times per step of every phase
  roofline               the dominant kernel of the surrogate step, its
                         algorithmic bytes over its event-timed duration
  k1, k2, mlp            the rooflines of integrate, the narrowphase and the
                         network beside it
  cpu_baseline           the FP64 oracle (a C restatement of the Go
                         reference) on a bounded sample of the same worlds,
                         on all host cores and on one
  resolver_split         per leg, the split between the one- and the
                         four-wavefront resolver kernels that the autotuner
                         (cbz_autotune_resolver) picked on the settled state,
                         with every candidate's ms/step
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s, 4096 worlds: surrogate vs exact GPU resolver at ≤1e-3 m error"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, FMA counted as two
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA
SAT_AXIS_FLOP = 75
PLANE_CHECK_FLOP = 40
K1_LAUNCHES = 50
DEFAULT_WINDOW_START = 300

# Algorithmic bytes of K1 per body-step: an integrated body reads 38 reals
# (pos 3, quat 4, vel 3, rot 3, acc 3, inv_mass 1, iit_body 9, the old
# iit_world 9 that angularAcceleration = iitWorld * torqueAccum reads, motion
# 1, the two Pow factors) + its flag byte and writes 50 reals (pos, quat, vel,
# rot, last_acc, transform 12, iit_world 9, collider transform 12, motion) +
# the flag byte; a sleeping body reads its flag byte and transform (12) and
# writes the collider transform (12).  Reals are 8 bytes (FP64) or 4 (FP32).
K1_READ_REALS, K1_WRITE_REALS = 38, 50

# Algorithmic bytes of the resolver per resolve, for its roofline: every
# contact record read once (ids 8 B, point 3, normal 3, pen 1) and its
# resolver internals written once (pen 1, c2w 9, rel 6, cvel 3, ddv 1), and
# every body of a world with contacts read once (pos 3, quat 4, vel 3, rot 3,
# inv_mass 1, iit_world 9, last_acc 3, flags 1 B) and written once (pos 3,
# quat 4, vel 3, rot 3, motion 1, flags 1 B).  The serial iterations re-read
# what an iteration changed; those re-reads are not algorithmic.
RES_CONTACT_READ, RES_CONTACT_WRITE = 7, 20
RES_BODY_READ, RES_BODY_WRITE = 26, 14


def traffic_ratio(precision, path=os.path.join(ROOT, "profiles", "k_integrate_hbm_traffic.md")):
    """HBM traffic / algorithmic bytes of K1, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (copy-calibrated), or None."""
    try:
        for line in open(path):
            if line.startswith(f"traffic_over_algorithmic_{precision}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


# --- CPU baseline ------------------------------------------------------------

_SCENE = None


def _cpu_world(args):
    """Oracle run of one world (worker process; the scene is inherited)."""
    from oracle.cbz_oracle import OracleWorld

    w, steps, window_start = args
    sc = _SCENE
    o = OracleWorld(sc, w)
    o.calculate_derived()
    t0 = time.perf_counter()
    first = min(window_start, steps)
    if first:
        o.step(sc.dt, first, sc.iter_mult)
    before = o.counters()
    if steps > first:
        o.step(sc.dt, steps - first, sc.iter_mult)
    after = o.counters()
    # the work counters of the window's steps alone
    return o.n * steps, time.perf_counter() - t0, {k: after[k] - before[k] for k in after}, o.n * (steps - first)


def _host_cpu():
    """nproc and the CPU model of the host the baseline ran on."""
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"nproc {os.cpu_count()}, {model}"


def cpu_baseline(scene, steps, budget_s, window_start=0):
    """Time the oracle on an evenly spaced sample of worlds, all `steps`
    steps from the initial state, with one process per host core (up to 16);
    the single-core rate is the sum of the per-world CPU times.  Also
    returns the oracle's work counters over steps window_start..steps-1 and
    the body-steps they cover."""
    import multiprocessing as mp

    from oracle import cbz_oracle

    global _SCENE
    _SCENE = scene
    n_worlds = scene.n_worlds
    cbz_oracle.load()
    cores = max(1, min(16, os.cpu_count() or 1))
    # a first world gives the per-body-step cost to size the sample
    probe = _cpu_world((0, steps, window_start))
    per_bs = probe[1] / max(probe[0], 1)
    mean_n = 132.0
    n_sample = int(max(cores, min(n_worlds, budget_s * cores / max(per_bs * mean_n * steps, 1e-9))))
    stride = max(1, n_worlds // n_sample)
    worlds = list(range(0, n_worlds, stride))[:n_sample]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        res = pool.map(_cpu_world, [(w, steps, window_start) for w in worlds])
    wall = time.perf_counter() - t0
    bs = sum(r[0] for r in res)
    cpu_s = sum(r[1] for r in res)
    counters = {k: sum(r[2][k] for r in res) for k in res[0][2]} if res else {}
    window_bs = sum(r[3] for r in res)
    return {
        "value": bs / wall,
        "unit": "body-steps/s",
        "cores": cores,
        "value_1core": bs / cpu_s,
        "kind": "port",
        "sample": f"{len(worlds)} of {n_worlds} worlds (every {stride}th), steps 0..{steps - 1} from the initial "
                  f"state, {bs} body-steps in {wall:.1f} s wall on {cores} processes ({cpu_s:.1f} s of CPU time, "
                  f"which gives value_1core); host: {_host_cpu()}",
    }, counters, window_bs


def narrowphase_rate(counters, sample_body_steps, n_bodies, ms_per_step):
    """K2's FP64 VALU rate: the reference's narrowphase work per step, from
    the oracle's op counters on the CPU sample (scaled to the whole scene by
    body-steps), over the FP64 narrowphase kernels' time per step.  Work
    model (the reference's arithmetic in colliders.go:445-477 tryAxis /
    penetrationOnAxis, :762-770 transformToAxis, :308-367 the half-space
    check): 75 flops per SAT axis tested, 40 per body-plane check; contact
    generation is not counted, so the rate is a lower bound."""
    scale = n_bodies / max(sample_body_steps, 1)
    per_step = {k: counters[k] * scale for k in ("pairs", "plane_checks", "sat_axes", "sat_skipped", "contacts")}
    flop = SAT_AXIS_FLOP * per_step["sat_axes"] + PLANE_CHECK_FLOP * per_step["plane_checks"]
    tf = flop / (ms_per_step / 1000.0) / 1e12 if ms_per_step > 0 else 0.0
    return {
        "kernels": "k_np_cull + k_np_eval + scans + k_np_emit_hits (FP64 state)",
        "bound": "fp64-valu", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": tf / FP64_PEAK_TFLOPS, "frac_of_unfused_peak": tf / (FP64_PEAK_TFLOPS / 2),
        "flop_per_step": flop, "ms_per_step": ms_per_step, "reference_work_per_step": per_step,
        "work_window": "the oracle's work counters on the CPU sample over the timed window's steps, scaled "
                       "to the scene by body-steps",
        "flop_model": f"{SAT_AXIS_FLOP} per SAT axis tested + {PLANE_CHECK_FLOP} per body-plane check "
                      "(contact generation not counted: a lower bound)",
    }


# --- GPU legs ----------------------------------------------------------------

def _log(msg):
    print(msg, file=sys.stderr, flush=True)


RESOLVER_SPLIT = "auto"  # --resolver-split: "auto", "default" or "min,max"
TUNE_STEPS = 4  # timed steps per candidate split


def _start(w, sc, ck):
    if ck is None:
        w.load(sc)
    else:
        w.restore(ck)
    if RESOLVER_SPLIT not in ("auto", "default"):
        w.set_resolver_split(*(int(x) for x in RESOLVER_SPLIT.split(",")))
    w.synchronize()


TUNE_CK = None  # the settled state the split is tuned on


def _tune(w, sc, precision, surrogate=False, tol=0.06, polish=4):
    """The resolver split of this leg: measured by the autotuner from the
    settled state (the world's state is then loaded from it; callers
    restore their window's start afterwards), or as given."""
    if RESOLVER_SPLIT != "auto":
        return {"split": list(w.resolver_split()), "tuned": False}
    from cubez_amd import autotune
    if TUNE_CK is not None:
        w.restore(TUNE_CK)
    r = autotune.tune(w, sc.dt, sc.iter_mult, precision, surrogate, steps=TUNE_STEPS, tol=tol, polish=polish,
                      use_cache=False)
    ms = r["ms_per_step"]
    return {"split": list(r["split"]), "tuned": True, "tune_steps": TUNE_STEPS, "ms_per_step": ms,
            "gain_vs_one_wave_only": ms["0,0"] / min(ms.values()) if "0,0" in ms else None}


def residual_penetration(w):
    """Penetration left in the current state: a narrowphase of it (no
    integration), every contact's depth: count, mean, p99 and max (m)."""
    w.narrowphase()
    p = w.contacts(-1, internals=False)["pen"]
    if not len(p):
        return {"contacts": 0}
    return {"contacts": int(len(p)), "mean_m": float(p.mean()), "p99_m": float(np.percentile(p, 99)),
            "max_m": float(p.max())}


def settle(World, sc, start):
    """The exact FP64 step from the initial state to step `start`: the
    checkpoint every leg starts its window from."""
    if start <= 0:
        return None, 0.0
    t0 = time.perf_counter()
    with World(sc, precision="f64") as w:
        w.step(sc.dt, start, sc.iter_mult)
        w.synchronize()
        ck = w.checkpoint()
    return ck, time.perf_counter() - t0


def run_path(World, sc, ck, precision, steps, warmup):
    """Exact resolver: warm up from the window's start, restore it, time
    `steps` steps (graph replay), then the same steps again with HIP events
    between kernels."""
    w = World(sc, precision=precision)
    _start(w, sc, ck)
    split = _tune(w, sc, precision)
    _start(w, sc, ck)
    if warmup > 0:
        w.step(sc.dt, warmup, sc.iter_mult)
        _start(w, sc, ck)
    t0 = time.perf_counter()
    w.step(sc.dt, steps, sc.iter_mult)
    w.synchronize()
    wall = time.perf_counter() - t0
    final_pos = w.get("pos")
    resid = residual_penetration(w)
    _start(w, sc, ck)
    w.set_profiling(True)
    w.step(sc.dt, steps, sc.iter_mult)
    w.set_profiling(False)
    ms, cnt = w.phase_times()
    crit = critical_path(w)
    # K1 alone: back-to-back launches between two events, on the window's
    # start state
    _start(w, sc, ck)
    k1_ms, k1_bodies = w.integrate_timed(sc.dt, K1_LAUNCHES)
    per_cu, n_cu = w.resolver_occupancy()
    w.close()
    sp, integrated, contacts, iters = (int(x) for x in cnt)
    names = ("integrate", "np_count", "np_scans", "np_emit", "resolve", "step_profiled")
    resident = min(per_cu * n_cu, sc.n_worlds)
    resolve_ns = ms[4] / sp * 1e6
    return {
        "value": sc.n_bodies * steps / wall,
        "ms_per_step": 1000.0 * wall / steps,
        "kernels": {k: round(ms[i] / sp, 5) for i, k in enumerate(names)},
        "steps_profiled": sp,
        "integrated_per_step": integrated / sp,
        "contacts_per_step": contacts / sp,
        "contacts_per_s": contacts / sp * steps / wall,
        "iters_per_step": iters / sp,
        # the resolver's latency measures: one wavefront per world,
        # worlds_per_cu of them resident per CU
        "resolver": {
            "iterations_per_world_per_step": iters / sp / sc.n_worlds,
            "worlds_per_cu": per_cu,
            "resident_worlds": resident,
            "waves_per_simd": resident / (4.0 * n_cu),
            "ns_per_iteration_aggregate": resolve_ns / (iters / sp) if iters else None,
            "ns_per_iteration_per_world": resolve_ns * resident / (iters / sp) if iters else None,
        },
        "critical_path_last_step": crit,
        "resolver_split": split,
        "k1_ms_per_launch": k1_ms,
        "k1_bodies_per_launch": k1_bodies / K1_LAUNCHES,
        "residual_penetration": resid,
        "_final_pos": final_pos,
    }


def run_surrogate(World, sc, ck, model, steps, warmup, tol, polish, zero=False):
    """Surrogate resolver (FP32 state): timed steps from the window's start,
    then the same steps again with HIP events around its phases."""
    params, norm = model
    w = World(sc, precision="f32")
    w.sur_set_norm(norm)
    w.sur_set_params(params)
    _start(w, sc, ck)
    split = _tune(w, sc, "f32", surrogate=True, tol=tol, polish=polish) if not zero else None
    _start(w, sc, ck)
    if warmup > 0:
        w.step_sur(sc.dt, warmup, sc.iter_mult, tol, polish, zero)
        _start(w, sc, ck)
    t0 = time.perf_counter()
    fb = w.step_sur(sc.dt, steps, sc.iter_mult, tol, polish, zero)
    w.synchronize()
    wall = time.perf_counter() - t0
    out = {"value": sc.n_bodies * steps / wall, "ms_per_step": 1000.0 * wall / steps,
           "fallback_rate": float(fb.sum()) / (steps * sc.n_worlds)}
    if zero:
        w.close()
        return out
    # read-back of the state across the host boundary, after the timed steps
    t1 = time.perf_counter()
    for k in ("pos", "quat", "vel", "rot"):
        w.get(k)
    download_s = time.perf_counter() - t1
    t1 = time.perf_counter()
    final_pos = w.get("pos")
    w.get("quat")
    pos_quat_s = time.perf_counter() - t1
    resid = residual_penetration(w)
    # scene upload from host buffers (with the derived data)
    t1 = time.perf_counter()
    w.load(sc)
    w.synchronize()
    upload_s = time.perf_counter() - t1
    _start(w, sc, ck)
    w.set_profiling(True)
    w.step_sur(sc.dt, steps, sc.iter_mult, tol, polish, zero)
    w.set_profiling(False)
    ms, cnt = w.sur_phase_times()
    crit = critical_path(w)
    w.close()
    n_res, contacts, fallbacks, vel_iters, pos_fallbacks, pos_iters = (int(x) for x in cnt)
    names = ("prepare", "features", "network", "per_body_sums", "position_phase", "velocity_phase")
    out.update({
        "transfer": {"wall_s": wall, "upload_s": upload_s, "download_state_s": download_s,
                     "download_pos_quat_s": pos_quat_s},
        "resolver_ms_per_step": {k: round(ms[i] / max(n_res, 1), 5) for i, k in enumerate(names)},
        "contacts_per_step": contacts / max(n_res, 1),
        "velocity_iterations_per_step": vel_iters / max(n_res, 1),
        "position_iterations_per_step": pos_iters / max(n_res, 1),
        "position_fallback_rate": pos_fallbacks / max(n_res * sc.n_worlds, 1),
        "network_ms_total": ms[2],
        "contacts_total": contacts,
        "residual_penetration": resid,
        "critical_path_last_step": crit,
        "resolver_split": split,
        "_final_pos": final_pos,
    })
    return out


def critical_path(w):
    """The last step's resolver launches, from the per-world workgroup
    timestamps: each launch's span (first start to last end) and its
    longest world (duration, contacts, iterations), in microseconds."""
    t = w.resolver_world_times()
    counts = w.contact_counts()
    iters = w.resolver_iterations()
    out = {}
    for name, k in (("position_launch", 0), ("velocity_launch", 2)):
        st, en = t[:, k], t[:, k + 1]
        ok = st >= 0
        if not ok.any():
            continue
        dur = np.where(ok, en - st, -1.0)
        wmax = int(np.argmax(dur))
        out[name] = {"span_us": float(en[ok].max() - st[ok].min()), "longest_world_us": float(dur[wmax]),
                     "longest_world": wmax, "its_contacts": int(counts[wmax]),
                     "its_iterations": [int(x) for x in iters[wmax]],
                     "median_world_us": float(np.median(dur[ok]))}
    return out


def deviation(a, b):
    d = np.linalg.norm(a - b, axis=1)
    return {"mean_m": float(d.mean()), "p99_m": float(np.percentile(d, 99)), "max_m": float(d.max())}


def probe_k1(World, sc, launches=20):
    """K1 launches and a calibration copy for hardware-counter runs (rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE): the copy moves a known number of bytes,
    which scales the counters of K1."""
    import torch

    for prec in ("f32", "f64"):
        with World(sc, precision=prec) as w:
            for _ in range(launches):
                w.integrate(sc.dt)
            w.synchronize()
        print(f"probe: {launches} x k_integrate<{prec}> on {sc.n_bodies} bodies", flush=True)
    n = 64 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(launches):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"probe: {launches} copies of {4 * n} bytes (read {4 * n}, write {4 * n})", flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--window-start", type=int, default=None,
                    help="first step of the timed window (default: 0 for 600 or more steps, else 300)")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--model", default=None, help="surrogate weights (default: the packaged model)")
    ap.add_argument("--exact-only", action="store_true", help="time the exact resolver only")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle wall time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--resolver-split", default="auto",
                    help="the resolver's split between its kernels: 'auto' (the autotuner, per leg, from the "
                         "window's start state), 'default' (the library's), or 'min,max' (the max largest worlds "
                         "of at least min contacts take the multi-wave kernel; 0,0: none)")
    ap.add_argument("--probe-k1", action="store_true",
                    help="for hardware-counter runs: 20 launches of K1 (FP32 and FP64) on the workload, then 20 "
                         "copies of a 256 MiB float32 buffer as the calibration kernel; prints no JSON")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: worlds are independent, so a multi-GPU run would shard "
              "them with no exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cubez_amd import World, scenes, surrogate

    global RESOLVER_SPLIT
    RESOLVER_SPLIT = args.resolver_split
    sc = scenes.batched(args.seed, args.worlds)
    if args.probe_k1:
        probe_k1(World, sc)
        return
    nb = sc.n_bodies
    model_path = args.model or surrogate.DEFAULT_MODEL
    use_sur = not args.exact_only
    if use_sur and not os.path.exists(model_path):
        print(f"no surrogate model at {model_path} (python -m cubez_amd.surrogate train); "
              "use --exact-only to time the exact resolver alone", file=sys.stderr)
        sys.exit(1)
    tol, polish = surrogate.DEFAULT_TOL, surrogate.DEFAULT_POLISH
    start = args.window_start if args.window_start is not None else (0 if args.steps >= 600 else DEFAULT_WINDOW_START)

    _log(f"bench: settling to step {start}")
    ck, settle_s = settle(World, sc, start)
    # the resolver split is tuned on the settled regime whatever the window
    # (from step 0 the worlds are in free fall with few contacts)
    global TUNE_CK
    TUNE_CK = ck if start >= DEFAULT_WINDOW_START or RESOLVER_SPLIT != "auto" else \
        settle(World, sc, DEFAULT_WINDOW_START)[0]
    _log("bench: exact FP32")
    exact32 = run_path(World, sc, ck, "f32", args.steps, args.warmup)
    _log("bench: exact FP64")
    exact64 = run_path(World, sc, ck, "f64", args.steps, args.warmup)
    sur = zero = None
    if use_sur:
        model = surrogate.load(model_path)
        _log("bench: surrogate")
        sur = run_surrogate(World, sc, ck, model, args.steps, args.warmup, tol, polish)
        _log("bench: zero prediction")
        zero = run_surrogate(World, sc, ck, model, args.steps, args.warmup, tol, polish, zero=True)

    rs = 4
    window = (f"steps {start}..{start + args.steps - 1}" + (" (the whole workload)" if start == 0 and args.steps >= 600
                                                               else f" (a window of the settled regime: the state "
                                                                    f"of the exact FP64 run at step {start}, loaded "
                                                                    f"into every leg)"))
    workload = (f"batched worlds: {args.worlds} worlds, plane + n~U{{8..256}} cubes each ({nb} bodies), {window} at "
                f"dt=1/60, all-pairs narrowphase, maxIterations=4*C, FP32 state with FP64 SAT and contact basis, ")
    ptol = surrogate.DEFAULT_POS_TOL
    workload += (f"resolver: BF16 MLP surrogate for both phases, each polished by the exact worst-first loop "
                 f"(positions to positionEpsilon+{ptol} m, velocities to velocityEpsilon+{tol} m/s) within "
                 f"{polish}*C iterations, else the world's phase is re-solved by the plain method"
                 if use_sur else "resolver: exact")
    main_path = sur if use_sur else exact32
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
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload, "worlds": args.worlds, "bodies": nb, "seed": args.seed,
            "window_start": start, "window_steps": args.steps,
            "resolver": "surrogate+fallback" if use_sur else "exact", "tolerance_m_per_s": tol,
            "polish_per_contact": polish, "position_tolerance_m": surrogate.DEFAULT_POS_TOL,
            "network": "40-256-256-256-13 SiLU, BF16 MFMA, FP32 accumulate",
        },
        "settle_s": settle_s,
    }
    kern32 = exact32["kernels"]
    # per-phase kernel times of the surrogate step (ms/step): integrate and
    # narrowphase are the FP32 exact path's kernels on the same window
    phases = {"integrate": kern32["integrate"], "narrowphase": kern32["np_count"] + kern32["np_scans"] + kern32["np_emit"]}
    if use_sur:
        out["speedup_vs_exact_f64"] = sur["value"] / exact64["value"]
        out["speedup_vs_exact_f32"] = sur["value"] / exact32["value"]
        out["fallback_rate"] = sur["fallback_rate"]
        out["velocity_iterations_per_step"] = {"surrogate": sur["velocity_iterations_per_step"]}
        out["position_iterations_per_step"] = {"surrogate": sur["position_iterations_per_step"]}
        out["position_fallback_rate"] = sur["position_fallback_rate"]
        out["critical_path_last_step"] = {"surrogate": sur["critical_path_last_step"],
                                          "exact_f64": exact64["critical_path_last_step"],
                                          "exact_f32": exact32["critical_path_last_step"]}
        out["resolver_split"] = {"surrogate": sur["resolver_split"], "exact_f64": exact64["resolver_split"],
                                 "exact_f32": exact32["resolver_split"],
                                 "what": "(min_contacts, max_worlds): the max_worlds largest worlds of at least "
                                         "min_contacts contacts resolve on the 4-wavefront kernel (0,0: none); "
                                         "picked per leg by the autotuner from the window's start state"}
        out["contacts_per_step"] = sur["contacts_per_step"]
        out["contacts_per_s"] = sur["contacts_per_step"] * sur["value"] / nb
        phases.update(sur["resolver_ms_per_step"])
        out["zero_prediction"] = {k: zero[k] for k in ("value", "ms_per_step", "fallback_rate")}
        tr = sur["transfer"]
        out["pcie_inclusive"] = {
            "job_value": nb * args.steps / (tr["upload_s"] + tr["wall_s"] + tr["download_state_s"]),
            "job": "scene upload from host buffers + the timed steps + read-back of pos/quat/vel/rot",
            "per_step_readback_value": nb / (tr["wall_s"] / args.steps + tr["download_pos_quat_s"]),
            "per_step_readback": "pos and quat read back to the host after every step",
            "unit": "body-steps/s", "upload_ms": 1000 * tr["upload_s"],
            "download_state_ms": 1000 * tr["download_state_s"], "download_pos_quat_ms": 1000 * tr["download_pos_quat_s"],
        }
        out["error"] = {
            "bar": f"per contact, at most positionEpsilon + {ptol * 1000:.1f} mm of penetration after the "
                   f"position phase and velocityEpsilon + {tol} m/s of residual closing speed (= "
                   f"{tol / 60.0 * 1000:.1f} mm of extra penetration per 1/60 s step) after the velocity phase; "
                   "a world above either, or whose position moves separate a contact by more than twice its "
                   "deepest penetration, is re-solved by that plain phase",
            "residual_penetration": {"surrogate": sur["residual_penetration"],
                                     "exact_f32": exact32["residual_penetration"],
                                     "exact_f64": exact64["residual_penetration"]},
            "surrogate_minus_exact_f64_mean_penetration_m": sur["residual_penetration"].get("mean_m", 0.0)
                                                           - exact64["residual_penetration"].get("mean_m", 0.0),
            "position_deviation_after_window": {
                "surrogate_vs_exact_f32": deviation(sur["_final_pos"], exact32["_final_pos"]),
                "surrogate_vs_exact_f64": deviation(sur["_final_pos"], exact64["_final_pos"]),
                "exact_f32_vs_exact_f64": deviation(exact32["_final_pos"], exact64["_final_pos"]),
            },
        }
    else:
        out["contacts_per_step"] = exact32["contacts_per_step"]
        out["contacts_per_s"] = exact32["contacts_per_s"]
    out["kernels_ms_per_step"] = {k: round(v, 5) for k, v in phases.items()}
    for k, p in (("exact_f64", exact64), ("exact_f32", exact32)):
        out[k] = {x: p[x] for x in ("value", "ms_per_step", "kernels", "iters_per_step", "contacts_per_step",
                                     "resolver")}

    # K1 roofline (FP32, the kernel the surrogate path launches): algorithmic
    # bytes per launch / average launch time of K1_LAUNCHES back-to-back
    # launches on the window's start state, timed with HIP events on the
    # library's stream
    aw = exact32["k1_bodies_per_launch"]
    k1_bytes = aw * (K1_READ_REALS * rs + 1 + K1_WRITE_REALS * rs + 1) + (nb - aw) * (1 + 24 * rs)
    k1_gbs = k1_bytes / (exact32["k1_ms_per_launch"] / 1000.0) / 1e9
    ratio = traffic_ratio("f32")
    out["k1"] = {
        "kernel": "k_integrate<float> (Integrate + CalculateDerivedData)", "bound": "hbm", "achieved": k1_gbs,
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": k1_gbs / HBM_PEAK_GBS,
        "traffic": ratio * k1_bytes if ratio else None, "bytes_per_launch": k1_bytes,
        "ms_per_launch": exact32["k1_ms_per_launch"], "awake_bodies": aw,
    }
    if use_sur:
        flop = 2 * sum(i * o for i, o in surrogate.LAYERS)  # 288,768 per contact
        tfl = flop * sur["contacts_total"] / (sur["network_ms_total"] / 1000.0) / 1e12 if sur["network_ms_total"] else 0
        out["mlp"] = {"kernel": "k_mlp_forward (fused 4-layer forward)", "bound": "mfma", "flop_per_contact": flop,
                      "ms_per_step": phases["network"], "achieved": tfl, "peak": BF16_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": tfl / BF16_PEAK_TFLOPS}
    # the dominant kernel of the headline step and its roofline
    if use_sur:
        dom = max(("position_phase", "velocity_phase", "network", "narrowphase", "integrate", "features"),
                  key=lambda k: phases[k])
    else:
        dom = "resolve"
    C = main_path["contacts_per_step"]
    res_bytes = C * (RES_CONTACT_READ * rs + 8 + RES_CONTACT_WRITE * rs) + nb * (RES_BODY_READ * rs + 1 +
                                                                                  RES_BODY_WRITE * rs + 1)
    if dom in ("position_phase", "velocity_phase", "resolve"):
        ms = phases[dom] if use_sur else kern32["resolve"]
        ach = res_bytes / (ms / 1000.0) / 1e9
        roof = {"kernel": {"position_phase": "k_resolve_f32<polish> (adjustPositions from the network's moves, "
                                             "with the exact fallback)",
                           "velocity_phase": "k_resolve_f32<polish> (adjustVelocities from the network's changes, "
                                             "with the exact fallback)",
                           "resolve": "k_resolve_f32<plain>"}[dom],
                "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                "traffic": None, "bytes_per_launch": res_bytes, "ms_per_launch": ms,
                "note": "serial worst-first iterations per world: latency-bound, see exact_*.resolver"}
    elif dom == "network":
        roof = dict(out["mlp"])
        roof["traffic"] = None
    elif dom == "narrowphase":
        roof = {"kernel": "narrowphase (k_np_cull + k_np_eval + scans + k_np_emit_hits, FP64 SAT)",
                "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "ms_per_launch": phases[dom]}
    else:
        roof = dict(out["k1"])
    roof["phase"] = dom
    out["roofline"] = roof
    if not args.no_cpu:
        _log("bench: CPU baseline")
        out["cpu_baseline"], counters, sample_bs = cpu_baseline(sc, start + args.steps, args.cpu_budget, start)
        if counters:
            k64 = exact64["kernels"]
            out["k2"] = narrowphase_rate(counters, sample_bs, nb, k64["np_count"] + k64["np_scans"] + k64["np_emit"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
