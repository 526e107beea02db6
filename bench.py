#!/usr/bin/env python3
"""Benchmark of the cubez simulation step on one MI355X.

Workload: 4096 independent worlds, each a ground plane and n ~ U{8..256}
cubes (the batched-worlds configuration), advanced with the reference's
step -- Integrate + CalculateDerivedData, the all-pairs narrowphase,
ResolveContacts(4 x contacts) -- for K steps at dt = 1/60 from the same
initial state.  A step is one pass of that hot path over all worlds.  The
metric is body-steps/s (bodies x steps / seconds) with the state resident in
HBM, for the learned-surrogate resolver with exact fallback (FP32 state,
tolerance 0.06 m/s of residual closing speed = 1 mm of penetration per
step), next to the exact FP64 GPU resolver it is compared with.

Prints one JSON line.  Besides the metric it reports:
  exact_f64, exact_f32   the exact resolver (hipGraph-replayed step)
  zero_prediction        the surrogate's polish started from a zero
                         prediction: the same method without the network
  fallback_rate          worlds re-solved by the plain method / world-steps
  mlp                    the network's forward GEMMs: FLOP, time, MFMA share
  roofline               K1 (integrate + derived data, HBM-bound) from its
                         average launch time measured with HIP events
  cpu_baseline           the FP64 oracle (a C restatement of the Go
                         reference) on a bounded sample of the same worlds
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
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, FMA counted as two
SAT_AXIS_FLOP = 75
PLANE_CHECK_FLOP = 40
K1_LAUNCHES = 50

# Algorithmic bytes of K1 per body-step: an integrated body reads 29 reals
# (pos 3, quat 4, vel 3, rot 3, acc 3, inv_mass 1, iit_body 9, motion 1, the
# two Pow factors) + its flag byte and writes 50 reals (pos, quat, vel, rot,
# last_acc, transform 12, iit_world 9, collider transform 12, motion) + the
# flag byte; a sleeping body reads its flag byte and transform (12) and
# writes the collider transform (12).  Reals are 8 bytes (FP64) or 4 (FP32).


_SCENE = None
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_integrate_hbm_traffic.md")


def traffic_ratio(precision):
    """HBM traffic / algorithmic bytes of K1, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE measurement (copy-calibrated), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith(f"traffic_over_algorithmic_{precision}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def _cpu_world(args):
    """Oracle run of one world (worker process; the scene is inherited)."""
    from oracle.cbz_oracle import OracleWorld

    w, steps = args
    sc = _SCENE
    o = OracleWorld(sc, w)
    o.calculate_derived()
    t0 = time.perf_counter()
    o.step(sc.dt, steps, sc.iter_mult)
    return o.n * steps, time.perf_counter() - t0, o.counters()


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
    counters = {k: sum(r[2][k] for r in res) for k in res[0][2]} if res else {}
    return {
        "value": bs / wall,
        "unit": "body-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{len(worlds)} of {n_worlds} worlds (every {stride}th), all {steps} steps, "
                  f"{bs} body-steps in {wall:.1f} s wall on {cores} processes; host: {_host_cpu()}",
    }, counters, bs


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


def narrowphase_rate(counters, sample_body_steps, n_bodies, kernels):
    """K2's FP64 VALU rate: the reference's narrowphase work on the same
    steps, from the oracle's op counters on the CPU sample (scaled to the
    whole scene by body-steps), over the FP64 narrowphase kernels' time per
    step.  Work model (per unit, the reference's arithmetic in
    colliders.go:445-477 tryAxis / penetrationOnAxis, :762-770
    transformToAxis, :308-367 the half-space check): 75 flops per SAT axis
    tested (normalise, two projections, the distance), 40 per body-plane check;
    contact generation is not counted, so the rate is a lower bound."""
    scale = n_bodies / max(sample_body_steps, 1)
    per_step = {k: counters[k] * scale for k in ("pairs", "plane_checks", "sat_axes", "sat_skipped", "contacts")}
    flop = SAT_AXIS_FLOP * per_step["sat_axes"] + PLANE_CHECK_FLOP * per_step["plane_checks"]
    ms = kernels["np_count"] + kernels["np_scans"] + kernels["np_emit"]
    tf = flop / (ms / 1000.0) / 1e12 if ms > 0 else 0.0
    return {
        "kernels": "k_np_cull + k_np_eval + scans + k_np_emit_hits (FP64 state, same 600 steps)",
        "bound": "fp64-valu", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": tf / FP64_PEAK_TFLOPS, "frac_of_unfused_peak": tf / (FP64_PEAK_TFLOPS / 2),
        "flop_per_step": flop, "ms_per_step": ms, "reference_work_per_step": per_step,
        "flop_model": f"{SAT_AXIS_FLOP} per SAT axis tested + {PLANE_CHECK_FLOP} per body-plane check "
                      "(contact generation not counted: a lower bound)",
    }


def run_path(World, sc, precision, steps, warmup):
    """Exact resolver: warm up, then time `steps` steps from the initial
    state (graph replay), then the same steps again with HIP events between
    kernels."""
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
    # K1 alone: back-to-back launches on the initial state between two events
    w.load(sc)
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
        "integrated": integrated,
        "contacts_per_step": contacts / sp,
        "contacts_per_s": contacts / (ms[5] / 1000.0) if ms[5] > 0 else None,
        "iters_per_step": iters / sp,
        # the resolver's latency measures (its phase has no roofline): one
        # wavefront per world, worlds_per_cu of them resident per CU
        "resolver": {
            "iterations_per_world_per_step": iters / sp / sc.n_worlds,
            "worlds_per_cu": per_cu,
            "resident_worlds": resident,
            "waves_per_simd": resident / (4.0 * n_cu),
            "ns_per_iteration_aggregate": resolve_ns / (iters / sp) if iters else None,
            "ns_per_iteration_per_world": resolve_ns * resident / (iters / sp) if iters else None,
        },
        "k1_ms_per_launch": k1_ms,
        "k1_bodies_per_launch": k1_bodies / K1_LAUNCHES,
    }


def run_surrogate(World, sc, model, steps, warmup, tol, polish, zero=False):
    """Surrogate resolver (FP32 state): timed steps from the initial state,
    then the same steps again with HIP events around its phases."""
    params, norm = model
    w = World(sc, precision="f32")
    w.sur_set_norm(norm)
    w.sur_set_params(params)
    if warmup > 0:
        w.step_sur(sc.dt, warmup, sc.iter_mult, tol, polish, zero)
    w.load(sc)
    w.synchronize()
    t0 = time.perf_counter()
    fb = w.step_sur(sc.dt, steps, sc.iter_mult, tol, polish, zero)
    w.synchronize()
    wall = time.perf_counter() - t0
    # host-buffer transfers across the boundary: the scene upload (with the
    # derived data), then, after the profiled steps, the state read-back
    t1 = time.perf_counter()
    w.load(sc)
    w.synchronize()
    upload_s = time.perf_counter() - t1
    w.set_profiling(True)
    w.step_sur(sc.dt, steps, sc.iter_mult, tol, polish, zero)
    w.set_profiling(False)
    ms, cnt = w.sur_phase_times()
    t1 = time.perf_counter()
    for k in ("pos", "quat", "vel", "rot"):
        w.get(k)
    download_s = time.perf_counter() - t1
    t1 = time.perf_counter()
    for k in ("pos", "quat"):
        w.get(k)
    pos_quat_s = time.perf_counter() - t1
    w.close()
    n_res, contacts, fallbacks, vel_iters = (int(x) for x in cnt)
    names = ("position_phase", "features", "network", "per_body_sums", "velocity_phase")
    return {
        "value": sc.n_bodies * steps / wall,
        "ms_per_step": 1000.0 * wall / steps,
        "fallback_rate": float(fb.sum()) / (steps * sc.n_worlds),
        "transfer": {"wall_s": wall, "upload_s": upload_s, "download_state_s": download_s,
                     "download_pos_quat_s": pos_quat_s},
        "resolver_ms_per_step": {k: round(ms[i] / max(n_res, 1), 5) for i, k in enumerate(names)},
        "contacts_per_step": contacts / max(n_res, 1),
        "velocity_iterations_per_step": vel_iters / max(n_res, 1),
        "network_ms_total": ms[2],
        "contacts_total": contacts,
    }


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
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--model", default=None, help="surrogate weights (default: the packaged model)")
    ap.add_argument("--exact-only", action="store_true", help="time the exact resolver only")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle wall time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-k1", action="store_true",
                    help="for hardware-counter runs: 20 launches of K1 (FP32 and FP64) on the workload, then 20 "
                         "copies of a 256 MiB float32 buffer as the calibration kernel; prints no JSON")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: worlds are independent, so a multi-GPU run would shard "
              "them with no exchange; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from cubez_amd import World, scenes, surrogate

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

    exact32 = run_path(World, sc, "f32", args.steps, args.warmup)
    exact64 = run_path(World, sc, "f64", args.steps, args.warmup)
    sur = zero = None
    if use_sur:
        model = surrogate.load(model_path)
        sur = run_surrogate(World, sc, model, args.steps, args.warmup, tol, polish)
        zero = run_surrogate(World, sc, model, args.steps, args.warmup, tol, polish, zero=True)

    kern = exact32["kernels"]
    rs = 4
    # K1 roofline (FP32, the kernel the surrogate path launches): algorithmic
    # bytes per launch / average launch time of K1_LAUNCHES back-to-back
    # launches on the workload's initial state, timed with HIP events on the
    # library's stream
    k1_awake = 29 * rs + 1 + 50 * rs + 1
    k1_sleep = 1 + 12 * rs + 12 * rs
    aw = exact32["k1_bodies_per_launch"]
    k1_bytes = aw * k1_awake + (nb - aw) * k1_sleep
    achieved = k1_bytes / (exact32["k1_ms_per_launch"] / 1000.0) / 1e9
    main_path = sur if use_sur else exact32
    workload = (f"batched worlds: {args.worlds} worlds, plane + n~U{{8..256}} cubes each ({nb} bodies), "
                f"first {args.steps} steps at dt=1/60 from the seeded initial state, all-pairs narrowphase, "
                f"maxIterations=4*C, FP32 state with FP64 SAT and contact basis, ")
    workload += (f"resolver: exact position phase + BF16 MLP surrogate for the velocity phase, polished by "
                 f"the exact worst-first loop to velocityEpsilon+{tol} m/s within {polish}*C iterations, "
                 f"else the world is re-solved by the plain method" if use_sur else "resolver: exact")
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
            "resolver": "surrogate+fallback" if use_sur else "exact", "tolerance_m_per_s": tol,
            "polish_per_contact": polish, "network": "40-256-256-256-12 SiLU, BF16 MFMA, FP32 accumulate",
        },
    }
    if use_sur:
        out["speedup_vs_exact_f64"] = sur["value"] / exact64["value"]
        out["speedup_vs_exact_f32"] = sur["value"] / exact32["value"]
        out["fallback_rate"] = sur["fallback_rate"]
        out["velocity_iterations_per_step"] = {"surrogate": sur["velocity_iterations_per_step"],
                                               "zero_prediction": zero["velocity_iterations_per_step"]}
        out["surrogate_resolver_ms_per_step"] = sur["resolver_ms_per_step"]
        out["contacts_per_step"] = sur["contacts_per_step"]
        out["contacts_per_s"] = sur["contacts_per_step"] * sur["value"] / nb
        flop = 2 * sum(i * o for i, o in surrogate.LAYERS)  # 288,768 per contact
        tfl = flop * sur["contacts_total"] / (sur["network_ms_total"] / 1000.0) / 1e12 if sur["network_ms_total"] else 0
        out["mlp"] = {"flop_per_contact": flop, "ms_per_step": sur["resolver_ms_per_step"]["network"],
                      "achieved_tflops": tfl, "peak_tflops": 2500.0, "mfma_frac": tfl / 2500.0,
                      "bound": "mfma", "kernel": "k_mlp_forward (fused 4-layer forward)"}
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
    else:
        out["contacts_per_step"] = exact32["contacts_per_step"]
        out["contacts_per_s"] = exact32["contacts_per_s"]
    out["exact_f64"] = {k: exact64[k] for k in ("value", "ms_per_step", "kernels", "iters_per_step", "resolver")}
    out["exact_f32"] = {k: exact32[k] for k in ("value", "ms_per_step", "kernels", "iters_per_step", "resolver")}
    out["kernels_ms_per_step"] = kern
    out["dominant_kernel"] = max(("integrate", "np_count", "np_scans", "np_emit", "resolve"), key=lambda k: kern[k])
    ratio = traffic_ratio("f32")
    out["roofline"] = {
        "kernel": "k_integrate (K1: Integrate + CalculateDerivedData)",
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": ratio * k1_bytes if ratio else None,
        "bytes_per_launch": k1_bytes, "ms_per_launch": exact32["k1_ms_per_launch"],
        "measured_on": f"{K1_LAUNCHES} back-to-back launches on the initial state (every body awake)",
        "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE of k_integrate<float>, copy-calibrated, "
                           f"traffic/algorithmic = {ratio} ({os.path.relpath(TRAFFIC_PROFILE, ROOT)})")
        if ratio else None,
    }
    if not args.no_cpu:
        out["cpu_baseline"], counters, sample_bs = cpu_baseline(sc, args.steps, args.cpu_budget)
        if counters:
            out["narrowphase"] = narrowphase_rate(counters, sample_bs, nb, exact64["kernels"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
