"""Host-side pieces of bench.py that need no GPU: the CPU-baseline leg (the
oracle on a sample of worlds) and the narrowphase rate built from the
oracle's op counters."""
import bench
from cubez_amd import scenes


def test_cpu_baseline_sample_and_counters():
    sc = scenes.batched(0, 16)
    base, counters, body_steps = bench.cpu_baseline(sc, 5, 0.5)
    assert set(base) == {"value", "value_1core", "unit", "cores", "kind", "sample"}
    assert base["kind"] == "port" and base["unit"] == "body-steps/s" and base["value"] > 0
    assert base["value_1core"] > 0
    # the sample covers whole worlds for every step
    assert body_steps % 5 == 0 and body_steps > 0
    # every body is integrated once per step on the first steps (all awake)
    assert counters["integrated"] == body_steps
    assert counters["pairs"] > 0 and counters["plane_checks"] == body_steps
    # every cube-cube pair tests at least its first (never degenerate) axis
    assert counters["sat_axes"] >= counters["pairs"]


def test_cpu_baseline_counts_the_window_only():
    """With a window starting at step 3 of 5, the work counters and the
    body-steps they cover are those of steps 3 and 4 alone."""
    sc = scenes.batched(0, 16)
    _, whole, bs_whole = bench.cpu_baseline(sc, 5, 0.5)
    _, first, bs_first = bench.cpu_baseline(sc, 3, 0.5)
    _, window, bs_window = bench.cpu_baseline(sc, 5, 0.5, window_start=3)
    assert bs_window * 5 == bs_whole * 2 and bs_first * 5 == bs_whole * 3
    for k in ("pairs", "plane_checks", "sat_axes", "contacts", "integrated"):
        assert window[k] == whole[k] - first[k], k


def test_narrowphase_rate_arithmetic():
    counters = {"pairs": 1000, "plane_checks": 100, "sat_axes": 3000, "sat_skipped": 0, "contacts": 50}
    r = bench.narrowphase_rate(counters, 200, 100, 1.0)  # the sample is two steps of the scene's work
    assert r["reference_work_per_step"]["sat_axes"] == 1500
    flop = bench.SAT_AXIS_FLOP * 1500 + bench.PLANE_CHECK_FLOP * 50
    assert r["flop_per_step"] == flop
    assert r["ms_per_step"] == 1.0
    assert abs(r["achieved"] - flop / 1e-3 / 1e12) < 1e-15
    assert abs(r["frac"] - r["achieved"] / 78.6) < 1e-15
    assert abs(r["frac_of_unfused_peak"] - 2 * r["frac"]) < 1e-15
