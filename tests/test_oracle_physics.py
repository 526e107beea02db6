"""Pins of the oracle's dynamics, narrowphase and resolver by closed forms,
invariants and brute force (the reference holds no tests for these).

Each check is chosen so that a plausible transcription mistake fails it:
damping applied before the acceleration, a body-frame instead of a
world-frame spin, a wrong SAT axis or sign, a non-conservative impulse."""
import math

import numpy as np
import pytest

from cubez_amd import scenes
from cubez_amd.scenes import Scene, _new_bodies, block_inertia, go_invert3, inertia_coeffs
from oracle.cbz_oracle import OracleWorld

DT = 1.0 / 60.0


def make_scene(specs, planes=((0.0, 1.0, 0.0, 0.0),), gravity=-9.8):
    """specs: list of dicts with shape ('cube'|'sphere'), half or r, mass, pos,
    and optional quat, vel, rot."""
    n = len(specs)
    f = _new_bodies(n)
    for i, s in enumerate(specs):
        cube = s.get("shape", "cube") == "cube"
        f["shape"][i] = 0 if cube else 1
        h = np.array(s["half"] if cube else [s["r"], 0, 0], dtype=float)
        f["half"][i] = h
        f["pos"][i] = s["pos"]
        f["quat"][i] = s.get("quat", (1, 0, 0, 0))
        f["vel"][i] = s.get("vel", (0, 0, 0))
        f["rot"][i] = s.get("rot", (0, 0, 0))
        f["acc"][i] = (0, gravity, 0)
        m = s.get("mass", 1.0)
        f["inv_mass"][i] = 1.0 / m if m else 0.0
        if cube:
            I = block_inertia(h[None, :], np.array([m]))
        else:
            c = 0.4 * m * h[0] * h[0]
            I = inertia_coeffs(np.array([c]), np.array([c]), np.array([c]))
        f["iit_body"][i] = go_invert3(I)[0]
    p = np.array(planes, dtype=float).reshape(-1, 4)
    return Scene(np.array([0, n], np.int32), np.array([0, len(p)], np.int32), p[:, :3].copy(), p[:, 3].copy(), f)


def world(specs, **kw):
    o = OracleWorld(make_scene(specs, **kw))
    o.calculate_derived()
    return o


# --- L1 dynamics -----------------------------------------------------------

def test_free_fall_closed_form():
    """Integrate (rigidbody.go:213-259): v_k = (v_{k-1} + g dt) L with
    L = Pow(0.95, dt), x_k = x_{k-1} + v_k dt, so v_k = g dt L (1-L^k)/(1-L)."""
    o = world([{"half": (1, 1, 1), "mass": 2.0, "pos": (0, 1000, 0)}], planes=())
    L = math.pow(0.95, DT)
    g = -9.8
    y = 1000.0
    for k in range(1, 121):
        o.integrate(DT)
        v = g * DT * L * (1 - L ** k) / (1 - L)
        y += v * DT
        assert math.isclose(o.get("vel")[0, 1], v, rel_tol=1e-12)
        assert math.isclose(o.get("pos")[0, 1], y, rel_tol=1e-12)
    assert o.get("vel")[0, 0] == 0.0 and o.get("vel")[0, 2] == 0.0


def test_spin_is_world_frame_and_decays():
    """AddScaledVector multiplies (0, w dt) on the left: the increment of a
    rotated body turns about the world axis of w; |w| decays as L^k."""
    s2 = math.sqrt(0.5)
    q0 = (s2, s2, 0.0, 0.0)  # 90 degrees about x
    w0 = np.array([0.0, 0.0, 2.0])
    o = world([{"half": (1, 1, 1), "mass": 1.0, "pos": (0, 100, 0), "quat": q0, "rot": tuple(w0)}],
              planes=(), gravity=0.0)
    R0 = o.get("transform")[0, :9].reshape(3, 3, order="F")
    o.integrate(DT)
    R1 = o.get("transform")[0, :9].reshape(3, 3, order="F")
    D = R1 @ R0.T
    axis = np.array([D[2, 1] - D[1, 2], D[0, 2] - D[2, 0], D[1, 0] - D[0, 1]])
    axis /= np.linalg.norm(axis)
    assert np.allclose(axis, [0, 0, 1], atol=1e-6)
    L = math.pow(0.95, DT)  # AngularDamping defaults to the linear 0.95 (rigidbody.go:108)
    for k in range(2, 50):
        o.integrate(DT)
        assert np.allclose(o.get("rot")[0], w0 * L ** k, rtol=1e-13)
    R = o.get("transform")[0, :9].reshape(3, 3, order="F")
    assert np.allclose(R.T @ R, np.eye(3), atol=1e-12)
    assert math.isclose(np.linalg.det(R), 1.0, rel_tol=1e-12)


def test_world_inverse_inertia():
    rng = np.random.default_rng(0)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    o = world([{"half": (0.5, 1.0, 2.0), "mass": 3.0, "pos": (0, 10, 0), "quat": tuple(q)}], planes=())
    R = o.get("transform")[0, :9].reshape(3, 3, order="F")
    Ib = np.diag([0.3 * 3 * (1 + 4), 0.3 * 3 * (0.25 + 4), 0.3 * 3 * (0.25 + 1)])
    ref = R @ np.linalg.inv(Ib) @ R.T
    assert np.allclose(o.get("iit_world")[0].reshape(3, 3, order="F"), ref, rtol=1e-12, atol=1e-15)


def test_sleep_after_motion_decays():
    """motion = b motion + (1-b) |v|^2 with b = Pow(0.5, dt), from 0.6 at rest:
    0.6 * 0.5^(k/60) < 0.3 first holds at k = 61 (60 is the boundary)."""
    o = world([{"half": (1, 1, 1), "mass": 1.0, "pos": (0, 100, 0)}], planes=(), gravity=0.0)
    asleep_at = None
    for k in range(1, 80):
        o.integrate(DT)
        if not (o.get("flags")[0] & 1):
            asleep_at = k
            break
    assert asleep_at in (60, 61)
    assert np.all(o.get("vel")[0] == 0) and np.all(o.get("rot")[0] == 0)


# --- L2 narrowphase ----------------------------------------------------------

def test_flat_cube_gives_four_contacts_edge_gives_two():
    o = world([{"half": (1, 1, 1), "mass": 1, "pos": (0, 1 - 0.001, 0)}])
    c = o.narrowphase()
    assert len(c) == 4
    assert np.allclose(c["pen"], 0.001, rtol=1e-9)
    assert np.all(c["normal"] == [0, 1, 0]) and np.all(c["body1"] == -1)
    s = math.sin(math.pi / 8)
    q = (math.cos(math.pi / 8), 0.0, 0.0, s)  # 45 degrees about z: resting on an edge
    o = world([{"half": (1, 1, 1), "mass": 1, "pos": (0, math.sqrt(2) - 0.001, 0), "quat": q}])
    c = o.narrowphase()
    assert len(c) == 2
    assert np.allclose(c["pen"], 0.001, atol=1e-9)


def test_sphere_plane_and_sphere_sphere_textbook():
    o = world([{"shape": "sphere", "r": 0.5, "mass": 1, "pos": (0, 0.4, 0)}])
    c = o.narrowphase()
    assert len(c) == 1 and math.isclose(c["pen"][0], 0.1, rel_tol=1e-12)
    o = world([{"shape": "sphere", "r": 1.0, "mass": 1, "pos": (0, 10, 0)},
               {"shape": "sphere", "r": 1.0, "mass": 1, "pos": (1.5, 10, 0)}])
    c = o.narrowphase()
    assert len(c) == 2  # ordered pairs: (0,1) and (1,0)
    assert np.allclose(c["pen"], 0.5)
    assert np.allclose(c["normal"][0], [-1, 0, 0]) and np.allclose(c["normal"][1], [1, 0, 0])
    assert (c["body0"][0], c["body1"][0]) == (0, 1)


def box_vertices(R, t, h):
    s = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], dtype=float)
    return (s * h) @ R.T + t


def sat_min(RA, tA, hA, RB, tB, hB):
    """Independent 15-axis separating-axis evaluation in numpy."""
    axes = [RA[:, i] for i in range(3)] + [RB[:, i] for i in range(3)]
    axes += [np.cross(RA[:, i], RB[:, k]) for i in range(3) for k in range(3)]
    d = tB - tA
    best, pens = None, []
    for a in axes:
        if a @ a < 1e-7:
            pens.append(None)
            continue
        a = a / np.linalg.norm(a)
        p = np.sum(hA * np.abs(RA.T @ a)) + np.sum(hB * np.abs(RB.T @ a)) - abs(d @ a)
        pens.append(p)
    valid = [p for p in pens if p is not None]
    return min(valid), pens


def boxes_overlap_lp(RA, tA, hA, RB, tB, hB, shrink):
    """Exact convex overlap test: is there a point inside both boxes shrunk by
    `shrink` (a linear feasibility problem)?"""
    from scipy.optimize import linprog
    Aub, bub = [], []
    for R, t, h in ((RA, tA, hA), (RB, tB, hB)):
        for i in range(3):
            Aub.append(R[:, i]); bub.append(R[:, i] @ t + h[i] - shrink)
            Aub.append(-R[:, i]); bub.append(-(R[:, i] @ t) + h[i] - shrink)
    r = linprog(np.zeros(3), A_ub=np.array(Aub), b_ub=np.array(bub), bounds=[(None, None)] * 3, method="highs")
    return r.status == 0


def test_cube_cube_sat_against_brute_force():
    rng = np.random.default_rng(7)
    checked = hits = 0
    for trial in range(150):
        specs = []
        for k in range(2):
            q = rng.standard_normal(4)
            q /= np.linalg.norm(q)
            specs.append({"half": tuple(rng.uniform(0.3, 1.5, 3)), "mass": 1.0,
                          "pos": tuple(rng.uniform(-1.2, 1.2, 3) + [0, 50, 0]), "quat": tuple(q)})
        o = world(specs, planes=())
        T = o.get("ctransform")
        R = [T[i, :9].reshape(3, 3, order="F") for i in range(2)]
        t = [T[i, 9:] for i in range(2)]
        h = [np.array(s["half"]) for s in specs]
        c = o.narrowphase()
        pen_min, _ = sat_min(R[0], t[0], h[0], R[1], t[1], h[1])
        inside = boxes_overlap_lp(R[0], t[0], h[0], R[1], t[1], h[1], 0.02)
        apart = not boxes_overlap_lp(R[0], t[0], h[0], R[1], t[1], h[1], -0.02)
        if inside:
            checked += 1
            hits += 1
            assert len(c) == 2, trial  # both ordered pairs report it
            assert np.allclose(c["pen"], pen_min, rtol=1e-9, atol=1e-12)
            assert np.allclose(np.linalg.norm(c["normal"], axis=1), 1.0)
        elif apart:
            checked += 1
            assert len(c) == 0, trial
    assert checked > 100 and hits > 20


def test_cube_sphere_penetration_closed_form():
    rng = np.random.default_rng(9)
    n_hit = 0
    for _ in range(100):
        q = rng.standard_normal(4)
        q /= np.linalg.norm(q)
        h = rng.uniform(0.5, 1.5, 3)
        r = rng.uniform(0.3, 1.0)
        cpos = rng.uniform(-1.8, 1.8, 3) + [0, 50, 0]
        o = world([{"half": tuple(h), "mass": 1, "pos": (0, 50, 0), "quat": tuple(q)},
                   {"shape": "sphere", "r": r, "mass": 1, "pos": tuple(cpos)}], planes=())
        T = o.get("ctransform")[0]
        R = T[:9].reshape(3, 3, order="F")
        local = R.T @ (cpos - T[9:])
        dist = math.sqrt(np.sum(np.maximum(np.abs(local) - h, 0) ** 2))
        c = o.narrowphase()
        if dist < r - 1e-6 and dist > 1e-6:
            n_hit += 1
            assert len(c) == 2  # (cube, sphere) and (sphere, cube) both put the cube first
            assert np.all(c["body0"] == 0) and np.all(c["body1"] == 1)
            assert np.allclose(c["pen"], r - dist, atol=1e-12)
        elif dist > r + 1e-6:
            assert len(c) == 0
    assert n_hit > 10


# --- L3 resolver -------------------------------------------------------------

def test_single_contact_position_then_velocity():
    """One sphere-plane contact: the position step removes the penetration
    (no angular part: r x n = 0), and the velocity step leaves the normal
    closing velocity at -restitution times the approach velocity."""
    o = world([{"shape": "sphere", "r": 0.5, "mass": 2.0, "pos": (0, 0.45, 0), "vel": (0.3, -2.0, 0.0)}])
    o.integrate(DT)
    c = o.narrowphase()
    assert len(c) == 1
    pen0 = c["pen"][0]
    y0 = o.get("pos")[0, 1]
    v0 = o.get("vel")[0].copy()
    out = o.resolve(c, DT, 1)
    assert math.isclose(o.get("pos")[0, 1], y0 + pen0, rel_tol=1e-12)
    assert abs(out["pen"][0]) < 1e-12
    v1 = o.get("vel")[0]
    la = -9.8 * DT  # lastFrameAcceleration * dt enters the desired velocity change
    want_vy = -0.1 * (v0[1] - la)
    assert math.isclose(v1[1], want_vy, rel_tol=1e-9)
    assert abs(v1[0]) < abs(v0[0])  # friction removes tangential velocity (sticking)


def test_position_tie_goes_to_the_lowest_index():
    """Two independent sphere-plane contacts with bit-equal penetration and
    one iteration: the reference's strict '>' from positionEpsilon resolves
    contact 0 (contact.go:238-248), and only its body moves."""
    o = world([{"shape": "sphere", "r": 0.5, "mass": 1.0, "pos": (-3, 0.45, 0)},
               {"shape": "sphere", "r": 0.5, "mass": 1.0, "pos": (3, 0.45, 0)}], gravity=0.0)
    c = o.narrowphase()
    assert len(c) == 2 and c["pen"][0] == c["pen"][1] and c["body0"][0] == 0
    y0 = o.get("pos")[:, 1].copy()
    o.resolve(c, DT, 1)
    y1 = o.get("pos")[:, 1]
    assert math.isclose(y1[0], y0[0] + c["pen"][0], rel_tol=1e-12)
    assert y1[1] == y0[1]


def test_velocity_tie_goes_to_the_lowest_index():
    """Equal desired velocity changes, penetrations below positionEpsilon, one
    iteration: only contact 0's body gets an impulse (contact.go:394-405)."""
    o = world([{"shape": "sphere", "r": 0.5, "mass": 1.0, "pos": (-3, 0.495, 0), "vel": (0, -1.0, 0)},
               {"shape": "sphere", "r": 0.5, "mass": 1.0, "pos": (3, 0.495, 0), "vel": (0, -1.0, 0)}],
              gravity=0.0)
    c = o.narrowphase()
    assert len(c) == 2
    o.resolve(c, DT, 1)
    v = o.get("vel")
    assert v[0, 1] > 0.0
    assert v[1, 1] == -1.0


def test_sat_tie_keeps_the_earlier_axis():
    """Axis-aligned unit cubes offset by (1.5, 1.5, 0) overlap 0.5 on x and
    on y: the axes 0, 1, 3 and 4 tie and the strict '<' of tryAxis
    (colliders.go:472) keeps axis 0, the first cube's x face, so the normal
    is -x (opposing the centre offset), the penetration 0.5 and the point
    the second cube's vertex (-1, +1, +1) in its frame: (0.5, 7.5, 1).  The
    later edge axis z x y = -x ties too and has the same normal; only the
    point (an edge midpoint) would tell it apart."""
    o = world([{"half": (1, 1, 1), "mass": 1.0, "pos": (0, 5, 0)},
               {"half": (1, 1, 1), "mass": 1.0, "pos": (1.5, 6.5, 0)}], planes=(), gravity=0.0)
    c = o.narrowphase_pairs([0], [1])
    assert len(c) == 1
    assert np.array_equal(c["normal"][0], [-1.0, 0.0, 0.0])
    assert c["pen"][0] == 0.5
    assert np.array_equal(c["point"][0], [0.5, 7.5, 1.0])


@pytest.mark.parametrize("vy, want", [(-0.2, 0.0), (-2.0, 0.2)])
def test_restitution_only_above_the_velocity_limit(vy, want):
    """calculateDesiredDeltaVelocity drops restitution when the closing speed
    is below 0.25 m/s (contact.go:87-112): a slow contact ends at rest along
    the normal, a fast one bounces at 0.1 of its approach speed."""
    o = world([{"shape": "sphere", "r": 0.5, "mass": 1.0, "pos": (0, 0.495, 0), "vel": (0, vy, 0)}], gravity=0.0)
    c = o.narrowphase()
    assert len(c) == 1
    o.resolve(c, DT, 4)
    assert math.isclose(o.get("vel")[0, 1], want, rel_tol=1e-12, abs_tol=1e-12)


def test_two_body_impulse_conserves_momentum():
    m0, m1 = 2.0, 5.0
    o = world([{"shape": "sphere", "r": 1.0, "mass": m0, "pos": (0, 10, 0), "vel": (3, 0.2, 0)},
               {"shape": "sphere", "r": 1.0, "mass": m1, "pos": (1.9, 10, 0), "vel": (-1, 0, 0.5)}],
              planes=(), gravity=0.0)
    o.integrate(DT)
    c = o.narrowphase()[:1]  # one contact of the ordered pair
    p0 = m0 * o.get("vel")[0] + m1 * o.get("vel")[1]
    o.resolve(c, DT, 1)
    p1 = m0 * o.get("vel")[0] + m1 * o.get("vel")[1]
    assert np.allclose(p0, p1, atol=1e-12)


def test_resolver_iteration_budget_and_guard():
    sc = scenes.cubedrop(1, 24)
    o = OracleWorld(sc)
    o.calculate_derived()
    o.step(sc.dt, 240, 4)
    o.integrate(sc.dt)
    c = o.narrowphase()
    before = o.counters()
    o.resolve(c, sc.dt, 3)
    after = o.counters()
    assert after["pos_iters"] - before["pos_iters"] <= 3
    assert after["vel_iters"] - before["vel_iters"] <= 3
    st = o.get("pos")
    o.resolve(c, 0.0, 100)  # duration <= 0: no-op (contact.go:210-212)
    assert np.array_equal(st, o.get("pos"))


@pytest.mark.parametrize("seed", [0, 1])
def test_cubedrop_settles(seed):
    """Whole-trajectory sanity: after 600 steps every cube rests on the
    ground (height ~ its half-extent) and most have fallen asleep."""
    sc = scenes.cubedrop(seed)
    o = OracleWorld(sc)
    o.calculate_derived()
    o.step(sc.dt, 600, 4)
    y = o.get("pos")[:, 1]
    assert np.all(y > 0.5) and np.all(y < 8.0)
    assert np.mean((o.get("flags") & 1) == 0) > 0.5
