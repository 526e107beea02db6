"""Pins of the oracle's box-box contact generation and of the resolver steps
that whole-trajectory checks leave open: the edge-edge contact
(colliders.go:642-709), the SAT's near-parallel threshold (colliders.go:461),
the inertia split and angular limit of applyPositionChange
(contact.go:286-386), the derived data a position change refreshes
(contact.go:380-382), matchAwakeState (contact.go:185-202) and the desired
velocity change of two awake bodies (contact.go:87-112).

Expected values come from geometry and rigid-body mechanics computed here
(numpy, an LP feasibility test), chosen so that a plausible transcription
mistake in the oracle fails one of them (tests/test_oracle_mutants.py plants
such mistakes and runs these checks)."""
import math

import numpy as np
import pytest

from tests.test_oracle_physics import DT, boxes_overlap_lp, world

SQRT2 = math.sqrt(2.0)


def q_axis(angle, axis):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    return (math.cos(angle / 2), *(math.sin(angle / 2) * a))


def q_from_matrix(R):
    """Unit quaternion (w, x, y, z) of a rotation matrix."""
    w = math.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    x = math.copysign(math.sqrt(max(0.0, 1 + R[0, 0] - R[1, 1] - R[2, 2])) / 2, R[2, 1] - R[1, 2])
    y = math.copysign(math.sqrt(max(0.0, 1 - R[0, 0] + R[1, 1] - R[2, 2])) / 2, R[0, 2] - R[2, 0])
    z = math.copysign(math.sqrt(max(0.0, 1 - R[0, 0] - R[1, 1] + R[2, 2])) / 2, R[1, 0] - R[0, 1])
    q = np.array([w, x, y, z])
    return q / np.linalg.norm(q)


def rot_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def boxes(o, halves):
    T = o.get("ctransform")
    return [(T[i, :9].reshape(3, 3, order="F"), T[i, 9:], np.asarray(halves[i], float)) for i in range(len(halves))]


# --- box-box contact generation ----------------------------------------------

@pytest.mark.parametrize("d", [0.1, 0.37])
def test_edge_edge_crossing_contact(d):
    """Cube one is turned 45 degrees about z (its top edge runs along z at
    height sqrt2), cube two 45 degrees about x (its bottom edge runs along x)
    and sits above, the edges crossing at depth d.  The minimal SAT axis is
    the edge cross product y; the contact normal points from box two towards
    box one (-y, colliders.go:653-655); the point is the midpoint of the
    closest points of the two edges, (0, sqrt2 - d/2, 0) (contactPoint,
    colliders.go:519-573, from the edge midpoints chosen at :660-677).  With
    the operands swapped the normal turns and the point is the same."""
    o = world([{"half": (1, 1, 1), "mass": 1, "pos": (0, 0, 0), "quat": q_axis(math.pi / 4, (0, 0, 1))},
               {"half": (1, 1, 1), "mass": 1, "pos": (0, 2 * SQRT2 - d, 0), "quat": q_axis(math.pi / 4, (1, 0, 0))}],
              planes=(), gravity=0.0)
    for one, two, ny in ((0, 1, -1.0), (1, 0, 1.0)):
        c = o.narrowphase_pairs([one], [two])
        assert len(c) == 1
        assert np.allclose(c["normal"][0], [0.0, ny, 0.0], atol=1e-12)
        assert math.isclose(c["pen"][0], d, rel_tol=1e-12)
        assert np.allclose(c["point"][0], [0.0, SQRT2 - d / 2, 0.0], atol=1e-12)
        assert c["body0"][0] == one and c["body1"][0] == two


def test_box_box_normal_points_from_the_second_box_to_the_first():
    """For every cube-cube contact of random overlapping pairs (face and
    edge cases alike), moving the contact's Bodies[0] by (pen + 0.25) along
    the contact normal separates the boxes: the normal is the minimal-overlap
    direction and points from Bodies[1] towards Bodies[0]
    (colliders.go:485-488, :653-655; a vertex of box one on a face of box two
    swaps the bodies, :632-637).  A normal of the wrong sign pushes the body
    further in."""
    rng = np.random.default_rng(21)
    seen = 0
    for trial in range(120):
        halves, specs = [], []
        for k in range(2):
            q = rng.standard_normal(4)
            halves.append(tuple(rng.uniform(0.3, 1.5, 3)))
            specs.append({"half": halves[-1], "mass": 1.0, "pos": tuple(rng.uniform(-1.2, 1.2, 3) + [0, 50, 0]),
                          "quat": tuple(q / np.linalg.norm(q))})
        o = world(specs, planes=())
        c = o.narrowphase_pairs([0], [1])
        if len(c) == 0:
            continue
        bx = boxes(o, halves)
        if c["body0"][0] == 1:
            bx = bx[::-1]
        (RA, tA, hA), (RB, tB, hB) = bx
        n, pen = c["normal"][0], c["pen"][0]
        if pen < 0.02:
            continue
        seen += 1
        assert not boxes_overlap_lp(RA, tA + (pen + 0.25) * n, hA, RB, tB, hB, -0.02), trial
        # and the overlap is real: moved by half the penetration it still overlaps
        assert boxes_overlap_lp(RA, tA + 0.5 * pen * n, hA, RB, tB, hB, 0.0), trial
    assert seen > 40


def _rods(s, L=200.0, w=0.1, phi=5.3270513720613675, psi=5.344784187709867, gap=0.04):
    """Two square rods of half-size (L, w, w): rod one along x at the origin,
    rod two along b = (cos t, sin t cos phi, sin t sin phi) with sin t = s,
    rolled by psi about b, its centre at the distance along u = x cross b
    (normalised) that leaves `gap` between their projections on u."""
    th = math.asin(s)
    b = np.array([math.cos(th), math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi)])
    u = np.cross([1.0, 0.0, 0.0], b)
    u /= np.linalg.norm(u)
    v = np.cross(b, u)
    y = math.cos(psi) * u + math.sin(psi) * v
    z = np.cross(b, y)
    R = np.stack([b, y, z], axis=1)
    D = w * (abs(u[1]) + abs(u[2])) + w * (abs(y @ u) + abs(z @ u)) + gap
    o = world([{"half": (L, w, w), "mass": 1, "pos": (0, 0, 0)},
               {"half": (L, w, w), "mass": 1, "pos": tuple(D * u), "quat": tuple(q_from_matrix(R))}],
              planes=(), gravity=0.0)
    (RA, tA, hA), (RB, tB, hB) = boxes(o, [(L, w, w), (L, w, w)])
    apart = not boxes_overlap_lp(RA, tA, hA, RB, tB, hB, -0.01)
    return o, apart


def test_sat_skips_exactly_the_nearly_parallel_edge_axes():
    """tryAxis skips an axis whose squared length is below Epsilon = 1e-7
    (colliders.go:460-462).  Two long rods at an angle t apart, separated by
    4 cm only along the cross product of their long edges (|x cross b|^2 =
    sin^2 t; every other SAT axis overlaps):
      sin^2 t = 1.16e-7: the axis is tried, so the pair is separate;
      sin^2 t = 0.90e-7: the axis is skipped and the reference reports a
      contact between boxes that an LP shows to be apart.
    So the threshold lies in (0.90e-7, 1.16e-7]."""
    o, apart = _rods(math.sqrt(1.16e-7))
    assert apart
    assert len(o.narrowphase_pairs([0], [1])) == 0
    o, apart = _rods(math.sqrt(0.90e-7))
    assert apart
    assert len(o.narrowphase_pairs([0], [1])) == 1


# --- applyPositionChange -------------------------------------------------------

def _vertex_on_plane(pen, mass=8.0, half=(1.0, 0.6, 0.8), q=None, awake=True):
    """A cube turned so that exactly one vertex is below the plane y = 0, by
    pen; returns the world, the one contact, and in numpy the vertex's offset
    r from the centre and the world inverse inertia tensor."""
    q = q_axis(0.7, (1.0, 0.35, 0.6)) if q is None else q
    R = rot_matrix(q)
    h = np.array(half)
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * h
    world_off = corners @ R.T
    k = int(np.argmin(world_off[:, 1]))
    ys = np.sort(world_off[:, 1])
    assert ys[1] - ys[0] > pen + 0.05  # only one vertex goes below the plane
    centre = np.array([0.3, -world_off[k, 1] - pen, -0.2])
    o = world([{"half": tuple(half), "mass": mass, "pos": tuple(centre), "quat": tuple(q)}], gravity=0.0)
    if not awake:
        o.a["flags"][0] &= ~np.uint8(1)
    c = o.narrowphase()
    assert len(c) == 1 and math.isclose(c["pen"][0], pen, rel_tol=1e-9)
    r = c["point"][0] - centre
    # the world inverse inertia is an input here (pinned against R I^-1 R^T
    # by test_oracle_physics.test_world_inverse_inertia)
    iw = o.get("iit_world")[0].reshape(3, 3, order="F")
    return o, c, r, iw


def _expected_moves(pen, mass, r, iw, limit_factor=0.2):
    """applyPositionChange for one body on a half-space, from the mechanics:
    the penetration is split between linear and angular motion in the ratio
    of the body's linear (1/m) and angular (n . ((I^-1 (r x n)) x r)) inverse
    inertia along the normal, the angular part capped at 0.2 |r_perp|."""
    n = np.array([0.0, 1.0, 0.0])
    ang_i = float(n @ np.cross(iw @ np.cross(r, n), r))
    lin_i = 1.0 / mass
    ang = pen * ang_i / (ang_i + lin_i)
    lin = pen * lin_i / (ang_i + lin_i)
    cap = limit_factor * np.linalg.norm(r - (r @ n) * n)
    clamped = abs(ang) > cap
    if clamped:
        total = ang + lin
        ang = math.copysign(cap, ang)
        lin = total - ang
    return lin, ang, ang_i, clamped


@pytest.mark.parametrize("pen, want_clamp", [(0.04, False), (0.6, True)])
def test_position_change_splits_by_inertia_within_the_angular_limit(pen, want_clamp):
    """One position iteration on a single vertex contact moves the centre
    along the normal by the linear share of the penetration; the angular
    share is capped at angularLimit = 0.2 times the contact offset
    perpendicular to the normal, the rest going to the linear move.  The
    small penetration stays under the cap (0.02 |r_perp| would not), the
    large one is capped."""
    mass = 8.0
    o, c, r, iw = _vertex_on_plane(pen, mass=mass)
    lin, ang, ang_i, clamped = _expected_moves(pen, mass, r, iw)
    assert clamped == want_clamp
    if not want_clamp:  # a cap ten times smaller would engage here
        assert ang > 0.02 * math.hypot(r[0], r[2])
    y0 = o.get("pos")[0, 1]
    o.resolve(c, DT, 1)
    assert math.isclose(o.get("pos")[0, 1] - y0, lin, rel_tol=1e-9)


def test_position_change_removes_its_own_penetration_and_turns_by_the_inverse_inertia():
    """After one position iteration the resolved contact's penetration is 0
    (the linear move plus the angular move, angularChange =
    I^-1 (r x n) * angularMove / angularInertia, adds up to it exactly), and
    the orientation is q + 1/2 (0, dtheta) q, normalised, with that
    dtheta."""
    pen, mass = 0.04, 8.0
    o, c, r, iw = _vertex_on_plane(pen, mass=mass)
    lin, ang, ang_i, _ = _expected_moves(pen, mass, r, iw)
    n = np.array([0.0, 1.0, 0.0])
    dtheta = iw @ np.cross(r, n) * (ang / ang_i)
    q = o.get("quat")[0].copy()
    out = o.resolve(c, DT, 1)
    assert abs(out["pen"][0]) < 1e-12
    w, x, y, z = q
    dq = 0.5 * np.array([-dtheta @ [x, y, z],
                         dtheta[0] * w + dtheta[1] * z - dtheta[2] * y,
                         dtheta[1] * w + dtheta[2] * x - dtheta[0] * z,
                         dtheta[2] * w + dtheta[0] * y - dtheta[1] * x])
    want = q + dq
    want /= np.linalg.norm(want)
    assert np.allclose(o.get("quat")[0], want, atol=1e-12)


def test_position_change_refreshes_derived_data_of_sleeping_bodies_only():
    """applyPositionChange recomputes the transform and world inverse
    inertia of a body that is asleep (contact.go:380-382); an awake body's
    derived data stay as they were until its next Integrate."""
    pen = 0.04
    o, c, r, iw = _vertex_on_plane(pen, awake=True)
    T0, I0 = o.get("transform").copy(), o.get("iit_world").copy()
    o.resolve(c, DT, 1)
    assert not np.array_equal(o.get("pos"), o.get("transform")[:, 9:])  # the position moved
    assert np.array_equal(o.get("transform"), T0) and np.array_equal(o.get("iit_world"), I0)

    o, c, r, iw = _vertex_on_plane(pen, awake=False)
    I0 = o.get("iit_world").copy()
    o.resolve(c, DT, 1)
    assert (o.get("flags")[0] & 1) == 0  # a half-space contact wakes nobody
    T = o.get("transform")[0]
    assert np.array_equal(T[9:], o.get("pos")[0])
    R = T[:9].reshape(3, 3, order="F")
    assert np.allclose(R, rot_matrix(o.get("quat")[0]), atol=1e-14)
    assert not np.array_equal(o.get("iit_world"), I0)


# --- matchAwakeState -------------------------------------------------------------

@pytest.mark.parametrize("sleeper", [0, 1])
def test_contact_wakes_the_sleeping_partner(sleeper):
    """A contact between an awake and a sleeping body wakes the sleeping one
    (SetAwake(true): IsAwake, motion = 2 sleepEpsilon = 0.6) when the
    resolver takes it; the awake one keeps its motion."""
    o = world([{"half": (1, 1, 1), "mass": 2.0, "pos": (0, 5, 0)},
               {"half": (1, 1, 1), "mass": 3.0, "pos": (0.3, 6.9, 0.2)}], planes=(), gravity=0.0)
    o.a["flags"][sleeper] &= ~np.uint8(1)
    o.a["motion"][:] = [1.5, 1.5]
    c = o.narrowphase_pairs([0], [1])
    assert len(c) == 1 and c["pen"][0] > 0.05
    o.resolve(c, DT, 1)
    assert np.all((o.get("flags") & 1) == 1)
    assert o.get("motion")[sleeper] == 0.6
    assert o.get("motion")[1 - sleeper] == 1.5


# --- calculateDesiredDeltaVelocity -------------------------------------------------

def test_desired_velocity_change_of_two_awake_bodies():
    """Two awake spheres meeting head-on along x with different
    lastFrameAccelleration: the desired change is -v - e (v - v_acc) with v
    the closing velocity along the normal, e = 0.1 (|v| >= 0.25) and
    v_acc = (a0 dt - a1 dt) . n, body one's acceleration entering with a
    minus sign (contact.go:93-102)."""
    a0, a1 = np.array([3.0, -9.8, 0.0]), np.array([-7.0, 4.0, 1.0])
    o = world([{"shape": "sphere", "r": 1.0, "mass": 2.0, "pos": (0, 10, 0), "vel": (1.0, 0.0, 0.0)},
               {"shape": "sphere", "r": 1.0, "mass": 5.0, "pos": (1.99, 10, 0), "vel": (-0.5, 0.0, 0.0)}],
              planes=())
    o.a["acc"][:] = np.concatenate([a0, a1])
    o.integrate(DT)
    assert np.array_equal(o.get("last_acc"), np.stack([a0, a1]))
    c = o.narrowphase_pairs([0], [1])
    assert len(c) == 1 and c["body0"][0] == 0
    n = c["normal"][0]
    v = o.get("vel")
    closing = float((v[0] - v[1]) @ n)
    assert abs(closing) >= 0.25
    v_acc = float((a0 * DT) @ n - (a1 * DT) @ n)
    want = -closing - 0.1 * (closing - v_acc)
    out = o.resolve(c, DT, 0)  # maxIterations 0: prepareContacts only
    assert math.isclose(out["ddv"][0], want, rel_tol=1e-12)
