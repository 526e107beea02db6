"""Pins of the oracle's per-contact friction and restitution and of its
frictionless impulse (calculateFrictionlessImpulse, contact.go:498-531),
which the reference never reaches because every contact it makes carries
Friction 0.9 (colliders.go:201).

The expected values come from rigid-body mechanics computed here in numpy
(the response of each body to a unit impulse applied at the contact), not
from the oracle's formulas, so a dropped term, a wrong sign or the wrong
body fails them."""
import math

import numpy as np

from oracle.cbz_oracle import FRICTIONLESS_TWO_BODY
from tests.test_oracle_physics import DT, world


def _two_cubes(masses=(2.0, 5.0)):
    q = (math.cos(0.3), math.sin(0.3) * 0.6, math.sin(0.3) * 0.8, 0.0)
    o = world([{"half": (1.0, 0.7, 0.5), "mass": masses[0], "pos": (0, 5, 0), "quat": q, "vel": (2.0, 0.5, 0.0),
                "rot": (0.3, -0.2, 0.1)},
               {"half": (0.6, 0.9, 1.1), "mass": masses[1], "pos": (1.4, 5.6, 0.2), "vel": (-1.5, 0.0, 0.4),
                "rot": (0.0, 0.4, -0.3)}], planes=(), gravity=0.0)
    o.masses = masses
    c = o.narrowphase_pairs([0], [1])
    assert len(c) == 1 and c["body0"][0] >= 0 and c["body1"][0] >= 0
    c["pen"][0] = 0.0  # below positionEpsilon: the position phase does nothing
    return o, c


def _one_frictionless_impulse(flags):
    o, c = _two_cubes()
    c["friction"][0] = 0.0
    c["restitution"][0] = 0.3
    b0, b1 = int(c["body0"][0]), int(c["body1"][0])
    if b0 != 0:  # keep body 0 of the test = Bodies[0]
        o.masses = o.masses[::-1]
    p = c["point"][0]
    n = c["normal"][0]
    x = o.get("pos")
    r0, r1 = p - x[b0], p - x[b1]
    vel0 = o.get("vel").copy()
    perm = [b0, b1]

    def rel():
        v, w = o.get("vel")[perm], o.get("rot")[perm]
        return (v[0] + np.cross(w[0], r0)) - (v[1] + np.cross(w[1], r1))

    closing = float(n @ rel())
    assert abs(closing) >= 0.25  # restitution applies
    ddv = -closing - 0.3 * closing  # contact.go:87-112, lastFrameAcceleration = 0

    def k(i, r):
        """n . (change of body i's contact-point velocity) for a unit
        impulse n applied to it at offset r"""
        inv_m = 1.0 / o.masses[i]
        iw = o.get("iit_world")[perm[i]].reshape(3, 3, order="F")
        return float(n @ (n * inv_m + np.cross(iw @ np.cross(r, n), r)))

    k0, k1 = k(0, r0), k(1, r1)
    o.resolve(c, DT, 1, flags=flags)
    dv = o.get("vel") - vel0
    return n, float(n @ rel()) - closing, ddv, k0, k1, dv[perm], np.array(o.masses)


def test_frictionless_two_body_flag_closes_by_the_desired_change():
    """With the second body's terms (CBZO_FRICTIONLESS_TWO_BODY) one impulse
    changes the normal closing speed by exactly desiredDeltaVelocity."""
    n, change, ddv, k0, k1, dv, m = _one_frictionless_impulse(FRICTIONLESS_TWO_BODY)
    assert math.isclose(change, ddv, rel_tol=1e-12)
    assert np.allclose(m[0] * dv[0] + m[1] * dv[1], 0.0, atol=1e-12)


def test_frictionless_reference_quirk_leaves_out_the_second_body():
    """The reference's inverted test (contact.go:512) sizes the impulse by
    body 0's response alone: the closing speed changes by ddv (k0 + k1) / k0,
    the impulse stays along the normal and momentum is still conserved."""
    n, change, ddv, k0, k1, dv, m = _one_frictionless_impulse(0)
    assert k1 > 0.01 * k0
    assert math.isclose(change, ddv * (k0 + k1) / k0, rel_tol=1e-11)
    t = dv[0] - (dv[0] @ n) * n
    assert np.linalg.norm(t) < 1e-12 * np.linalg.norm(dv[0])
    assert np.allclose(m[0] * dv[0] + m[1] * dv[1], 0.0, atol=1e-12)


def _sphere_on_plane(vel, friction, restitution):
    o = world([{"shape": "sphere", "r": 0.5, "mass": 2.0, "pos": (0, 0.495, 0), "vel": vel}], gravity=0.0)
    c = o.narrowphase()
    assert len(c) == 1 and c["body1"][0] == -1
    c["friction"][0] = friction
    c["restitution"][0] = restitution
    return o, c


def test_frictionless_half_space_contact_keeps_tangential_velocity():
    """A frictionless sphere-plane contact: the impulse is normal only, so
    the sliding velocity and the spin are untouched, and the bounce is
    -restitution times the approach speed."""
    for flags in (0, FRICTIONLESS_TWO_BODY):
        o, c = _sphere_on_plane((1.5, -2.0, 0.7), 0.0, 0.35)
        o.resolve(c, DT, 4, flags=flags)
        v = o.get("vel")[0]
        assert v[0] == 1.5 and v[2] == 0.7
        assert math.isclose(v[1], 0.35 * 2.0, rel_tol=1e-12)
        assert np.array_equal(o.get("rot")[0], [0.0, 0.0, 0.0])
        # the reference would have dereferenced nil here (contact.go:512)
        assert o.counters()["panics"] >= 1


def test_per_contact_restitution():
    o, c = _sphere_on_plane((0.0, -3.0, 0.0), 0.9, 0.6)
    o.resolve(c, DT, 4)
    assert math.isclose(o.get("vel")[0, 1], 0.6 * 3.0, rel_tol=1e-12)


def test_per_contact_friction_slides_on_the_cone():
    """Low friction and fast sliding: the impulse lies on the friction cone,
    |J_t| = mu |J_n| (contact.go:585-600).  A sphere's linear velocity
    change is J / m, so |dv_t| / |dv_n| = mu."""
    mu = 0.1
    o, c = _sphere_on_plane((4.0, -2.0, 1.0), mu, 0.0)
    v0 = o.get("vel")[0].copy()
    o.resolve(c, DT, 1)
    dv = o.get("vel")[0] - v0
    assert math.isclose(math.hypot(dv[0], dv[2]) / abs(dv[1]), mu, rel_tol=1e-9)
    # opposing the sliding direction
    assert dv[0] < 0 and dv[2] < 0 and math.isclose(dv[0] / dv[2], 4.0, rel_tol=1e-9)


def test_high_friction_sticks():
    """Slow sliding under friction 0.9: the sticking impulse removes the
    contact point's tangential velocity (rolling without slipping)."""
    o, c = _sphere_on_plane((0.2, -2.0, -0.1), 0.9, 0.0)
    r = c["point"][0] - o.get("pos")[0]
    o.resolve(c, DT, 1)
    v, w = o.get("vel")[0], o.get("rot")[0]
    vp = v + np.cross(w, r)
    assert abs(vp[0]) < 1e-12 and abs(vp[2]) < 1e-12
