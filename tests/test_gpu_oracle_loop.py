"""The oracle in the loop for the callers and the rows around the step:
the ballistic example's own pair order with its static, infinite-mass
backboard; the sort-and-sweep broadphase against the oracle's all-pairs
narrowphase; a checkpoint resume against the oracle's uninterrupted run;
and the NewRigidBody defaults cbz_set_scene gives every body.  Every
comparison is bit-exact (FP64)."""
import numpy as np
import pytest

from cubez_amd import World, scenes
from oracle.cbz_oracle import OracleWorld
from tests.test_gpu_parity import STATE, assert_state_equal, same

pytestmark = pytest.mark.gpu


def test_ballistic_example_pair_order_with_static_backboard():
    """updateCallback of examples/ballistic.go:85-91: updateObjects
    integrates the cube and the bullets (the backboard never), then
    generateContacts checks the example's pairs in its order, then
    ResolveContacts(8 * contacts) when any were found.  The bullets hit the
    backboard and the cube, bounce, and come to rest on the ground."""
    nb = 12
    sc = scenes.ballistic_example(nb)
    one, two = scenes.ballistic_example_pairs(nb)
    hits = {"backboard": 0, "cube": 0, "bullets": 0}
    with World(sc) as w:
        o = OracleWorld(sc)
        o.calculate_derived()
        board0 = (w.get("pos")[1].copy(), w.get("quat")[1].copy())
        for step in range(150):
            w.integrate(sc.dt)
            o.integrate(sc.dt)
            n = w.narrowphase_pairs(one, two)
            oc = o.narrowphase_pairs(one, two)
            assert n == len(oc), step
            if n:
                gc = w.contacts(0, internals=False)
                assert np.array_equal(gc["body0"], oc["body0"]) and np.array_equal(gc["body1"], oc["body1"])
                assert same(gc["pen"], oc["pen"]) and same(gc["normal"], oc["normal"].reshape(-1, 3))
                b = set(zip(gc["body0"].tolist(), gc["body1"].tolist()))
                hits["backboard"] += any(1 in p for p in b)
                hits["cube"] += any(0 in p and -1 not in p and 1 not in p for p in b)
                hits["bullets"] += any(p[0] >= 2 and p[1] >= 2 for p in b)
                w.resolve(sc.dt, 8, 0)
                o.resolve(oc, sc.dt, 8 * len(oc))
            if step % 10 == 9:
                assert_state_equal(w, o, 0, sc.n_bodies)
        assert_state_equal(w, o, 0, sc.n_bodies)
        # the backboard took hits and stayed where it was, to the bit
        assert hits["backboard"] > 0 and hits["cube"] > 0
        assert np.array_equal(w.get("pos")[1], board0[0]) and np.array_equal(w.get("quat")[1], board0[1])
        print("steps with contacts on: backboard", hits["backboard"], "cube", hits["cube"],
              "bullet-bullet", hits["bullets"])


def test_per_call_answers_of_a_batched_pair_list():
    """What a Go caller's per-pair calls need from one batched narrowphase
    (the lazy flush of a drop-in package): the ballistic example's frame,
    its pairs run as one cbz_narrowphase_pairs, then cbz_pair_contact_counts
    splits the contacts back into the example's calls.  Every call's bool
    ("at least one contact appended", colliders.go:720-747) and the contacts
    it appends equal the oracle's CheckForCollisions on that one pair, bit
    for bit, for 120 frames."""
    nb = 12
    sc = scenes.ballistic_example(nb)
    one, two = scenes.ballistic_example_pairs(nb)
    found_calls = 0
    with World(sc) as w:
        o = OracleWorld(sc)
        o.calculate_derived()
        for step in range(120):
            w.integrate(sc.dt)
            o.integrate(sc.dt)
            n = w.narrowphase_pairs(one, two)
            counts = w.pair_contact_counts(len(one))
            assert int(counts.sum()) == n, step
            gc = w.contacts(0, internals=False) if n else None
            off = 0
            for k in range(len(one)):
                oc = o.narrowphase_pairs(one[k:k + 1], two[k:k + 1])
                assert (counts[k] > 0) == (len(oc) > 0), (step, k)
                assert counts[k] == len(oc), (step, k)
                if len(oc):
                    found_calls += 1
                    sl = slice(off, off + counts[k])
                    assert np.array_equal(gc["body0"][sl], oc["body0"]) and np.array_equal(gc["body1"][sl], oc["body1"])
                    assert same(gc["pen"][sl], oc["pen"]) and same(gc["normal"][sl], oc["normal"].reshape(-1, 3))
                    assert same(gc["point"][sl], oc["point"].reshape(-1, 3))
                off += counts[k]
            # the frame's resolve, from the same batched contacts
            oc = o.narrowphase_pairs(one, two)
            if n:
                w.resolve(sc.dt, 8, 0)
                o.resolve(oc, sc.dt, 8 * len(oc))
        assert_state_equal(w, o, 0, sc.n_bodies)
    assert found_calls > 100


def test_broadphase_matches_the_oracle_all_pairs():
    """cbz_narrowphase_broad in the step loop of a 1,024-body ballistic
    world against the oracle's all-pairs step: the same contact counts every
    step and the same state."""
    sc = scenes.ballistic(2, 512, 512)
    with World(sc) as w:
        o = OracleWorld(sc)
        o.calculate_derived()
        for step in range(40):
            w.integrate(sc.dt)
            total, checked = w.narrowphase_broad()
            assert checked < sc.n_bodies * (sc.n_bodies - 1)  # it did cull
            w.resolve(sc.dt, 4, 0)
            oc = o.step(sc.dt, 1, 4)
            assert total == int(oc[0]), step
        assert_state_equal(w, o, 0, sc.n_bodies)


def test_checkpoint_resume_matches_the_oracle_uninterrupted():
    sc = scenes.cubedrop(3)
    with World(sc) as a:
        a.step(sc.dt, 160, 4)
        data = a.checkpoint()
    with World() as r:
        r.restore(data)
        rc = r.step(sc.dt, 160, 4, counts=True)[:, 0]
        o = OracleWorld(sc)
        o.calculate_derived()
        oc = o.step(sc.dt, 320, 4)
        assert np.array_equal(rc, oc[160:])
        assert_state_equal(r, o, 0, sc.n_bodies)


def test_scene_defaults_are_new_rigid_body():
    """cbz_set_scene alone: every body as NewRigidBody leaves it
    (rigidbody.go:104-114): identity orientation, LinearDamping and
    AngularDamping both defaultLinearDamping 0.95 (:107-108),
    defaultAcceleration (0,-9.78,0) (:18), an identity world inverse
    inertia tensor (:110), CanSleep and awake with motion 2*sleepEpsilon =
    0.6 (:111-112, :182-187); everything else zero; contact friction 0.9
    and restitution 0.1 (colliders.go:201-202)."""
    sc = scenes.cubedrop(0, 5)
    with World() as w:
        w._chk(w._lib.cbz_set_scene(w._h, 1, np.ascontiguousarray(sc.body_off, np.int32).ctypes.data,
                                    np.ascontiguousarray(sc.plane_off, np.int32).ctypes.data,
                                    np.ascontiguousarray(sc.plane_normal, np.float64).ctypes.data,
                                    np.ascontiguousarray(sc.plane_offset, np.float64).ctypes.data))
        w.n_worlds, w.n_bodies = 1, 5
        eye = np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1])
        want = {
            "quat": np.tile([1.0, 0, 0, 0], (5, 1)), "acc": np.tile([0, -9.78, 0], (5, 1)),
            "lin_damp": np.full(5, 0.95), "ang_damp": np.full(5, 0.95), "iit_world": np.tile(eye, (5, 1)),
            "motion": np.full(5, 0.6), "flags": np.full(5, 3, np.uint8), "friction": np.full(5, 0.9),
            "restitution": np.full(5, 0.1),
            "pos": np.zeros((5, 3)), "vel": np.zeros((5, 3)), "rot": np.zeros((5, 3)), "inv_mass": np.zeros(5),
            "iit_body": np.zeros((5, 9)), "last_acc": np.zeros((5, 3)),
        }
        for k, v in want.items():
            g = w.get(k)
            assert np.array_equal(np.asarray(g).reshape(np.shape(v)), v), (k, g)
