"""Per-contact friction and restitution on the HIP path, against the FP64
oracle (bit for bit).

The reference gives every contact Friction 0.9 and Restitution 0.1
(colliders.go:201-202); with materials on, a contact carries the means of
its body's and its partner's (body or plane) values, and a contact of
friction 0 takes calculateFrictionlessImpulse (contact.go:498-531), in the
reference's form (second body left out, :512) or, with the flag, with it.
The oracle resolves the same records, with the friction and restitution the
combination rule gives, so both the narrowphase's material output and the
resolver's use of it are checked."""
import numpy as np
import pytest

from cubez_amd import World, scenes
from cubez_amd._native import CubezError
from oracle.cbz_oracle import FRICTIONLESS_TWO_BODY, OracleWorld
from tests.test_gpu_parity import STATE, assert_state_equal, same

pytestmark = pytest.mark.gpu


def combine(cs, f, e, pf, pe):
    """The mean of the two sides' materials; f, e are the world's per-body
    values (world-local indices), pf, pe its plane's."""
    b0, b1 = cs["body0"], cs["body1"]
    f1 = np.where(b1 >= 0, f[np.maximum(b1, 0)], pf)
    e1 = np.where(b1 >= 0, e[np.maximum(b1, 0)], pe)
    return (f[b0] + f1) * 0.5, (e[b0] + e1) * 0.5


def materials(sc, seed):
    """Per-body friction in [0, 1.2] with a quarter of the bodies (and every
    body of world 1) at 0; restitution in [0, 0.5]; plane values per world,
    some planes frictionless."""
    rng = np.random.default_rng(seed)
    n = sc.n_bodies
    f = rng.uniform(0.0, 1.2, n)
    f[rng.random(n) < 0.25] = 0.0
    if sc.n_worlds > 1:
        f[sc.world_slice(1)] = 0.0
    e = rng.uniform(0.0, 0.5, n)
    pf = np.array([[0.0, 0.5, 0.9, 0.0, 1.0, 0.3][w % 6] for w in range(sc.n_worlds)])
    pe = np.array([[0.1, 0.3, 0.0, 0.5, 0.2, 0.05][w % 6] for w in range(sc.n_worlds)])
    return f, e, pf, pe


def test_default_materials_reproduce_the_reference():
    """Materials on with every value at its default: the same trajectory,
    bit for bit, as the constant path, and contacts reading 0.9 / 0.1."""
    sc = scenes.batched(3, 8, 8, 96)
    with World(sc) as a, World(sc) as b:
        b.set_materials(True)
        ca = a.step(sc.dt, 150, 4, counts=True)
        cb = b.step(sc.dt, 150, 4, counts=True)
        assert np.array_equal(ca, cb)
        for k in STATE:
            assert same(a.get(k), b.get(k)), k
        c = b.contacts(0)
        assert len(c["pen"]) > 0
        assert np.all(c["friction"] == 0.9) and np.all(c["restitution"] == 0.1)
        c = a.contacts(0)
        assert np.all(c["friction"] == 0.9) and np.all(c["restitution"] == 0.1)


@pytest.mark.parametrize("two_body", [False, True])
def test_material_steps_match_oracle(two_body):
    sc = scenes.batched(11, 4, 24, 64)
    f, e, pf, pe = materials(sc, 1)
    steps = 200
    with World(sc) as w:
        w.set_materials(True, frictionless_two_body=two_body)
        w.set("friction", f)
        w.set("restitution", e)
        w.set_plane_materials(pf, pe)
        gcounts = w.step(sc.dt, steps, 4, counts=True)
        flags = FRICTIONLESS_TWO_BODY if two_body else 0
        zero_two_body = 0
        for wi in range(sc.n_worlds):
            sl = sc.world_slice(wi)
            o = OracleWorld(sc, wi)
            o.calculate_derived()
            for s in range(steps):
                o.integrate(sc.dt)
                cs = o.narrowphase()
                assert len(cs) == gcounts[s, wi], f"world {wi} step {s} contact count"
                cs["friction"], cs["restitution"] = combine(cs, f[sl], e[sl], pf[wi], pe[wi])
                zero_two_body += int(((cs["friction"] == 0) & (cs["body1"] >= 0)).sum())
                o.resolve(cs, sc.dt, 4 * len(cs), flags=flags)
            assert_state_equal(w, o, sl.start, sl.stop - sl.start)
        assert zero_two_body > 0  # the frictionless two-body path ran


def test_narrowphase_writes_the_combined_materials():
    sc = scenes.batched(11, 4, 24, 64)
    f, e, pf, pe = materials(sc, 2)
    with World(sc) as w:
        w.step(sc.dt, 150, 4)
        w.set_materials(True)
        w.set("friction", f)
        w.set("restitution", e)
        w.set_plane_materials(pf, pe)
        w.integrate(sc.dt)
        w.narrowphase()
        for wi in range(sc.n_worlds):
            sl = sc.world_slice(wi)
            c = w.contacts(wi, internals=False)
            assert len(c["pen"]) > 0
            loc = {"body0": c["body0"] - sl.start, "body1": np.where(c["body1"] >= 0, c["body1"] - sl.start, -1)}
            wf, we = combine(loc, f[sl], e[sl], pf[wi], pe[wi])
            assert same(c["friction"], wf) and same(c["restitution"], we), wi


@pytest.mark.parametrize("two_body", [False, True])
def test_contact_material_override_matches_oracle(two_body):
    """Per-contact values set after the narrowphase (a Go caller assigning
    Contact.Friction / Restitution before ResolveContacts)."""
    sc = scenes.cubedrop(0)
    with World(sc) as w:
        o = OracleWorld(sc)
        o.calculate_derived()
        w.step(sc.dt, 200, 4)
        o.step(sc.dt, 200, 4)
        w.set_materials(True, frictionless_two_body=two_body)
        w.integrate(sc.dt)
        o.integrate(sc.dt)
        w.narrowphase()
        cs = o.narrowphase()
        rng = np.random.default_rng(5)
        fr = rng.uniform(0.0, 1.0, len(cs))
        fr[rng.random(len(cs)) < 0.4] = 0.0
        re = rng.uniform(0.0, 0.8, len(cs))
        w.set_contact_materials(fr, re)
        g = w.contacts(0, internals=False)
        assert same(g["friction"], fr) and same(g["restitution"], re)
        cs["friction"], cs["restitution"] = fr, re
        w.resolve(sc.dt, 4)
        out = o.resolve(cs, sc.dt, 4 * len(cs), flags=FRICTIONLESS_TWO_BODY if two_body else 0)
        assert_state_equal(w, o, 0, sc.n_bodies)
        gc = w.contacts(0)
        for k in ("pen", "cvel", "ddv"):
            assert same(gc[k], out[k]), k


def test_material_graph_replays_equal_eager_steps():
    sc = scenes.batched(4, 6, 8, 64)
    f, e, pf, pe = materials(sc, 3)
    with World(sc) as a, World(sc) as b:
        for w in (a, b):
            w.set_materials(True)
            w.set("friction", f)
            w.set("restitution", e)
            w.set_plane_materials(pf, pe)
        a.step(sc.dt, 120, 4)  # captured graph, replayed
        b.step(sc.dt, 120, 4, counts=True)  # eager
        for k in STATE:
            assert same(a.get(k), b.get(k)), k


def test_fp32_materials_track_fp64_early():
    sc = scenes.batched(4, 6, 8, 64)
    f, e, pf, pe = materials(sc, 4)
    runs = {}
    for prec in ("f64", "f32"):
        with World(sc, precision=prec) as w:
            w.set_materials(True)
            w.set("friction", f)
            w.set("restitution", e)
            w.set_plane_materials(pf, pe)
            w.step(sc.dt, 40, 4, counts=True)
            runs[prec] = w.get("pos")
    assert np.all(np.isfinite(runs["f32"]))
    assert np.max(np.abs(runs["f32"] - runs["f64"])) < 1e-2


def test_learned_resolver_refuses_materials():
    sc = scenes.cubedrop(0)
    with World(sc) as w:
        w.set_materials(True)
        w.integrate(sc.dt)
        w.narrowphase()
        with pytest.raises(CubezError):
            w.sur_resolve(sc.dt)
        with pytest.raises(CubezError):
            w.set_materials(False, frictionless_two_body=True)
        with pytest.raises(CubezError):
            w.set_plane_materials([-1.0])
