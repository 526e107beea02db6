"""The resolver's multi-wave kernel (four wavefronts per world, for the
largest worlds) against the one-wavefront kernel: every kind of resolve
gives the same bits through either.

cbz_set_resolver_split(ctx, m, k) sends the k largest worlds of at least m
contacts (at most 2048, and 256 bodies) to the multi-wave kernel; 0 sends
none.  Each test settles a scene of large worlds with the one-wavefront
kernel, then runs the same steps from that checkpoint with the split off,
at a few sizes of the multi-wave share (up to every eligible world), and
compares the whole state, the contact internals and the iteration counts."""
import numpy as np
import pytest

from cubez_amd import World, scenes, surrogate

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "quat", "vel", "rot", "iit_world", "transform", "motion", "flags")


def _scene():
    # worlds of 160..256 cubes: after settling, many hold 500..2000 contacts
    return scenes.batched(3, 24, n_min=160, n_max=256)


_SETTLED = {}


def _settled(precision):
    if precision not in _SETTLED:
        sc = _scene()
        with World(sc, precision=precision) as w:
            w.set_resolver_split(0, 0)
            w.step(sc.dt, 240, 4)
            _SETTLED[precision] = (sc, w.checkpoint())
    return _SETTLED[precision]


def _snapshot(w):
    st = {f: w.get(f) for f in FIELDS}
    st["iters"] = w.resolver_iterations()
    c = w.contacts(-1)
    for k in ("body0", "body1", "pen", "ddv", "cvel", "c2w", "rel"):
        st["c_" + k] = c[k]
    return st


def _same(a, b, what):
    for k in a:
        assert a[k].shape == b[k].shape, (what, k)
        assert np.array_equal(a[k], b[k], equal_nan=True), (what, k, np.nanmax(np.abs(
            a[k].astype(np.float64) - b[k].astype(np.float64))))


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_multiwave_resolve_equals_one_wave(precision):
    sc, ck = _settled(precision)
    out = {}
    for split in ((0, 0), (257, 5), (1, 4096)):
        with World(precision=precision) as w:
            w.restore(ck)
            w.set_resolver_split(*split)
            assert w.resolver_split() == split
            counts = w.step(sc.dt, 6, 4, counts=True)
            out[split] = _snapshot(w)
            if split == (0, 0):
                # the scene exercises the kernel: worlds above 512 contacts
                assert counts.max() > 512, counts.max()
    _same(out[(0, 0)], out[(257, 5)], "5 largest worlds")
    _same(out[(0, 0)], out[(1, 4096)], "every world")


def test_multiwave_captured_steps_equal_one_wave():
    # cbz_step without counts replays a captured graph: the fork to the
    # second stream and the join are part of it
    sc, ck = _settled("f64")
    out = {}
    for split in ((0, 0), (1, 4096)):
        with World(precision="f64") as w:
            w.restore(ck)
            w.set_resolver_split(*split)
            w.step(sc.dt, 5, 4)
            w.step(sc.dt, 5, 4)
            out[split] = _snapshot(w)
    _same(out[(0, 0)], out[(1, 4096)], "graph")


def test_multiwave_materials_equal_one_wave():
    sc, ck = _settled("f64")
    rng = np.random.default_rng(5)
    nb = sc.n_bodies
    fric = rng.uniform(0.0, 1.0, nb)
    fric[rng.random(nb) < 0.2] = 0.0  # frictionless contacts take the other impulse
    rest = rng.uniform(0.0, 0.5, nb)
    out = {}
    for split in ((0, 0), (1, 4096)):
        with World(precision="f64") as w:
            w.restore(ck)
            w.set_resolver_split(*split)
            w.set_materials(True)
            w.set("friction", fric)
            w.set("restitution", rest)
            w.step(sc.dt, 4, 4, counts=True)
            out[split] = _snapshot(w)
    _same(out[(0, 0)], out[(1, 4096)], "materials")


def test_multiwave_labels_equal_one_wave():
    sc, ck = _settled("f32")
    out = {}
    for split in ((0, 0), (1, 4096)):
        with World(precision="f32") as w:
            w.restore(ck)
            w.set_resolver_split(*split)
            w.sur_reserve(1 << 18)
            for _ in range(3):
                w.integrate(sc.dt)
                w.narrowphase()
                w.sur_collect(sc.dt, 4)
            X, Y = w.sur_dataset()
            out[split] = {"X": X, "Y": Y, **_snapshot(w)}
    _same(out[(0, 0)], out[(1, 4096)], "labels")


def test_multiwave_polish_equals_one_wave():
    params, norm = surrogate.load()
    sc, ck = _settled("f32")
    out = {}
    for split in ((0, 0), (1, 4096)):
        with World(precision="f32") as w:
            w.restore(ck)
            w.set_resolver_split(*split)
            w.sur_set_norm(norm)
            w.sur_set_params(params)
            fb = w.step_sur(sc.dt, 4, 4, tol=0.06, polish=1)
            out[split] = {"fb": np.asarray(fb), **_snapshot(w)}
    _same(out[(0, 0)], out[(1, 4096)], "polish")


def test_partial_split_on_many_worlds_resolves_every_world():
    # 1024 settled worlds, a few dozen of them on the multi-wave kernel:
    # the selection must hand every world to exactly one kernel (a world
    # flagged for the multi-wave kernel but not in its list would not be
    # resolved at all), whatever the split, in the exact and the surrogate
    # step (whose two launches select by different costs)
    sc = scenes.batched(0, 1024)
    with World(sc, precision="f32") as w:
        w.set_resolver_split(0, 0)
        w.step(sc.dt, 240, 4)
        ck = w.checkpoint()
    params, norm = surrogate.load()
    for surrogate_step in (False, True):
        out = {}
        for split in ((0, 0), (257, 32), (257, 100), (600, 7)):
            with World(precision="f32") as w:
                w.restore(ck)
                w.set_resolver_split(*split)
                if surrogate_step:
                    w.sur_set_norm(norm)
                    w.sur_set_params(params)
                    fb, counts = w.step_sur(sc.dt, 4, 4, tol=0.06, polish=1, counts=True)
                    out[split] = {"fb": np.asarray(fb), **_snapshot(w)}
                else:
                    counts = w.step(sc.dt, 4, 4, counts=True)
                    out[split] = _snapshot(w)
                # every world with penetrating contacts ran its position phase
                it = w.resolver_iterations()
                pen = w.contacts(-1)["pen"]
                assert len(pen) and it[:, 0].sum() > 0
        for split in ((257, 32), (257, 100), (600, 7)):
            _same(out[(0, 0)], out[split], f"split {split}, surrogate {surrogate_step}")


def test_multiwave_position_polish_rejections_equal_one_wave():
    """The position polish in both kernels, with predictions shifted up so
    that many worlds over-separate a contact and are rejected: the same bits
    through either kernel, the same worlds rejected."""
    params, norm = surrogate.load()
    norm = norm.copy()
    norm[80 + 12] += 0.02  # out_mean of the resolved penetration
    sc, ck = _settled("f32")
    out = {}
    for split in (0, 1):
        with World(precision="f32") as w:
            w.restore(ck)
            w.set_resolver_split(split)
            w.sur_set_norm(norm)
            w.sur_set_params(params)
            w.sur_set_position(0.0, 4)
            fb = w.step_sur(sc.dt, 3, 4, tol=0.06, polish=4)
            out[split] = {"fb": np.asarray(fb), "pmask": w.sur_position_mask(), **_snapshot(w)}
    _same(out[0], out[1], "position rejections")
    print("worlds given the plain position phase:", int(out[0]["pmask"].sum()), "of", len(out[0]["pmask"]))
