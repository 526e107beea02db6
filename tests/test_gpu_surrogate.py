"""The learned resolver surrogate on the GPU: the MFMA forward and backward
kernels against an FP64 numpy MLP, the data logger against the exact
resolver, the fallback contract, and what training gains."""
import numpy as np
import pytest

from cubez_amd import World, scenes, surrogate

pytestmark = pytest.mark.gpu

NO = surrogate.N_OUT  # outputs: 12 velocity-phase changes, 1 resolved penetration


def out_norm(mean=0.0, std=1.0):
    """A normalisation vector: identity inputs, the given output mean/std."""
    return np.concatenate([np.zeros(40), np.ones(40), np.broadcast_to(mean, NO),
                           np.broadcast_to(std, NO)]).astype(np.float32)


def bf16(x):
    """Round float32 to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def ref_forward(params, norm, X):
    """FP64 forward with the GPU's roundings: BF16 inputs, weights and hidden
    activations, FP32-exact accumulation modelled in FP64."""
    mi, si, mo, so = norm[:40], norm[40:80], norm[80:80 + NO], norm[80 + NO:80 + 2 * NO]
    h = bf16((X - mi) / si).astype(np.float64)
    acts, zs = [h], []
    layers = surrogate.unpack(params)
    for l, (W, b) in enumerate(layers):
        z = h @ bf16(W).astype(np.float64).T + b
        zs.append(z)
        if l < len(layers) - 1:
            h = bf16(silu(z).astype(np.float32)).astype(np.float64)
            acts.append(h)
        else:
            h = z
    return h, h * so + mo, acts, zs


def ref_grad(params, norm, X, Y):
    """FP64 backprop of mean((y - t)^2) with normalised targets."""
    mo, so = norm[80:80 + NO], norm[80 + NO:80 + 2 * NO]
    yhat, _, acts, zs = ref_forward(params, norm, X)
    t = (Y - mo) / so
    n = yhat.size
    loss = np.sum((yhat - t) ** 2) / n
    dz = 2.0 * (yhat - t) / n
    layers = surrogate.unpack(params)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads[l] = (dz.T @ acts[l], dz.sum(axis=0))
        if l > 0:
            da = dz @ bf16(W).astype(np.float64)
            s = 1.0 / (1.0 + np.exp(-zs[l - 1]))
            dz = da * s * (1.0 + zs[l - 1] * (1.0 - s))
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return flat, loss


@pytest.fixture(scope="module")
def model():
    rng = np.random.default_rng(0)
    w = World(scenes.batched(0, 2))
    n = w.sur_num_params()
    params = np.zeros(n, np.float32)
    off = 0
    for n_in, n_out in surrogate.LAYERS:
        params[off:off + n_in * n_out] = rng.uniform(-1, 1, n_in * n_out) * np.sqrt(6.0 / n_in)
        off += n_in * n_out
        params[off:off + n_out] = rng.uniform(-0.1, 0.1, n_out)
        off += n_out
    norm = np.concatenate([rng.normal(0, 1, 40), rng.uniform(0.5, 2, 40), rng.normal(0, 1, NO),
                           rng.uniform(0.5, 2, NO)]).astype(np.float32)
    w.sur_set_norm(norm)
    w.sur_set_params(params)
    yield w, params, norm
    w.close()


def test_forward_matches_fp64_reference(model):
    w, params, norm = model
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1.5, (1000, 40)).astype(np.float32)
    got = w.sur_predict(X)
    _, want, _, _ = ref_forward(params, norm, X.astype(np.float64))
    err = np.abs(got - want) / (np.abs(want).max(axis=0) + 1e-6)
    assert err.max() < 2e-2, err.max()
    assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.9999


def test_fused_forward_equals_layered_forward(model):
    """The fused inference kernel (all layers in one launch, activations in
    LDS) and the per-layer training kernels give the same bits, including a
    ragged last block of rows."""
    w, params, norm = model
    rng = np.random.default_rng(3)
    for rows in (1, 63, 64, 65, 1000, 4099):
        X = rng.normal(0, 1.5, (rows, 40)).astype(np.float32)
        assert np.array_equal(w.sur_predict(X), w.sur_predict(X, layered=True)), rows


def test_gradient_matches_fp64_reference(model):
    w, params, norm = model
    rng = np.random.default_rng(2)
    X = rng.normal(0, 1.0, (700, 40)).astype(np.float32)
    Y = rng.normal(0, 1.0, (700, NO)).astype(np.float32)
    g, loss = w.sur_grad(X, Y)
    gref, lref = ref_grad(params.astype(np.float64), norm.astype(np.float64), X.astype(np.float64),
                          Y.astype(np.float64))
    assert abs(loss - lref) / lref < 1e-2
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 3e-2, rel
    # per layer as well (a transposed operand would fail here)
    off = 0
    for n_in, n_out in surrogate.LAYERS:
        k = n_in * n_out + n_out
        a, b = g[off:off + k], gref[off:off + k]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 5e-2
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.998
        off += k


def _settled(seed=3, worlds=12, steps=240):
    sc = scenes.batched(seed, worlds, 8, 120)
    return sc, steps


def test_collect_is_the_exact_resolve_and_labels_sum_to_body_changes():
    sc, steps = _settled()
    with World(sc) as a, World(sc) as b:
        a.step(sc.dt, steps, 4)
        b.step(sc.dt, steps, 4)
        a.integrate(sc.dt)
        b.integrate(sc.dt)
        a.narrowphase()
        b.narrowphase()
        v_before = a.get("vel").copy()
        w_before = a.get("rot").copy()
        a.sur_reserve(200_000)
        n = a.sur_collect(sc.dt, 4)
        b.resolve(sc.dt, 4)
        for f in ("pos", "quat", "vel", "rot", "flags"):
            assert np.array_equal(a.get(f), b.get(f)), f
        X, Y = a.sur_dataset()
        assert len(X) == n == int(a.contact_counts().sum())
        dv = np.zeros_like(v_before)
        dw = np.zeros_like(w_before)
        off = 0
        for wi in range(sc.n_worlds):
            c = a.contacts(wi, internals=False)
            m = len(c["pen"])
            for j in range(m):
                y = Y[off + j].astype(np.float64)
                dv[c["body0"][j]] += y[0:3]
                dw[c["body0"][j]] += y[3:6]
                if c["body1"][j] >= 0:
                    dv[c["body1"][j]] += y[6:9]
                    dw[c["body1"][j]] += y[9:12]
            off += m
        assert np.allclose(dv, a.get("vel") - v_before, atol=1e-5, rtol=1e-4)
        assert np.allclose(dw, a.get("rot") - w_before, atol=1e-5, rtol=1e-4)
        # the position label: the penetration resolved at the contact, every
        # selection's depth (each above positionEpsilon = 0.01); so a world's
        # labelled contacts are at most its position iterations, and its
        # label sum exceeds 0.01 per iteration
        pos = Y[:, 12].astype(np.float64)
        assert (pos >= 0).all()
        assert (pos[pos > 0] >= np.float32(0.01)).all()
        it = a.resolver_iterations()[:, 0]
        counts = a.contact_counts()
        off = np.concatenate([[0], np.cumsum(counts)])
        for wi in range(sc.n_worlds):
            lw = pos[off[wi]:off[wi + 1]]
            assert (lw > 0).sum() <= it[wi], wi
            assert lw.sum() >= 0.01 * it[wi] - 1e-6, wi
        assert it.sum() > 0


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_collected_features_are_the_prepared_contacts(precision):
    """The feature rows (one workgroup per world, the per-body contact counts
    in LDS) against the same 40 values computed here from a prepare-only
    resolve of the same state: normal, both lever arms, the bodies' relative
    velocity and rotation, inverse masses, world inverse inertias,
    penetration, desiredDeltaVelocity, the closing speed and 1/contacts per
    body, each rounded to FP32 as the kernel rounds it."""
    sc, steps = _settled()
    with World(sc, precision=precision) as a, World(sc, precision=precision) as b:
        for w in (a, b):
            w.step(sc.dt, steps, 4)
            w.integrate(sc.dt)
            w.narrowphase()
        a.sur_reserve(200_000)
        n = a.sur_collect(sc.dt, 4)
        X, _ = a.sur_dataset()
        b.resolve(sc.dt, 0, 0)  # prepare, no iterations
        c = b.contacts(-1, internals=True)
        assert n == len(c["pen"]) > 0
        f = np.float32
        vel, rot = b.get("vel").astype(f), b.get("rot").astype(f)
        im, iw = b.get("inv_mass").astype(f), b.get("iit_world").astype(f)
        b0, b1 = c["body0"], c["body1"]
        assert (b0 >= 0).all()
        has1 = b1 >= 0
        b1c = np.where(has1, b1, 0)
        h = has1[:, None]
        deg = np.bincount(b0, minlength=len(vel)) + np.bincount(b1[has1], minlength=len(vel))
        want = np.zeros((n, 40), f)
        want[:, 0:3] = c["normal"].astype(f)
        want[:, 3:9] = c["rel"].astype(f)
        want[:, 9:12] = vel[b0] - np.where(h, vel[b1c], f(0))
        want[:, 12:15] = rot[b0] - np.where(h, rot[b1c], f(0))
        want[:, 15] = im[b0]
        want[:, 16] = np.where(has1, im[b1c], f(0))
        want[:, 17:26] = iw[b0]
        want[:, 26:35] = np.where(h, iw[b1c], f(0))
        want[:, 35] = c["pen"].astype(f)
        want[:, 36] = c["ddv"].astype(f)
        want[:, 37] = c["cvel"][:, 0].astype(f)
        want[:, 38] = f(1) / deg[b0].astype(f)
        want[:, 39] = np.where(has1, f(1) / np.maximum(deg[b1c], 1).astype(f), f(0))
        X = np.asarray(X, f)
        assert np.array_equal(X[:, :38], want[:, :38])
        # 1/deg: within an ulp whatever the division's rounding
        assert np.allclose(X[:, 38:], want[:, 38:], rtol=2e-7, atol=0)


def _plain_and_sur(tol, params_fn, polish=1, zero=False, norm=None, seed=5, pos=(-1.0, 4)):
    """Five resolves of the same contacts, plain and with the surrogate
    (position polish off unless `pos` gives its tolerance and budget)."""
    sc, steps = _settled(seed)
    with World(sc) as p, World(sc) as s:
        p.step(sc.dt, steps, 4)
        s.step(sc.dt, steps, 4)
        if norm is None:
            norm = out_norm()
        s.sur_set_position(*pos)
        s.sur_set_norm(norm)
        s.sur_set_params(params_fn(s.sur_num_params()))
        for _ in range(5):
            p.integrate(sc.dt)
            s.integrate(sc.dt)
            p.narrowphase()
            s.narrowphase()
            p.resolve(sc.dt, 4)
            nf = s.sur_resolve(sc.dt, 4, tol, polish=polish, zero_prediction=zero)
        return p.state(), s.state(), nf, s.sur_world_mask(), sc


def test_negative_tolerance_falls_back_everywhere_and_equals_plain():
    ps, ss, nf, mask, sc = _plain_and_sur(-1.0, lambda n: np.zeros(n, np.float32))
    assert mask.all()
    for k in ps:
        assert np.array_equal(ps[k], ss[k]), k


def test_nan_network_falls_back_and_equals_plain():
    # with the position polish on: a NaN prediction sends both phases plain
    ps, ss, nf, mask, sc = _plain_and_sur(1e30, lambda n: np.full(n, np.nan, np.float32), pos=(1e-3, 4))
    assert nf > 0
    for k in ps:
        assert np.array_equal(ps[k], ss[k]), k


def test_infinite_tolerance_accepts_every_world():
    ps, ss, nf, mask, sc = _plain_and_sur(1e30, lambda n: np.zeros(n, np.float32))
    assert nf == 0


def test_prediction_that_pushes_bodies_apart_is_rejected():
    """A prediction that drives contacts apart (5 m/s up for every contact's
    first body) leaves no contact closing, so the closing-speed test alone
    would accept it; the separation bound rejects it.  Rejected worlds are
    the plain resolver's result bit for bit, and no accepted world keeps the
    push."""
    norm = out_norm()
    norm[80 + 1] = 5.0  # out_mean of dv_body0.y
    sc, steps = _settled(11, worlds=16, steps=120)
    rejected = 0
    with World(sc) as p, World(sc) as s:
        p.step(sc.dt, steps, 4)
        s.step(sc.dt, steps, 4)
        s.sur_set_position(-1.0)
        s.sur_set_norm(norm)
        s.sur_set_params(np.zeros(s.sur_num_params(), np.float32))
        for _ in (0, 1, 2):
            p.integrate(sc.dt)
            s.integrate(sc.dt)
            p.narrowphase()
            s.narrowphase()
            p.resolve(sc.dt, 4)
            s.sur_resolve(sc.dt, 4, 0.06, polish=1)
            mask = s.sur_world_mask()
            ps, ss = p.state(), s.state()
            for w in range(sc.n_worlds):
                b0, b1 = sc.body_off[w], sc.body_off[w + 1]
                if mask[w]:
                    rejected += 1
                    for k in ps:
                        assert np.array_equal(ps[k][b0:b1], ss[k][b0:b1]), (w, k)
                else:
                    assert ss["vel"][b0:b1, 1].max() < 1.0, w
            for k in set(ps) | {"last_acc", "motion", "flags"}:
                s.set(k, p.get(k))
    assert rejected > 0


def test_zero_prediction_at_zero_tolerance_is_the_plain_loop():
    """Polishing from zero predictions to positionEpsilon and velocityEpsilon
    with the plain budget is the reference's loops: bit-identical to plain
    (a position polish the separation bound rejects is restored and re-run
    plain, which gives the same bits)."""
    ps, ss, nf, mask, sc = _plain_and_sur(0.0, lambda n: np.zeros(n, np.float32), polish=4, zero=True,
                                          pos=(0.0, 4))
    for k in ps:
        assert np.array_equal(ps[k], ss[k]), k


def test_polished_worlds_meet_tolerance_and_fallbacks_equal_plain():
    """A weak (small random) network: worlds the polish accepts end with every
    desired velocity change within velocityEpsilon + tol; worlds it rejects
    are bit-identical to the plain resolver."""
    rng = np.random.default_rng(7)
    tol = 0.06
    sc, steps = _settled(9, worlds=16)
    with World(sc) as p, World(sc) as s:
        p.step(sc.dt, steps, 4)
        s.step(sc.dt, steps, 4)
        s.sur_set_position(-1.0)
        s.sur_set_norm(out_norm(0.0, 0.05))
        s.sur_set_params((rng.normal(0, 0.05, s.sur_num_params())).astype(np.float32))
        for _ in (0, 1, 2):
            p.integrate(sc.dt)
            s.integrate(sc.dt)
            p.narrowphase()
            s.narrowphase()
            p.resolve(sc.dt, 4)
            s.sur_resolve(sc.dt, 4, tol, polish=1)
            mask = s.sur_world_mask()
            counts = s.contact_counts()
            ps, ss = p.state(), s.state()
            accepted = fell_back = 0
            for w in range(sc.n_worlds):
                b0, b1 = sc.body_off[w], sc.body_off[w + 1]
                if counts[w] == 0:
                    continue
                if mask[w]:
                    fell_back += 1
                    for k in ps:
                        assert np.array_equal(ps[k][b0:b1], ss[k][b0:b1]), (w, k)
                else:
                    accepted += 1
                    ddv = s.contacts(w)["ddv"]
                    assert ddv.max() <= 0.01 + tol + 1e-12, (w, ddv.max())
            print("accepted", accepted, "fell back", fell_back)
            # continue both from the plain state so each step compares like with like
            for k in set(ps) | {"last_acc", "motion", "flags"}:
                s.set(k, p.get(k))


POS_FIELDS = ("pos", "quat", "transform", "iit_world")  # what the position phase writes


def _position_polish_steps(norm, params, seed, ptol=1e-3, steps=3):
    """Per resolve: the plain resolver and the surrogate (velocity tolerance
    infinite, so every world keeps its velocity polish and only the position
    phase decides) from the same state; yields (plain state, surrogate state,
    position mask, contact counts, surrogate world)."""
    sc, n0 = _settled(seed, worlds=16)
    with World(sc) as p, World(sc) as s:
        p.step(sc.dt, n0, 4)
        s.step(sc.dt, n0, 4)
        s.sur_set_position(ptol, 4)
        s.sur_set_norm(norm)
        s.sur_set_params(params)
        for _ in range(steps):
            p.integrate(sc.dt)
            s.integrate(sc.dt)
            p.narrowphase()
            s.narrowphase()
            pen0 = s.contacts(-1, internals=False)["pen"]
            p.resolve(sc.dt, 4)
            s.sur_resolve(sc.dt, 4, 1e30, polish=4)
            yield sc, p.state(), s.state(), s.sur_position_mask(), s.contact_counts(), s, pen0
            for k in set(p.state()) | {"last_acc", "motion", "flags"}:
                s.set(k, p.get(k))


def test_position_polish_meets_tolerance_and_fallbacks_equal_plain():
    """A weak (small random) network's position moves: worlds the polish
    accepts end with every penetration within positionEpsilon + tol and no
    contact separated beyond twice the world's deepest starting penetration + tol;
    worlds it rejects have the plain position phase's positions,
    orientations and derived data bit for bit."""
    rng = np.random.default_rng(17)
    tol = 1e-3
    norm = out_norm(0.0, 0.02)
    accepted = fell_back = 0
    with World(scenes.batched(0, 2)) as tmp:
        params = rng.normal(0, 0.05, tmp.sur_num_params()).astype(np.float32)
    for sc, ps, ss, pmask, counts, s, pen0 in _position_polish_steps(norm, params, seed=9, ptol=tol):
        off = np.concatenate([[0], np.cumsum(counts)])
        pen = s.contacts(-1)["pen"]
        for w in range(sc.n_worlds):
            b0, b1 = sc.body_off[w], sc.body_off[w + 1]
            if counts[w] == 0:
                continue
            if pmask[w]:
                fell_back += 1
                for k in POS_FIELDS:
                    assert np.array_equal(ps[k][b0:b1], ss[k][b0:b1]), (w, k)
            else:
                accepted += 1
                pw = pen[off[w]:off[w + 1]]
                floor = -(2.0 * max(pen0[off[w]:off[w + 1]].max(), 0.0) + tol)
                assert pw.max() <= 0.01 + tol + 1e-12, (w, pw.max())
                assert pw.min() >= floor - 1e-12, (w, pw.min(), floor)
    print("position polish: accepted", accepted, "fell back", fell_back)
    assert accepted > 0


def test_position_prediction_that_pushes_bodies_apart_is_rejected():
    """Predicted resolved penetration of 0.3 m at every contact lifts bodies
    far apart; every penetration is then within tolerance, so only the
    separation bound can see it.  Rejected worlds have the plain position
    phase's result bit for bit, and no accepted world keeps the lift."""
    norm = out_norm()
    norm[80 + 12] = 0.3  # out_mean of the resolved penetration
    rejected = 0
    with World(scenes.batched(0, 2)) as tmp:
        params = np.zeros(tmp.sur_num_params(), np.float32)
    for sc, ps, ss, pmask, counts, s, pen0 in _position_polish_steps(norm, params, seed=11):
        off = np.concatenate([[0], np.cumsum(counts)])
        pen = s.contacts(-1)["pen"]
        for w in range(sc.n_worlds):
            b0, b1 = sc.body_off[w], sc.body_off[w + 1]
            if counts[w] == 0:
                continue
            if pmask[w]:
                rejected += 1
                for k in POS_FIELDS:
                    assert np.array_equal(ps[k][b0:b1], ss[k][b0:b1]), (w, k)
            else:
                pw = pen[off[w]:off[w + 1]]
                assert pw.min() >= -(2.0 * max(pen0[off[w]:off[w + 1]].max(), 0.0) + 1e-3) - 1e-12, w
    assert rejected > 0


def test_training_gains_over_the_method_without_the_network():
    """Train briefly on contacts of 12 seeds; on fresh seeds the trained
    network must leave the exact loop fewer velocity iterations to do than
    the same polish started from a zero prediction, and its loss must fall."""
    with World(scenes.batched(0, 2)) as w:
        surrogate.collect(w, range(12), n_worlds=24, at=(200, 300, 400), reserve=800_000)
        losses = surrogate.train(w, steps=800, batch=2048, lr=1e-3, seed=1)
        trained = w.sur_get_params()
        norm = w.sur_get_norm()
        # per-batch losses are heavy-tailed: compare medians
        assert np.median(losses[-200:]) < 0.8 * np.median(losses[:20])

    def velocity_iterations(zero):
        it = fall = 0
        for seed in (1000, 1001):
            sc = scenes.batched(seed, 24)
            with World(sc) as e:
                e.step(sc.dt, 300, 4)
                e.sur_set_norm(norm)
                e.sur_set_params(trained)
                for _ in range(20):
                    fall += int(e.step_sur(sc.dt, 1, 4, surrogate.DEFAULT_TOL, polish=1, zero_prediction=zero)[0])
                    it += int(e.resolver_iterations()[:, 1].sum())
        return it, fall

    it_net, fb_net = velocity_iterations(False)
    it_zero, fb_zero = velocity_iterations(True)
    print("velocity iterations: network", it_net, "zero prediction", it_zero, "fallbacks", fb_net, fb_zero)
    assert it_net < it_zero


def test_packaged_model_gains_in_both_phases():
    """The packaged model on fresh seeds (settled worlds): the polish from its
    predictions leaves the exact loops fewer position and fewer velocity
    iterations than the same polish from zero predictions."""
    params, norm = surrogate.load()

    def iterations(zero):
        pos = vel = 0
        for seed in (1000, 1001):
            sc = scenes.batched(seed, 24)
            with World(sc, precision="f32") as e:
                e.step(sc.dt, 300, 4)
                e.sur_set_norm(norm)
                e.sur_set_params(params)
                for _ in range(10):
                    e.step_sur(sc.dt, 1, 4, surrogate.DEFAULT_TOL, surrogate.DEFAULT_POLISH, zero_prediction=zero)
                    it = e.resolver_iterations()
                    pos += int(it[:, 0].sum())
                    vel += int(it[:, 1].sum())
        return pos, vel

    pos_net, vel_net = iterations(False)
    pos_zero, vel_zero = iterations(True)
    print("position iterations", pos_net, pos_zero, "velocity iterations", vel_net, vel_zero)
    assert pos_net < pos_zero
    assert vel_net < vel_zero


def test_collection_schedule_of_1024_worlds_trains():
    """The collection that once ended in an illegal memory access at the
    first training GEMM (64 seeds x 16 worlds, snapshots at steps 60, 120,
    200, 300, 450 and 600, about 2 M rows), then training on it.  The fault
    came from k_gemm reading operand rows and bias entries past its buffers
    (layer-0 weight gradient, output-layer bias), which this schedule's
    allocations exposed; the GEMMs now read only inside them."""
    with World(scenes.batched(0, 2)) as w:
        n = surrogate.collect(w, range(64), n_worlds=16, at=(60, 120, 200, 300, 450, 600), reserve=3_200_000)
        assert n > 1_000_000
        losses = surrogate.train(w, steps=200, batch=4096, lr=1e-3, seed=3)
        assert np.all(np.isfinite(losses))
        assert np.all(np.isfinite(w.sur_get_params()))


def test_captured_surrogate_step_equals_eager_steps():
    """cbz_step_sur without counts replays one captured graph per step (no
    host round trip: the feature and network grids read the contact count on
    the device); with counts it launches eagerly.  Replays in chunks of 3, 1
    and 4 steps, with eager work and synchronisations between them, give the
    eager path's state bit for bit and its per-step fallback counts."""
    params, norm = surrogate.load()
    sc = scenes.batched(5, 48)
    with World(sc, precision="f32") as g, World(sc, precision="f32") as e:
        for w in (g, e):
            w.sur_set_norm(norm)
            w.sur_set_params(params)
            w.step(sc.dt, 200, sc.iter_mult)  # into the contact-rich regime
        for chunk in (3, 1, 4):
            fg = g.step_sur(sc.dt, chunk, sc.iter_mult, surrogate.DEFAULT_TOL, surrogate.DEFAULT_POLISH)
            fe, counts = e.step_sur(sc.dt, chunk, sc.iter_mult, surrogate.DEFAULT_TOL, surrogate.DEFAULT_POLISH,
                                    counts=True)
            assert np.array_equal(fg, fe)
            assert np.array_equal(g.contact_counts(), counts[-1])
            for k in ("pos", "quat", "vel", "rot", "motion", "flags"):
                assert np.array_equal(g.get(k), e.get(k)), (chunk, k)
            g.narrowphase()  # eager work on the graph's context between replays
            e.narrowphase()
            g.synchronize()
