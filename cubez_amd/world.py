"""Batched worlds on one GPU, over the C ABI (include/cubez_amd.h).

``World`` holds many independent scenes (worlds) in HBM and runs the
reference's step on all of them: Integrate + CalculateDerivedData
(rigidbody.go:213-272), the ordered-pair narrowphase (colliders.go:720) and
ResolveContacts (contact.go:208), each phase one launch over every world.
"""
import ctypes

import numpy as np

from . import _native as N


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class World:
    def __init__(self, scene=None, precision="f64", device=0, contact_capacity=0):
        self._lib = N.load()
        cfg = N.Config(N.CBZ_F64 if precision == "f64" else N.CBZ_F32, device, int(contact_capacity))
        h = ctypes.c_void_p()
        rc = self._lib.cbz_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise N.CubezError(rc, "cbz_create failed (no HIP device, or bad configuration)")
        self._h = h
        self.precision = precision
        self.scene = None
        if scene is not None:
            self.load(scene)

    # -- plumbing -----------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            msg = self._lib.cbz_last_error(self._h)
            raise N.CubezError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cbz_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- scene ---------------------------------------------------------------
    def load(self, scene, derive=True):
        """Upload a Scene and (by default) compute the derived data of every
        body, as the reference's set-up calls CalculateDerivedData."""
        bo = np.ascontiguousarray(scene.body_off, np.int32)
        po = np.ascontiguousarray(scene.plane_off, np.int32)
        pn = np.ascontiguousarray(scene.plane_normal, np.float64).reshape(-1)
        pof = np.ascontiguousarray(scene.plane_offset, np.float64)
        self._chk(self._lib.cbz_set_scene(self._h, len(bo) - 1, _ptr(bo), _ptr(po), _ptr(pn), _ptr(pof)))
        self.scene = scene
        self.n_worlds = len(bo) - 1
        self.n_bodies = int(bo[-1])
        for name, arr in scene.fields.items():
            self.set(name, arr)
        if derive:
            self.calculate_derived()

    def set(self, name, arr):
        fid, width, kind = N.FIELDS[name]
        dt = {"f": np.float64, "i": np.int32, "u": np.uint8}[kind]
        a = np.ascontiguousarray(arr, dtype=dt).reshape(self.n_bodies * width)
        self._chk(self._lib.cbz_set_field(self._h, fid, _ptr(a)))

    def get(self, name):
        fid, width, kind = N.FIELDS[name]
        dt = {"f": np.float64, "i": np.int32, "u": np.uint8}[kind]
        a = np.empty(self.n_bodies * width, dtype=dt)
        self._chk(self._lib.cbz_get_field(self._h, fid, _ptr(a)))
        return a.reshape(self.n_bodies, width) if width > 1 else a

    def state(self, names=("pos", "quat", "vel", "rot", "iit_world", "transform", "ctransform",
                           "last_acc", "motion", "flags")):
        return {k: self.get(k) for k in names}

    # -- phases -----------------------------------------------------------------
    def calculate_derived(self):
        self._chk(self._lib.cbz_calculate_derived(self._h))

    def integrate(self, dt):
        self._chk(self._lib.cbz_integrate(self._h, float(dt)))

    def narrowphase(self):
        t = ctypes.c_int64(0)
        self._chk(self._lib.cbz_narrowphase(self._h, ctypes.byref(t)))
        return t.value

    def narrowphase_broad(self):
        """The narrowphase behind the sort-and-sweep broadphase: (contacts,
        pairs checked); the contacts equal narrowphase()'s."""
        t = ctypes.c_int64(0)
        n = ctypes.c_int64(0)
        self._chk(self._lib.cbz_narrowphase_broad(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def pair_contact_counts(self, n_pairs):
        """Contacts each pair of the last narrowphase_pairs call appended
        (its CheckForCollisions bool is count > 0)."""
        out = np.empty(n_pairs, np.int32)
        self._chk(self._lib.cbz_pair_contact_counts(self._h, int(n_pairs), _ptr(out)))
        return out

    def narrowphase_pairs(self, one, two):
        one = np.ascontiguousarray(one, np.int32)
        two = np.ascontiguousarray(two, np.int32)
        t = ctypes.c_int64(0)
        self._chk(self._lib.cbz_narrowphase_pairs(self._h, len(one), _ptr(one), _ptr(two), ctypes.byref(t)))
        return t.value

    def contact_counts(self):
        c = np.empty(self.n_worlds, np.int64)
        self._chk(self._lib.cbz_contact_counts(self._h, _ptr(c)))
        return c

    def contacts(self, world=0, internals=True):
        """Contact records of one world, or of every world (world=-1)."""
        counts = self.contact_counts()
        n = int(counts.sum()) if world < 0 else int(counts[world])
        out = {
            "body0": np.empty(n, np.int32), "body1": np.empty(n, np.int32),
            "point": np.empty((n, 3)), "normal": np.empty((n, 3)), "pen": np.empty(n),
            "friction": np.empty(n), "restitution": np.empty(n),
        }
        if internals:
            out.update({"c2w": np.empty((n, 9)), "rel": np.empty((n, 6)), "cvel": np.empty((n, 3)),
                        "ddv": np.empty(n)})
        v = N.ContactView(*[out[k].ctypes.data if k in out else None for k in
                            ("body0", "body1", "point", "normal", "pen", "c2w", "rel", "cvel", "ddv",
                             "friction", "restitution")])
        self._chk(self._lib.cbz_get_contacts(self._h, int(world), ctypes.byref(v)))
        return out

    def set_contacts(self, counts, body0, body1, point, normal, pen):
        counts = np.ascontiguousarray(counts, np.int64)
        arrs = [np.ascontiguousarray(body0, np.int32), np.ascontiguousarray(body1, np.int32),
                np.ascontiguousarray(point, np.float64).reshape(-1),
                np.ascontiguousarray(normal, np.float64).reshape(-1), np.ascontiguousarray(pen, np.float64)]
        self._chk(self._lib.cbz_set_contacts(self._h, _ptr(counts), *[_ptr(a) for a in arrs]))

    def set_materials(self, on=True, frictionless_two_body=False):
        """Per-body, per-plane and per-contact friction and restitution
        (cbz_set_materials); off, every contact carries 0.9 and 0.1 as in the
        reference.  Drops the current contacts."""
        flags = (N.MATERIALS if on else 0) | (N.MATERIALS_FRICTIONLESS_TWO_BODY if frictionless_two_body else 0)
        self._chk(self._lib.cbz_set_materials(self._h, flags))

    def set_plane_materials(self, friction=None, restitution=None):
        f = None if friction is None else np.ascontiguousarray(friction, np.float64)
        e = None if restitution is None else np.ascontiguousarray(restitution, np.float64)
        self._chk(self._lib.cbz_set_plane_materials(self._h, None if f is None else _ptr(f),
                                                    None if e is None else _ptr(e)))

    def set_contact_materials(self, friction=None, restitution=None):
        """Friction / restitution of every current contact (all worlds
        concatenated in emission order)."""
        f = None if friction is None else np.ascontiguousarray(friction, np.float64)
        e = None if restitution is None else np.ascontiguousarray(restitution, np.float64)
        self._chk(self._lib.cbz_set_contact_materials(self._h, None if f is None else _ptr(f),
                                                      None if e is None else _ptr(e)))

    # -- checkpoints ------------------------------------------------------------
    def checkpoint(self):
        """The whole state as bytes (cbz_checkpoint_save)."""
        n = ctypes.c_int64(0)
        self._chk(self._lib.cbz_checkpoint_size(self._h, ctypes.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        self._chk(self._lib.cbz_checkpoint_save(self._h, _ptr(buf), n.value))
        return buf.tobytes()

    def restore(self, data):
        """Replace the state with a checkpoint (cbz_checkpoint_load)."""
        buf = np.frombuffer(bytes(data), np.uint8).copy()
        self._chk(self._lib.cbz_checkpoint_load(self._h, _ptr(buf), len(buf)))
        self.scene = None
        self.n_worlds = int(self._lib.cbz_num_worlds(self._h))
        self.n_bodies = int(self._lib.cbz_num_bodies(self._h))

    def resolve(self, dt, iter_per_contact=4, iter_fixed=0):
        self._chk(self._lib.cbz_resolve(self._h, float(dt), int(iter_per_contact), int(iter_fixed)))

    def resolver_occupancy(self):
        """(worlds resident per CU for the plain resolver, CU count)."""
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        self._chk(self._lib.cbz_resolver_occupancy(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def set_resolver_split(self, min_contacts, max_worlds=32):
        """Of the worlds of at least min_contacts contacts (at most 2048, and
        256 bodies), the max_worlds largest resolve on the kernel of four
        wavefronts per world; either 0: none (cbz_set_resolver_split; the
        same results either way)."""
        self._chk(self._lib.cbz_set_resolver_split(self._h, int(min_contacts), int(max_worlds)))

    def resolver_split(self):
        a, b = ctypes.c_int32(0), ctypes.c_int32(0)
        self._chk(self._lib.cbz_get_resolver_split(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune_resolver(self, dt, iter_per_contact=4, steps=5, candidates=((0, 0), (257, 16), (257, 32),
                                                                           (257, 64), (257, 128)),
                          surrogate=False, tol=0.06, polish=4):
        """Times `steps` steps from the current state for each resolver split
        (min_contacts, max_worlds), restoring the state around each run,
        keeps the fastest (cbz_autotune_resolver) and returns
        {"split": (min, max), "ms_per_step": {split: ms}}."""
        cand = np.asarray(candidates, np.int32).reshape(-1, 2)
        lo, hi = np.ascontiguousarray(cand[:, 0]), np.ascontiguousarray(cand[:, 1])
        ms = np.zeros(len(cand), np.float64)
        best = ctypes.c_int32(-1)
        self._chk(self._lib.cbz_autotune_resolver(self._h, float(dt), int(iter_per_contact), int(steps),
                                                  1 if surrogate else 0, float(tol), int(polish), len(cand),
                                                  _ptr(lo), _ptr(hi), _ptr(ms), ctypes.byref(best)))
        splits = [tuple(int(x) for x in r) for r in cand]
        return {"split": splits[best.value], "ms_per_step": {f"{a},{b}": float(m) for (a, b), m in zip(splits, ms)}}

    def resolver_world_times(self):
        """[n_worlds, 4] microseconds: start and end of each world's
        workgroup in the last position-only and velocity launches (-1:
        none)."""
        t = np.empty((self.n_worlds, 4))
        self._chk(self._lib.cbz_resolver_world_times(self._h, _ptr(t)))
        return t

    def resolver_iterations(self):
        it = np.empty(2 * self.n_worlds, np.int64)
        self._chk(self._lib.cbz_resolver_iterations(self._h, _ptr(it)))
        return it.reshape(self.n_worlds, 2)

    def step(self, dt, n_steps=1, iter_per_contact=4, counts=False):
        """n_steps fused steps; with counts=True returns [n_steps, n_worlds]
        contact counts."""
        buf = np.empty((n_steps, self.n_worlds), np.int64) if counts else None
        self._chk(self._lib.cbz_step(self._h, float(dt), int(n_steps), int(iter_per_contact), _ptr(buf)))
        return buf

    def integrate_timed(self, dt, n=50):
        """n back-to-back integrate launches timed with HIP events on the
        library's stream: (ms per launch, bodies integrated over the n)."""
        ms = np.zeros(1)
        cnt = np.zeros(1, np.int64)
        self._chk(self._lib.cbz_integrate_timed(self._h, float(dt), int(n), _ptr(ms), _ptr(cnt)))
        return float(ms[0]), int(cnt[0])

    def synchronize(self):
        self._chk(self._lib.cbz_synchronize(self._h))

    def set_profiling(self, on=True):
        self._chk(self._lib.cbz_set_profiling(self._h, 1 if on else 0))

    def phase_times(self):
        """Accumulated ms [integrate, np count, np scans, np emit, resolve,
        whole step] and counters [steps, bodies integrated, contacts, resolver
        iterations] of the profiled steps since the last call."""
        ms = np.zeros(6)
        cnt = np.zeros(4, np.int64)
        self._chk(self._lib.cbz_phase_times(self._h, _ptr(ms), _ptr(cnt)))
        return ms, cnt

    def selftest_f64(self, op, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.empty_like(a)
        self._chk(self._lib.cbz_selftest_f64(self._h, int(op), len(a), _ptr(a), _ptr(b), _ptr(out)))
        return out

    # -- learned resolver surrogate ---------------------------------------------
    N_FEAT, N_OUT = 40, 13  # outputs: 12 velocity-phase changes, 1 resolved penetration
    N_NORM = 2 * N_FEAT + 2 * N_OUT

    def selftest_sort(self, keys, vals=None):
        """The broadphase's device radix sort on host arrays
        (cbz_selftest_sort): (sorted keys, values in the same permutation)."""
        keys = np.ascontiguousarray(keys, np.uint64)
        ko = np.empty_like(keys)
        vo = None
        if vals is not None:
            vals = np.ascontiguousarray(vals, np.int32)
            vo = np.empty_like(vals)
        self._chk(self._lib.cbz_selftest_sort(self._h, len(keys), _ptr(keys), None if vals is None else _ptr(vals),
                                              _ptr(ko), None if vo is None else _ptr(vo)))
        return ko, vo

    def sur_num_params(self):
        return int(self._lib.cbz_sur_num_params())

    def sur_set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.sur_num_params()
        self._chk(self._lib.cbz_sur_set_params(self._h, _ptr(p)))

    def sur_get_params(self):
        p = np.empty(self.sur_num_params(), np.float32)
        self._chk(self._lib.cbz_sur_get_params(self._h, _ptr(p)))
        return p

    def sur_set_norm(self, n):
        n = np.ascontiguousarray(n, np.float32)
        assert n.size == self.N_NORM
        self._chk(self._lib.cbz_sur_set_norm(self._h, _ptr(n)))

    def sur_get_norm(self):
        n = np.empty(self.N_NORM, np.float32)
        self._chk(self._lib.cbz_sur_get_norm(self._h, _ptr(n)))
        return n

    def sur_reserve(self, rows):
        self._chk(self._lib.cbz_sur_reserve(self._h, int(rows)))

    def sur_collect(self, dt, iter_per_contact=4):
        n = ctypes.c_int64(0)
        self._chk(self._lib.cbz_sur_collect(self._h, float(dt), int(iter_per_contact), ctypes.byref(n)))
        return n.value

    def sur_dataset(self, download=True):
        n = ctypes.c_int64(0)
        self._chk(self._lib.cbz_sur_dataset(self._h, ctypes.byref(n), None, None))
        if not download:
            return n.value
        X = np.empty((n.value, self.N_FEAT), np.float32)
        Y = np.empty((n.value, self.N_OUT), np.float32)
        self._chk(self._lib.cbz_sur_dataset(self._h, ctypes.byref(n), _ptr(X), _ptr(Y)))
        return X, Y

    def sur_load_dataset(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        self._chk(self._lib.cbz_sur_load_dataset(self._h, len(X), _ptr(X), _ptr(Y)))

    def sur_init(self, seed=0):
        self._chk(self._lib.cbz_sur_init(self._h, int(seed)))

    def sur_train(self, steps, batch=4096, lr=1e-3, seed=0):
        losses = np.empty(max(int(steps), 1), np.float32)
        self._chk(self._lib.cbz_sur_train(self._h, int(steps), int(batch), float(lr), int(seed), _ptr(losses)))
        return losses[:steps]

    def sur_grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        g = np.empty(self.sur_num_params(), np.float32)
        loss = np.zeros(1, np.float32)
        self._chk(self._lib.cbz_sur_grad(self._h, len(X), _ptr(X), _ptr(Y), _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def sur_predict(self, X, layered=False):
        """Predictions for raw feature rows: the fused inference kernel, or
        (layered=True) the per-layer kernels training uses."""
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty((len(X), self.N_OUT), np.float32)
        self._chk(self._lib.cbz_sur_predict(self._h, len(X), _ptr(X), _ptr(out), 1 if layered else 0))
        return out

    SUR_ZERO_PREDICTION = 1

    def sur_resolve(self, dt, iter_per_contact=4, tol=0.06, polish=4, zero_prediction=False):
        """ResolveContacts with the surrogate: per phase, the predicted
        changes polished by the worst-first loop (positions to
        positionEpsilon + the position tolerance, see sur_set_position;
        velocities to velocityEpsilon + tol within polish * C iterations),
        else that plain phase.  Returns the number of worlds that fell back
        in either phase."""
        n = ctypes.c_int64(0)
        flags = self.SUR_ZERO_PREDICTION if zero_prediction else 0
        self._chk(self._lib.cbz_sur_resolve(self._h, float(dt), int(iter_per_contact), float(tol), int(polish), flags,
                                            ctypes.byref(n)))
        return n.value

    def sur_phase_times(self):
        """Accumulated ms [prepare, features, network, per-body sums,
        position phase, velocity phase] and counts [resolves, contacts,
        fallbacks (either phase), velocity iterations, position fallbacks,
        position iterations] of profiled surrogate resolves since the last
        call."""
        ms = np.zeros(6)
        cnt = np.zeros(6, np.int64)
        self._chk(self._lib.cbz_sur_phase_times(self._h, _ptr(ms), _ptr(cnt)))
        return ms, cnt

    def sur_world_mask(self):
        """Per world, 1 where the last surrogate resolve's velocity phase was
        the plain one."""
        m = np.empty(self.n_worlds, np.uint8)
        self._chk(self._lib.cbz_sur_world_mask(self._h, _ptr(m)))
        return m

    def sur_position_mask(self):
        """Per world, 1 where the last surrogate resolve's position phase was
        the plain one."""
        m = np.empty(self.n_worlds, np.uint8)
        self._chk(self._lib.cbz_sur_position_mask(self._h, _ptr(m)))
        return m

    def sur_set_position(self, tol=0.0, polish=4):
        """The position phase's polish: tolerance in m above positionEpsilon
        (negative: the plain position phase everywhere) and budget per
        contact."""
        self._chk(self._lib.cbz_sur_set_position(self._h, float(tol), int(polish)))

    def sur_position(self):
        t, p = ctypes.c_double(0), ctypes.c_int64(0)
        self._chk(self._lib.cbz_sur_get_position(self._h, ctypes.byref(t), ctypes.byref(p)))
        return t.value, p.value

    def step_sur(self, dt, n_steps=1, iter_per_contact=4, tol=0.06, polish=4, zero_prediction=False,
                 counts=False):
        """n_steps with the surrogate resolver; returns the per-step fallback
        counts (and [n_steps, n_worlds] contact counts with counts=True)."""
        buf = np.empty((n_steps, self.n_worlds), np.int64) if counts else None
        fb = np.empty(max(n_steps, 1), np.int64)
        flags = self.SUR_ZERO_PREDICTION if zero_prediction else 0
        self._chk(self._lib.cbz_step_sur(self._h, float(dt), int(n_steps), int(iter_per_contact), float(tol),
                                         int(polish), flags, _ptr(buf), _ptr(fb)))
        return (fb[:n_steps], buf) if counts else fb[:n_steps]
