"""ctypes wrapper of the CPU FP64 oracle (oracle/libcbz_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.

OracleWorld holds one world of a cubez_amd Scene in the oracle's
body-major arrays and runs the reference's step on it (cbz_oracle.c cites
the Go file:line of every function).
"""
import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CBZ_ORACLE_LIB points the tests at another build of the oracle (the
# planted-mistake builds of tests/test_oracle_mutants.py)
LIB_PATH = os.environ.get("CBZ_ORACLE_LIB") or os.path.join(HERE, "libcbz_oracle.so")

CONTACT_DTYPE = np.dtype([
    ("body0", np.int32), ("body1", np.int32), ("friction", np.float64), ("restitution", np.float64),
    ("point", np.float64, 3), ("normal", np.float64, 3), ("pen", np.float64), ("c2w", np.float64, 9),
    ("rel", np.float64, 6), ("cvel", np.float64, 3), ("ddv", np.float64),
], align=True)

# cbzo_resolve_ex option: the frictionless impulse of a two-body contact
# includes its second body (the reference omits it, contact.go:512)
FRICTIONLESS_TWO_BODY = 1

COUNTER_FIELDS = ("pairs", "plane_checks", "sat_axes", "sat_skipped", "contacts", "pos_iters",
                  "vel_iters", "integrated", "panics")


class Scene(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("np", ctypes.c_int32)] + [
        (k, ctypes.c_void_p) for k in (
            "shape", "half", "pos", "quat", "vel", "rot", "acc", "inv_mass", "iit_body", "iit_world",
            "transform", "ctransform", "offset", "last_acc", "motion", "pow_lin", "pow_ang", "flags",
            "plane_normal", "plane_offset")]


class Counters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in COUNTER_FIELDS]

    def as_dict(self):
        return {k: getattr(self, k) for k in COUNTER_FIELDS}


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sp = ctypes.POINTER(Scene)
        cp = ctypes.POINTER(Counters)
        L.cbzo_calculate_derived.argtypes = [sp]
        L.cbzo_integrate.argtypes = [sp, f64, f64, cp]
        L.cbzo_narrowphase.argtypes = [sp, P, i64, cp]
        L.cbzo_narrowphase.restype = i64
        L.cbzo_narrowphase_pairs.argtypes = [sp, i64, P, P, P, i64, cp]
        L.cbzo_narrowphase_pairs.restype = i64
        L.cbzo_resolve.argtypes = [sp, P, i64, i64, f64, cp]
        L.cbzo_resolve_ex.argtypes = [sp, P, i64, i64, f64, i32, cp]
        L.cbzo_step.argtypes = [sp, f64, f64, i32, i32, P, cp]
        L.cbzo_step.restype = ctypes.c_int
        L.cbzo_real_equal.argtypes = [f64, f64]
        L.cbzo_real_equal.restype = ctypes.c_int
        for name in ("cbzo_vec_dot", "cbzo_vec_magnitude", "cbzo_quat_len", "cbzo_mat3_det"):
            getattr(L, name).restype = f64
        L.cbzo_vec_dot.argtypes = [P, P]
        L.cbzo_vec_magnitude.argtypes = [P]
        L.cbzo_quat_len.argtypes = [P]
        L.cbzo_mat3_det.argtypes = [P]
        L.cbzo_quat_add_scaled_vector.argtypes = [P, P, f64]
        L.cbzo_quat_from_axis.argtypes = [f64, f64, f64, f64, P]
        L.cbzo_mat3_block_inertia.argtypes = [P, f64, P]
        L.cbzo_contact_point.argtypes = [P, P, f64, P, P, f64, ctypes.c_int, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def pow_inputs(lin_damp, ang_damp, dt):
    """The three math.Pow values of Integrate (rigidbody.go:233, 234, 250),
    from the C library's pow (the product evaluates them the same way)."""
    pl = np.array([math.pow(x, dt) for x in lin_damp], dtype=np.float64)
    pa = np.array([math.pow(x, dt) for x in ang_damp], dtype=np.float64)
    return pl, pa, math.pow(0.5, dt)


class OracleWorld:
    """One world of a Scene, held in the oracle's layout."""

    def __init__(self, scene, world=0):
        self.L = load()
        sl = scene.world_slice(world)
        f = scene.fields
        n = sl.stop - sl.start
        p0, p1 = int(scene.plane_off[world]), int(scene.plane_off[world + 1])
        self.n, self.np_ = n, p1 - p0
        self.body0 = sl.start
        g = lambda k, w: np.ascontiguousarray(f[k][sl], dtype=np.float64).reshape(n * w).copy()
        self.a = {
            "shape": np.ascontiguousarray(f["shape"][sl], np.int32).copy(),
            "half": g("half", 3), "pos": g("pos", 3), "quat": g("quat", 4), "vel": g("vel", 3),
            "rot": g("rot", 3), "acc": g("acc", 3), "inv_mass": g("inv_mass", 1),
            "iit_body": g("iit_body", 9), "offset": g("offset", 12),
            "motion": g("motion", 1),
            "iit_world": np.tile(np.array([1.0, 0, 0, 0, 1, 0, 0, 0, 1]), n),
            "transform": np.zeros(12 * n), "ctransform": np.zeros(12 * n), "last_acc": np.zeros(3 * n),
            "flags": np.ascontiguousarray(f["flags"][sl], np.uint8).copy(),
            "pow_lin": np.ones(n), "pow_ang": np.ones(n),
            "plane_normal": np.ascontiguousarray(scene.plane_normal[p0:p1], np.float64).reshape(-1).copy(),
            "plane_offset": np.ascontiguousarray(scene.plane_offset[p0:p1], np.float64).copy(),
        }
        for k in ("iit_world", "transform", "ctransform", "last_acc"):
            if k in f:
                self.a[k] = g(k, {"iit_world": 9, "transform": 12, "ctransform": 12, "last_acc": 3}[k])
        self.lin_damp = np.asarray(f["lin_damp"][sl], np.float64)
        self.ang_damp = np.asarray(f["ang_damp"][sl], np.float64)
        self.s = Scene(n, self.np_, *[_p(self.a[k]) for k in (
            "shape", "half", "pos", "quat", "vel", "rot", "acc", "inv_mass", "iit_body", "iit_world",
            "transform", "ctransform", "offset", "last_acc", "motion", "pow_lin", "pow_ang", "flags",
            "plane_normal", "plane_offset")])
        self.cnt = Counters()
        self._dt = None
        self.bias = 1.0

    def _pow(self, dt):
        if dt != self._dt:
            pl, pa, self.bias = pow_inputs(self.lin_damp, self.ang_damp, dt)
            self.a["pow_lin"][:] = pl
            self.a["pow_ang"][:] = pa
            self._dt = dt

    def calculate_derived(self):
        self.L.cbzo_calculate_derived(ctypes.byref(self.s))

    def integrate(self, dt):
        self._pow(dt)
        self.L.cbzo_integrate(ctypes.byref(self.s), dt, self.bias, ctypes.byref(self.cnt))

    def narrowphase(self, cap=None):
        cap = cap or 64 + 16 * self.n * (self.np_ + 1)
        out = np.zeros(cap, CONTACT_DTYPE)
        m = self.L.cbzo_narrowphase(ctypes.byref(self.s), _p(out), cap, ctypes.byref(self.cnt))
        if m > cap:
            return self.narrowphase(cap=m)
        return out[:m].copy()

    def narrowphase_pairs(self, one, two, cap=None):
        one = np.ascontiguousarray(one, np.int32)
        two = np.ascontiguousarray(two, np.int32)
        cap = cap or 64 + 8 * len(one)
        out = np.zeros(cap, CONTACT_DTYPE)
        m = self.L.cbzo_narrowphase_pairs(ctypes.byref(self.s), len(one), _p(one), _p(two), _p(out), cap,
                                          ctypes.byref(self.cnt))
        if m > cap:
            return self.narrowphase_pairs(one, two, cap=m)
        return out[:m].copy()

    def resolve(self, contacts, dt, max_iter, flags=0):
        """ResolveContacts on the given records (their friction and
        restitution included); flags: FRICTIONLESS_TWO_BODY."""
        c = np.ascontiguousarray(contacts, CONTACT_DTYPE).copy()
        self.L.cbzo_resolve_ex(ctypes.byref(self.s), _p(c), len(c), int(max_iter), dt, int(flags),
                               ctypes.byref(self.cnt))
        return c

    def step(self, dt, n_steps, iter_mult):
        self._pow(dt)
        counts = np.zeros(n_steps, np.int64)
        rc = self.L.cbzo_step(ctypes.byref(self.s), dt, self.bias, int(iter_mult), int(n_steps), _p(counts),
                              ctypes.byref(self.cnt))
        if rc != 0:
            raise RuntimeError("oracle step failed")
        return counts

    def get(self, name):
        w = {"pos": 3, "quat": 4, "vel": 3, "rot": 3, "iit_world": 9, "transform": 12, "ctransform": 12,
             "last_acc": 3, "motion": 1, "flags": 1}[name]
        a = self.a[name]
        return a.reshape(self.n, w).copy() if w > 1 else a.copy()

    def state(self, names=("pos", "quat", "vel", "rot", "iit_world", "transform", "ctransform",
                           "last_acc", "motion", "flags")):
        return {k: self.get(k) for k in names}

    def counters(self):
        return self.cnt.as_dict()
