"""ctypes binding of libcubez_amd.so (the C ABI declared in include/cubez_amd.h).

The library is built in-tree by ``__graft_entry__.build()``.  Nothing here
falls back to a CPU implementation: if the library is missing, or the HIP
runtime finds no GPU, the calls raise.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUBEZ_AMD_LIB") or os.path.join(_HERE, "libcubez_amd.so")

CBZ_OK = 0
CBZ_ERR_ARG = -1
CBZ_ERR_HIP = -2
CBZ_ERR_CAPACITY = -3
CBZ_ERR_STATE = -4

CBZ_F64 = 0
CBZ_F32 = 1
CBZ_CUBE = 0
CBZ_SPHERE = 1
FLAG_AWAKE = 1
FLAG_CAN_SLEEP = 2
FLAG_STATIC = 4

# field id -> (name, width, element kind: 'f' double, 'i' int32, 'u' uint8)
FIELDS = {
    "shape": (0, 1, "i"),
    "half": (1, 3, "f"),
    "pos": (2, 3, "f"),
    "quat": (3, 4, "f"),
    "vel": (4, 3, "f"),
    "rot": (5, 3, "f"),
    "acc": (6, 3, "f"),
    "inv_mass": (7, 1, "f"),
    "iit_body": (8, 9, "f"),
    "iit_world": (9, 9, "f"),
    "transform": (10, 12, "f"),
    "ctransform": (11, 12, "f"),
    "offset": (12, 12, "f"),
    "last_acc": (13, 3, "f"),
    "motion": (14, 1, "f"),
    "lin_damp": (15, 1, "f"),
    "ang_damp": (16, 1, "f"),
    "flags": (17, 1, "u"),
    "friction": (18, 1, "f"),
    "restitution": (19, 1, "f"),
}

# cbz_set_materials flags
MATERIALS = 1
MATERIALS_FRICTIONLESS_TWO_BODY = 2

# every symbol include/cubez_amd.h declares
EXPORTS = (
    "cbz_create", "cbz_destroy", "cbz_last_error", "cbz_field_width", "cbz_set_scene",
    "cbz_set_field", "cbz_get_field", "cbz_num_worlds", "cbz_num_bodies",
    "cbz_calculate_derived", "cbz_integrate", "cbz_narrowphase", "cbz_narrowphase_pairs",
    "cbz_contact_counts", "cbz_get_contacts", "cbz_set_contacts", "cbz_resolve",
    "cbz_resolver_iterations", "cbz_step", "cbz_synchronize", "cbz_set_profiling",
    "cbz_phase_times", "cbz_selftest_f64", "cbz_selftest_sort",
    "cbz_sur_num_params", "cbz_sur_set_params", "cbz_sur_get_params", "cbz_sur_set_norm", "cbz_sur_get_norm",
    "cbz_sur_reserve", "cbz_sur_collect", "cbz_sur_dataset", "cbz_sur_load_dataset", "cbz_sur_init",
    "cbz_sur_train", "cbz_sur_grad", "cbz_sur_predict", "cbz_sur_resolve", "cbz_sur_world_mask", "cbz_sur_phase_times", "cbz_step_sur", "cbz_integrate_timed",
    "cbz_set_materials", "cbz_set_plane_materials", "cbz_set_contact_materials",
    "cbz_checkpoint_size", "cbz_checkpoint_save", "cbz_checkpoint_load", "cbz_narrowphase_broad",
    "cbz_resolver_occupancy", "cbz_sur_check_gemms", "cbz_resolver_world_times", "cbz_set_resolver_split", "cbz_get_resolver_split",
    "cbz_autotune_resolver",
    "cbz_sur_predict_device", "cbz_sur_position_mask", "cbz_sur_set_position", "cbz_sur_get_position",
    "cbz_pair_contact_counts",
)


class Config(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("device", ctypes.c_int32),
                ("contact_capacity", ctypes.c_int64)]


class ContactView(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("body0", "body1", "point", "normal", "pen", "c2w", "rel", "cvel", "ddv", "friction",
                 "restitution")]


class CubezError(RuntimeError):
    """A failed cbz_* call; ``code`` is the CBZ_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"cubez_amd error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the built library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "cbz_create": ([ctypes.POINTER(Config), ctypes.POINTER(P)], i32),
        "cbz_destroy": ([P], i32),
        "cbz_last_error": ([P], ctypes.c_char_p),
        "cbz_field_width": ([i32], i32),
        "cbz_set_scene": ([P, i32, P, P, P, P], i32),
        "cbz_set_field": ([P, i32, P], i32),
        "cbz_get_field": ([P, i32, P], i32),
        "cbz_num_worlds": ([P], i32),
        "cbz_num_bodies": ([P], i64),
        "cbz_calculate_derived": ([P], i32),
        "cbz_integrate": ([P, f64], i32),
        "cbz_narrowphase": ([P, P], i32),
        "cbz_narrowphase_pairs": ([P, i64, P, P, P], i32),
        "cbz_contact_counts": ([P, P], i32),
        "cbz_get_contacts": ([P, i32, ctypes.POINTER(ContactView)], i32),
        "cbz_set_contacts": ([P, P, P, P, P, P, P], i32),
        "cbz_resolve": ([P, f64, i64, i64], i32),
        "cbz_resolver_iterations": ([P, P], i32),
        "cbz_step": ([P, f64, i32, i64, P], i32),
        "cbz_synchronize": ([P], i32),
        "cbz_set_profiling": ([P, i32], i32),
        "cbz_phase_times": ([P, P, P], i32),
        "cbz_selftest_f64": ([P, i32, i64, P, P, P], i32),
        "cbz_selftest_sort": ([P, i64, P, P, P, P], i32),
        "cbz_sur_num_params": ([], i64),
        "cbz_sur_set_params": ([P, P], i32),
        "cbz_sur_get_params": ([P, P], i32),
        "cbz_sur_set_norm": ([P, P], i32),
        "cbz_sur_get_norm": ([P, P], i32),
        "cbz_sur_reserve": ([P, i64], i32),
        "cbz_sur_collect": ([P, f64, i64, P], i32),
        "cbz_sur_dataset": ([P, P, P, P], i32),
        "cbz_sur_load_dataset": ([P, i64, P, P], i32),
        "cbz_sur_init": ([P, ctypes.c_uint64], i32),
        "cbz_sur_train": ([P, i32, i32, ctypes.c_float, ctypes.c_uint64, P], i32),
        "cbz_sur_grad": ([P, i32, P, P, P, P], i32),
        "cbz_sur_predict": ([P, i32, P, P, i32], i32),
        "cbz_sur_check_gemms": ([i32, P, i32], i32),
        "cbz_resolver_world_times": ([P, P], i32),
        "cbz_set_resolver_split": ([P, i32, i32], i32),
        "cbz_get_resolver_split": ([P, P, P], i32),
        "cbz_autotune_resolver": ([P, f64, i64, i32, i32, f64, i64, i32, P, P, P, P], i32),
        "cbz_sur_predict_device": ([P, i32, P, P], i32),
        "cbz_sur_resolve": ([P, f64, i64, f64, i64, i32, P], i32),
        "cbz_sur_world_mask": ([P, P], i32),
        "cbz_sur_position_mask": ([P, P], i32),
        "cbz_pair_contact_counts": ([P, i64, P], i32),
        "cbz_sur_set_position": ([P, f64, i64], i32),
        "cbz_sur_get_position": ([P, P, P], i32),
        "cbz_step_sur": ([P, f64, i32, i64, f64, i64, i32, P, P], i32),
        "cbz_sur_phase_times": ([P, P, P], i32),
        "cbz_integrate_timed": ([P, f64, i32, P, P], i32),
        "cbz_set_materials": ([P, i32], i32),
        "cbz_set_plane_materials": ([P, P, P], i32),
        "cbz_set_contact_materials": ([P, P, P], i32),
        "cbz_checkpoint_size": ([P, P], i32),
        "cbz_checkpoint_save": ([P, P, i64], i32),
        "cbz_checkpoint_load": ([P, P, i64], i32),
        "cbz_narrowphase_broad": ([P, P, P], i32),
        "cbz_resolver_occupancy": ([P, P, P], i32),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L
