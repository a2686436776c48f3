"""ctypes binding of libmi_rb.so (the C ABI declared in include/mi_rb.h).

The library is built in-tree by experiments_amd.build; loading fails loudly
when it is missing -- there is no CPU fallback for the step.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmi_rb.so")

OK, INVALID_ARG, CAPACITY, DEVICE, UNSUPPORTED_SHAPE, STATE = range(6)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "CAPACITY", 3: "DEVICE", 4: "UNSUPPORTED_SHAPE",
                5: "STATE", 6: "FORMAT"}

WARM_ZERO, WARM_PREVIOUS, WARM_LEARNED = 0, 1, 2
SAT_GPU, SAT_CPU_TWIN = 0, 1
REDUCE_SEQUENTIAL, REDUCE_SPLIT_PATH = 0, 1
INERTIA_REFERENCE, INERTIA_BOX = 0, 1
OVERFLOW_PAIRS, OVERFLOW_BODY_PAIRS, OVERFLOW_CONTACTS = 1, 2, 4
OVERFLOW_COLORS, OVERFLOW_TILE, OVERFLOW_POLYGON, OVERFLOW_LEVEL = 8, 16, 32, 64
OVERFLOW_LOOKBACK, OVERFLOW_UNSUPPORTED_PAIR = 128, 256
(BUF_SORTED_KV, BUF_AABB, BUF_PAIRS, BUF_CONTACTS, BUF_CONSTRAINTS, BUF_SOLVE_ORDER, BUF_COLORS,
 BUF_INV_INERTIA, BUF_PHASES, BUF_HULLS, BUF_VERTS, BUF_PLANES, BUF_FACE_INDEX, BUF_INDICES,
 BUF_EDGES, BUF_INV_MASS, BUF_INV_INERTIA_LOCAL, BUF_SCENE, BUF_IS_LARGE, BUF_SHAPE,
 BUF_FEATURES, BUF_PREV_KEYS, BUF_PREV_IMPULSES, BUF_PREDICTIONS, BUF_PREV_POINTS,
 BUF_CHILDREN, BUF_TILE_STATS) = range(27)
MLP_NW, MLP_NB = 64 * 16 + 64 * 64 + 3 * 64, 64 + 64 + 3
(STAGE_SETUP, STAGE_SORT, STAGE_PAIRS, STAGE_NARROWPHASE, STAGE_CONSTRAINTS, STAGE_COLOR,
 STAGE_SOLVE, STAGE_INTEGRATE) = range(8)


class Tuning(ctypes.Structure):
    _fields_ = [("solve_block", ctypes.c_int), ("narrow_waves", ctypes.c_int),
                ("sort_tile", ctypes.c_int), ("sort_digit_bits", ctypes.c_int)]


class Config(ctypes.Structure):
    _fields_ = [
        ("max_bodies", ctypes.c_int), ("max_shapes", ctypes.c_int),
        ("max_hull_vertices", ctypes.c_int), ("max_pairs", ctypes.c_int),
        ("max_contacts", ctypes.c_int), ("max_pairs_per_body", ctypes.c_int),
        ("dt", ctypes.c_float), ("solver_dt", ctypes.c_float), ("gravity", ctypes.c_float * 3),
        ("solver_iterations", ctypes.c_int), ("friction", ctypes.c_float),
        ("baumgarte", ctypes.c_float), ("position_drift", ctypes.c_float),
        ("contact_threshold", ctypes.c_float), ("angular_damping", ctypes.c_float),
        ("max_angular_motion", ctypes.c_float), ("cell_size", ctypes.c_float),
        ("grid_dims", ctypes.c_int * 3), ("residual_tolerance", ctypes.c_float),
        ("warm_start", ctypes.c_int), ("inertia_mode", ctypes.c_int),
        ("tile_size", ctypes.c_float), ("device", ctypes.c_int), ("sat_mode", ctypes.c_int),
        ("reduction_mode", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("num_bodies", ctypes.c_int), ("num_pairs", ctypes.c_int), ("num_contacts", ctypes.c_int),
        ("num_colors", ctypes.c_int), ("overflow_flags", ctypes.c_int),
        ("normal_iterations", ctypes.c_int), ("friction_iterations", ctypes.c_int),
        ("normal_residual", ctypes.c_float), ("friction_residual", ctypes.c_float),
        ("stage_ms", ctypes.c_float * 8), ("steps", ctypes.c_longlong),
        ("mean_normal_iterations", ctypes.c_float), ("mean_friction_iterations", ctypes.c_float),
        ("solve_bodies", ctypes.c_int), ("solve_kernel_ms", ctypes.c_float),
        ("solve_launches", ctypes.c_int), ("solve_events_ms", ctypes.c_float),
        ("face_pass_pairs", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [("cell_size", ctypes.c_float), ("inv_cell", ctypes.c_float),
                ("dims", ctypes.c_int * 3), ("cell_bits", ctypes.c_int), ("key_bits", ctypes.c_int),
                ("n_scenes", ctypes.c_int), ("n_large", ctypes.c_int), ("tile_mode", ctypes.c_int),
                ("n_tiles", ctypes.c_int), ("tiles_x", ctypes.c_int), ("tiles_z", ctypes.c_int),
                ("tile_inv", ctypes.c_float), ("tile_origin_x", ctypes.c_float),
                ("tile_origin_z", ctypes.c_float)]


class DeviceView(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_void_p), ("quat", ctypes.c_void_p), ("linvel", ctypes.c_void_p),
                ("angvel", ctypes.c_void_p), ("inv_mass", ctypes.c_void_p),
                ("pairs", ctypes.c_void_p), ("contacts", ctypes.c_void_p),
                ("stream", ctypes.c_void_p), ("num_bodies", ctypes.c_int)]


class SapDesc(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_int), ("max_pairs", ctypes.c_int)]


class GridDesc(ctypes.Structure):
    _fields_ = [("cell_size", ctypes.c_float), ("dims", ctypes.c_int * 3),
                ("max_pairs", ctypes.c_int), ("max_pairs_per_body", ctypes.c_int),
                ("levels", ctypes.c_int)]


class BulletShape(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("num_points", ctypes.c_int), ("first_point", ctypes.c_int),
                ("local_scaling", ctypes.c_float * 3), ("implicit_dims", ctypes.c_float * 3),
                ("margin", ctypes.c_float), ("plane_normal", ctypes.c_float * 3),
                ("plane_constant", ctypes.c_float)]


class BulletBody(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("inverse_mass", ctypes.c_float),
                ("basis", ctypes.c_float * 9), ("origin", ctypes.c_float * 3),
                ("quat", ctypes.c_float * 4), ("linvel", ctypes.c_float * 3),
                ("angvel", ctypes.c_float * 3), ("friction", ctypes.c_float),
                ("restitution", ctypes.c_float)]


class BulletSummary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("shapes", "shapes_registered", "unsupported_shapes",
                                           "bodies", "bodies_registered", "ground_bodies",
                                           "pointer_size", "big_endian")]


class MiRbError(RuntimeError):
    def __init__(self, fn, status):
        super().__init__("%s failed: %s" % (fn, STATUS_NAMES.get(status, status)))
        self.status = status


_lib = None
VP, I, F, CP = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_char_p
PVP = ctypes.POINTER(ctypes.c_void_p)
PI = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "mi_rb_config_reference": (None, [ctypes.POINTER(Config)]),
    "mi_rb_create": (I, [ctypes.POINTER(Config), PVP]),
    "mi_rb_destroy": (I, [VP]),
    "mi_rb_register_convex_hull": (I, [VP, VP, I, I, VP, PI]),
    "mi_rb_register_convex_polyhedron": (I, [VP, VP, I, I, VP, VP, I, VP, I, VP, I, I, PI]),
    "mi_rb_register_unsupported_shape": (I, [VP, I]),
    "mi_rb_register_sphere": (I, [VP, F, PI]),
    "mi_rb_register_plane": (I, [VP, VP, F, PI]),
    "mi_rb_register_compound": (I, [VP, I, VP, VP, VP, PI]),
    "mi_rb_register_concave_mesh": (I, [VP, VP, I, I, VP, I, VP, PI]),
    "mi_rb_obj_read": (I, [ctypes.c_char_p, VP, I, VP, I, PI, PI]),
    "mi_rb_register_concave_mesh_obj": (I, [VP, ctypes.c_char_p, VP, PI]),
    "mi_rb_register_body": (I, [VP, F, VP, VP, I, I, PI]),
    "mi_rb_upload": (I, [VP]),
    "mi_rb_set_transform": (I, [VP, I, VP, VP]),
    "mi_rb_set_linvel": (I, [VP, I, VP]),
    "mi_rb_set_body_state": (I, [VP, VP, VP, VP, VP]),
    "mi_rb_get_body_state": (I, [VP, VP, VP, VP, VP]),
    "mi_rb_step": (I, [VP, I]),
    "mi_rb_step_partial": (I, [VP, I]),
    "mi_rb_use_graph": (I, [VP, I]),
    "mi_rb_readback": (I, [VP]),
    "mi_rb_get_transform": (I, [VP, I, VP, VP]),
    "mi_rb_get_stats": (I, [VP, ctypes.POINTER(Stats)]),
    "mi_rb_device_ptrs": (I, [VP, ctypes.POINTER(DeviceView)]),
    "mi_rb_copy_transforms": (I, [VP, VP, I, I, VP]),
    "mi_rb_get_params": (I, [VP, ctypes.POINTER(Params)]),
    "mi_rb_read_buffer": (I, [VP, I, VP, ctypes.c_size_t, PI]),
    "mi_rb_save_state": (I, [VP, CP]),
    "mi_rb_load_state": (I, [VP, CP]),
    "mi_rb_snapshot_check": (I, [CP, ctypes.POINTER(Config), PI]),
    "mi_rb_set_warmstart_weights": (I, [VP, VP, VP]),
    "mi_rb_copy_training_batch": (I, [VP, VP, VP, I, PI]),
    "mi_rb_set_tuning": (I, [VP, ctypes.POINTER(Tuning)]),
    "mi_rb_set_solver": (I, [VP, I, ctypes.c_float, I]),
    "mi_rb_get_tuning": (I, [VP, ctypes.POINTER(Tuning)]),
    "mi_rb_autotune": (I, [VP, I, ctypes.POINTER(Tuning), VP]),
    "mi_rb_mlp_create": (I, [ctypes.c_uint64, PVP]),
    "mi_rb_mlp_destroy": (None, [VP]),
    "mi_rb_mlp_set_weights": (I, [VP, VP, VP]),
    "mi_rb_mlp_get_weights": (I, [VP, VP, VP]),
    "mi_rb_mlp_forward": (I, [VP, VP, I, VP]),
    "mi_rb_mlp_gradients": (I, [VP, VP, VP, I, VP, ctypes.POINTER(ctypes.c_float)]),
    "mi_rb_mlp_train": (I, [VP, VP, VP, I, I, I, ctypes.c_float, I, VP]),
    "mi_rb_last_step_ms": (I, [VP, ctypes.POINTER(ctypes.c_float)]),
    "mi_rb_set_solve_timing": (I, [VP, I]),
    "mi_rb_hull_build": (I, [VP, I, I, VP, VP, I, VP, VP, I, VP, I, VP, I, VP]),
    "mi_rb_radix_sort_create": (I, [I, PVP]),
    "mi_rb_radix_sort_destroy": (I, [VP]),
    "mi_rb_radix_sort_kv": (I, [VP, VP, I, I, VP]),
    "mi_rb_radix_sort_kv_noalloc": (I, [VP, I, I, VP]),
    "mi_rb_exclusive_scan_u32": (I, [VP, VP, VP, I, VP, VP]),
    "mi_rb_bound_search": (I, [VP, I, VP, I, I, VP]),
    "mi_rb_grid_create": (I, [ctypes.POINTER(GridDesc), I, PVP]),
    "mi_rb_grid_destroy": (I, [VP]),
    "mi_rb_grid_find_pairs": (I, [VP, VP, VP, VP, I, VP, VP, VP]),
    "mi_rb_grid_last_ms": (I, [VP, VP]),
    "mi_rb_grid_overflow": (I, [VP, PI]),
    "mi_rb_sap_create": (I, [ctypes.POINTER(SapDesc), I, PVP]),
    "mi_rb_sap_destroy": (I, [VP]),
    "mi_rb_sap_find_pairs": (I, [VP, VP, VP, VP, I, VP, VP, VP]),
    "mi_rb_sap_last_ms": (I, [VP, VP]),
    "mi_rb_sap_overflow": (I, [VP, PI]),
    "mi_rb_bullet_read": (I, [CP, VP, I, VP, I, VP, I, PI]),
    "mi_rb_load_bullet": (I, [VP, CP, I, ctypes.POINTER(BulletSummary)]),
}


def lib():
    """Load libmi_rb.so (built in-tree); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmi_rb.so is not built (run experiments_amd.build.build()); "
                               "the step has no CPU fallback")
        _lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def check(fn, status):
    if status != OK:
        raise MiRbError(fn, status)


def call(name, *args):
    st = getattr(lib(), name)(*args)
    check(name, st)
    return st
