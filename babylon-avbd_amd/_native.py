"""ctypes binding of libavbd_hip.so (the C ABI declared in include/avbd.h).

The library is the only compute path: when it is missing the binding raises instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libavbd_hip.so")

AVBD_OK = 0
AVBD_ERR_INVALID = -1
AVBD_ERR_DEVICE = -2
AVBD_ERR_DESTROYED = -3
AVBD_ERR_CAPACITY = -4

FLAG_MATCH_NEAREST = 1
FLAG_FACE_BIAS = 2
FLAG_REUSE_CONTACTS = 4
FLAG_START_AT_REST = 8
FLAG_MASS_PENALTY = 16
FLAGS_DEFAULT = 31
FLAG_CAPSULE_STEP_START_ARMS = 32  # this build's option, off by default (the reference turns capsule lever arms)

u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f32p = ctypes.POINTER(ctypes.c_float)


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("body_count", ctypes.c_uint32),
        ("bodies", f32p),
        ("joint_count", ctypes.c_uint32),
        ("joints", f32p),
        ("joint_info", i32p),
        ("ignore_count", ctypes.c_uint32),
        ("ignore_pairs", u32p),
        ("filters", u32p),
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("up", ctypes.c_float * 3),
        ("alpha", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("beta_lin", ctypes.c_float),
        ("beta_ang", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("hull_words", ctypes.c_uint32),
        ("hulls", u32p),
    ]


class HullInfo(ctypes.Structure):
    _fields_ = [
        ("vertex_count", ctypes.c_uint32),
        ("face_count", ctypes.c_uint32),
        ("edge_count", ctypes.c_uint32),
        ("words", ctypes.c_uint32),
        ("volume", ctypes.c_double),
        ("moments", ctypes.c_double * 3),
        ("size", ctypes.c_double * 3),
        ("radius", ctypes.c_double),
        ("center", ctypes.c_double * 3),
        ("rotation", ctypes.c_double * 4),
    ]


HULL_MAX_WORDS = 320
SHAPE_HULL = 3.0


class Options(ctypes.Structure):
    _fields_ = [
        ("color_cap", ctypes.c_uint32),
        ("color_rounds", ctypes.c_uint32),
        ("pair_capacity", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("adaptive", ctypes.c_uint32),
        ("broadphase", ctypes.c_uint32),
        ("bvh_rebuild_interval", ctypes.c_uint32),
        ("solver", ctypes.c_uint32),
        ("coloring_rounds", ctypes.c_uint32),
        ("world_block", ctypes.c_uint32),
        ("dataflow_share", ctypes.c_uint32),
    ]


class QueryOptions(ctypes.Structure):
    """AvbdQueryOptions (include/avbd.h)."""
    _fields_ = [("ignore", ctypes.POINTER(ctypes.c_uint32)), ("ignore_count", ctypes.c_uint32),
                ("collides_with", ctypes.c_uint32), ("membership", ctypes.c_uint32), ("use_membership", ctypes.c_int)]


class Counters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint32) for k in
                ("pairs", "contacts", "manifolds", "overflow", "clashes", "colors", "joints", "bodies",
                 "pivot_fallbacks", "solve_aborts")]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("max_active", ctypes.c_uint32),
        ("lookahead", ctypes.c_uint32),
        ("device", ctypes.c_int32),
        ("color_cap", ctypes.c_uint32),
        ("color_rounds", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("broadphase", ctypes.c_uint32),
        ("small_world_bodies", ctypes.c_uint32),
        ("small_world_block", ctypes.c_uint32),
        ("large_world_bodies", ctypes.c_uint32),
        ("large_world_share", ctypes.c_uint32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in
                ("submitted", "completed", "admissions", "evictions", "restores", "steps", "body_steps")] + \
               [(k, ctypes.c_double) for k in
                ("wall_seconds", "step_ms_p50", "step_ms_p99", "step_ms_max", "request_ms_p50", "request_ms_p99")] + \
               [("resident_bytes", ctypes.c_uint64), ("peak_resident_bytes", ctypes.c_uint64),
                ("resident_worlds", ctypes.c_uint32), ("worlds", ctypes.c_uint32), ("solve_aborts", ctypes.c_uint64)]


class Params(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("up", ctypes.c_float * 3),
        ("alpha", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("beta_lin", ctypes.c_float),
        ("beta_ang", ctypes.c_float),
        ("iterations", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


COMMAND_WORDS = 12
CMD_LINEAR_VELOCITY, CMD_ANGULAR_VELOCITY, CMD_IMPULSE, CMD_FORCE, CMD_TELEPORT, CMD_ANGULAR_IMPULSE, CMD_TORQUE = range(7)


# name -> (restype, argtypes): every entry point of include/avbd.h
SIGNATURES = {
    "avbd_last_error": (ctypes.c_char_p, []),
    "avbd_version": (ctypes.c_char_p, []),
    "avbd_device_count": (ctypes.c_int, []),
    "avbd_world_create": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]),
    "avbd_hull_from_points": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint32, u32p,
                                             ctypes.c_uint32, ctypes.POINTER(HullInfo)]),
    "avbd_hull_from_triangles": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, u32p, ctypes.c_uint32,
                                                ctypes.c_uint32, u32p, ctypes.c_uint32, ctypes.POINTER(HullInfo)]),
    "avbd_batch_create": (ctypes.c_int, [ctypes.POINTER(SceneDesc), ctypes.c_uint32, ctypes.POINTER(Options),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "avbd_world_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "avbd_world_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "avbd_world_step_timed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, f32p]),
    "avbd_world_step_timed_begin": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "avbd_world_step_timed_end": (ctypes.c_int, [ctypes.c_void_p, f32p, ctypes.c_uint32, u32p]),
    "avbd_world_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "avbd_world_body_count": (ctypes.c_int, [ctypes.c_void_p, u32p]),
    "avbd_world_joint_count": (ctypes.c_int, [ctypes.c_void_p, u32p]),
    "avbd_world_read_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, f32p]),
    "avbd_world_write_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, f32p]),
    "avbd_world_read_selected": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, ctypes.c_int, f32p]),
    "avbd_world_read_joints": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, f32p]),
    "avbd_world_write_joints": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, f32p]),
    "avbd_world_read_counters": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Counters)]),
    "avbd_world_read_pairs": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, u32p]),
    "avbd_world_read_contacts": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, u32p]),
    "avbd_world_read_colors": (ctypes.c_int, [ctypes.c_void_p, u32p]),
    "avbd_world_read_adjacency": (ctypes.c_int, [ctypes.c_void_p, u32p, u32p, ctypes.c_uint32, u32p]),
    "avbd_world_read_refs": (ctypes.c_int, [ctypes.c_void_p, f32p]),
    "avbd_world_read_steps": (ctypes.c_int, [ctypes.c_void_p, u32p]),
    "avbd_world_write_state": (ctypes.c_int, [ctypes.c_void_p, f32p, u32p, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p, u32p]),
    "avbd_world_set_coloring": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "avbd_world_adapt": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "avbd_world_release_joints": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32]),
    "avbd_world_disable_constraint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "avbd_world_set_world_anchor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, f32p]),
    "avbd_world_set_motor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, f32p, f32p]),
    "avbd_world_set_spring_material": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float,
                                                      ctypes.c_float]),
    "avbd_world_set_angular_constraint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, f32p, f32p,
                                                         ctypes.c_float, ctypes.c_float]),
    "avbd_world_joint_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int32)]),
    "avbd_world_get_coloring": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32),
                                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]),
    "avbd_world_profile_step": (ctypes.c_int, [ctypes.c_void_p, f32p]),
    "avbd_world_device_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "avbd_world_solver_kind": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "avbd_world_enqueue_commands": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32]),
    "avbd_world_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Params)]),
    "avbd_world_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(Params)]),
    "avbd_world_set_filters": (ctypes.c_int, [ctypes.c_void_p, u32p, u32p, u32p, ctypes.c_uint32]),
    "avbd_world_set_restitution": (ctypes.c_int, [ctypes.c_void_p, u32p, f32p, ctypes.c_uint32]),
    "avbd_world_get_restitution": (ctypes.c_int, [ctypes.c_void_p, f32p]),
    "avbd_world_enable_sleeping": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                                  ctypes.c_int]),
    "avbd_world_disable_sleeping": (ctypes.c_int, [ctypes.c_void_p]),
    "avbd_world_sleeping": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "avbd_world_wake_all": (ctypes.c_int, [ctypes.c_void_p]),
    "avbd_world_wake_bodies": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32]),
    "avbd_world_set_sleep_enabled": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.POINTER(ctypes.c_uint8),
                                                    ctypes.c_uint32]),
    "avbd_world_sleep_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, u32p]),
    "avbd_world_read_sleep": (ctypes.c_int, [ctypes.c_void_p, u32p, u32p]),
    "avbd_world_cast": (ctypes.c_int, [ctypes.c_void_p, f32p, ctypes.c_uint32, ctypes.c_void_p, f32p]),
    "avbd_world_cast_noalloc": (ctypes.c_int, [ctypes.c_void_p, f32p, ctypes.c_uint32, ctypes.c_void_p, f32p]),
    "avbd_world_watch_contacts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]),
    "avbd_world_set_watched": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_uint32]),
    "avbd_world_reset_contact_history": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32]),
    "avbd_world_read_contact_events": (ctypes.c_int, [ctypes.c_void_p, u32p, ctypes.c_uint32, u32p, u32p, u32p]),
    "avbd_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOptions), ctypes.POINTER(ctypes.c_void_p)]),
    "avbd_server_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "avbd_server_add_world": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SceneDesc), u32p]),
    "avbd_server_world_body_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, u32p]),
    "avbd_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "avbd_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "avbd_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "avbd_server_read_bodies": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, f32p]),
    "avbd_server_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    "avbd_server_reset_stats": (ctypes.c_int, [ctypes.c_void_p]),
}

_lib = None


class AvbdError(RuntimeError):
    """An error reported by the library (the reference throws Error with a message)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def load():
    """The loaded library; raises when it has not been built (there is no other compute path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != AVBD_OK:
        msg = load().avbd_last_error().decode() or f"avbd error {code}"
        raise AvbdError(code, msg)
    return code


def ptr(array, ctype):
    """Pointer to a contiguous numpy array (None for None)."""
    if array is None:
        return None
    return array.ctypes.data_as(ctypes.POINTER(ctype))
