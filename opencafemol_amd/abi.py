"""ctypes mirror of include/ocm.h and the loader of the HIP library.

The loader never falls back to anything: if ``libocm.so`` has not been built
it raises, so a caller on a GPU box cannot silently run something else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libocm.so")

OCM_OK = 0
OCM_ERR_INVALID = 1
OCM_ERR_OOM = 2
OCM_ERR_HIP = 3
OCM_ERR_NUMERIC = 4
OCM_ERR_NO_DEVICE = 5
OCM_ERR_STATE = 6
OCM_ERR_TIMEOUT = 7

HARMONIC_BOND = 1
GAUSSIAN_BOND = 2
GO_CONTACT = 3
HARMONIC_ANGLE = 4
FLEXIBLE_LOCAL_ANGLE = 5
GAUSSIAN_DIHEDRAL = 6
FLEXIBLE_LOCAL_DIHEDRAL = 7
CAPPED_GO_CONTACT = 8
COSINE_DIHEDRAL = 9
EXCLUDED_VOLUME = 16
DEBYE_HUCKEL = 17
LJ_REPULSIVE = 18
LJ_ATTRACTIVE = 19
WCA = 20
UNIFORM_WCA = 21
UNIFORM_LJ_ATTRACTIVE = 22
ISOLF_ATTRACTIVE = 23

BONDED_ARITY = {1: 2, 2: 2, 3: 2, 4: 3, 5: 3, 6: 4, 7: 4, 8: 2, 9: 4}
BONDED_NPARAMS = {1: 2, 2: 3, 3: 2, 4: 2, 5: 1, 6: 3, 7: 1, 8: 3, 9: 3}
PAIR_NPP = {16: 1, 17: 1, 18: 2, 19: 2, 20: 2, 21: 0, 22: 0, 23: 3}

STATE_POSITIONS = 1
STATE_VELOCITIES = 2
STATE_FORCES = 4
STATE_ENERGY = 8

# Every symbol include/ocm.h declares.
EXPORTED = [
    "ocm_options_default", "ocm_context_create", "ocm_context_destroy", "ocm_set_positions",
    "ocm_set_velocities", "ocm_set_velocities_to_temperature", "ocm_step", "ocm_prepare_steps", "ocm_get_state",
    "ocm_get_neighbor_pairs", "ocm_get_neighbor_pairs_noalloc", "ocm_set_timing", "ocm_get_timing", "ocm_get_list_stats",
    "ocm_get_inner_list_stats", "ocm_get_list_counts", "ocm_get_launch_shape", "ocm_get_resident",
    "ocm_batch_create", "ocm_batch_set_positions", "ocm_batch_set_velocities",
    "ocm_batch_get_state", "ocm_batch_system_status", "ocm_batch_size", "ocm_last_error",
    "ocm_batch_write_checkpoint", "ocm_batch_read_checkpoint", "ocm_write_checkpoint", "ocm_read_checkpoint",
    "ocm_version", "ocm_minimize", "ocm_server_options_default", "ocm_server_create",
    "ocm_server_destroy", "ocm_server_submit", "ocm_server_next_frame", "ocm_server_wait",
    "ocm_server_get_stats",
]


class ForceDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("group", C.c_int32), ("use_periodic", C.c_int32),
        ("n_terms", C.c_int32), ("indices", C.POINTER(C.c_int32)),
        ("params", C.POINTER(C.c_double)), ("n_globals", C.c_int32),
        ("n_categories", C.c_int32), ("globals", C.POINTER(C.c_double)),
        ("particle_params", C.POINTER(C.c_double)), ("category", C.POINTER(C.c_int32)),
        ("category_allow", C.POINTER(C.c_uint8)), ("cutoff", C.c_double),
        ("switch_distance", C.c_double), ("n_exclusions", C.c_int64),
        ("exclusions", C.POINTER(C.c_int32)),
    ]


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int32), ("n_forces", C.c_int32), ("masses", C.POINTER(C.c_double)),
        ("use_periodic_box", C.c_int32), ("reserved", C.c_int32), ("box", C.c_double * 3),
        ("forces", C.POINTER(ForceDesc)),
    ]


class IntegratorDesc(C.Structure):
    _fields_ = [("temperature", C.c_double), ("friction", C.c_double), ("dt", C.c_double),
                ("seed", C.c_uint64)]


class Options(C.Structure):
    _fields_ = [("device", C.c_int32), ("rebuild_interval", C.c_int32), ("skin", C.c_double),
                ("team_size", C.c_int32), ("use_graph", C.c_int32), ("chunk_steps", C.c_int32),
                ("prune_interval", C.c_int32), ("resident", C.c_int32), ("entries_per_lane", C.c_int32)]


class State(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_double)), ("velocities", C.POINTER(C.c_double)),
                ("forces", C.POINTER(C.c_double)), ("energy_by_group", C.c_double * 32),
                ("potential_energy", C.c_double), ("kinetic_energy", C.c_double),
                ("step", C.c_int64)]


class ServerOptions(C.Structure):
    _fields_ = [("engine", Options), ("hbm_budget_bytes", C.c_int64), ("max_active_jobs", C.c_int32),
                ("max_queued_frames", C.c_int32), ("quantum_steps", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("active", C.c_int64), ("queued", C.c_int64), ("repacks", C.c_int64),
                ("frames", C.c_int64), ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double),
                ("wait_p50_ms", C.c_double), ("wait_p99_ms", C.c_double), ("simulated_ns", C.c_double),
                ("busy_ms", C.c_double), ("repack_ms", C.c_double), ("makespan_ms", C.c_double),
                ("device_bytes", C.c_int64), ("budget_bytes", C.c_int64)]


class JobInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_particles", C.c_int32), ("steps_done", C.c_int64),
                ("submit_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double)]


class OcmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ocm error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path=LIB_PATH):
    """Loads libocm.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP library {path} is missing: build it with `python -m opencafemol_amd.build`")
    lib = C.CDLL(path)
    P = C.POINTER
    sig = {
        "ocm_options_default": (None, [P(Options)]),
        "ocm_context_create": (C.c_int, [P(SystemDesc), P(IntegratorDesc), P(Options), P(C.c_void_p)]),
        "ocm_batch_create": (C.c_int, [P(SystemDesc), P(IntegratorDesc), C.c_int32, P(Options), P(C.c_void_p)]),
        "ocm_context_destroy": (C.c_int, [C.c_void_p]),
        "ocm_set_positions": (C.c_int, [C.c_void_p, P(C.c_double), C.c_int32]),
        "ocm_set_velocities": (C.c_int, [C.c_void_p, P(C.c_double), C.c_int32]),
        "ocm_batch_set_positions": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), C.c_int32]),
        "ocm_batch_set_velocities": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), C.c_int32]),
        "ocm_set_velocities_to_temperature": (C.c_int, [C.c_void_p, C.c_double, C.c_uint64]),
        "ocm_step": (C.c_int, [C.c_void_p, C.c_int64]),
        "ocm_prepare_steps": (C.c_int, [C.c_void_p, C.c_int64]),
        "ocm_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, P(State)]),
        "ocm_batch_get_state": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32, P(State)]),
        "ocm_get_neighbor_pairs": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_int32), C.c_int64, P(C.c_int64)]),
        "ocm_batch_write_checkpoint": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p]),
        "ocm_batch_read_checkpoint": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p]),
        "ocm_write_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
        "ocm_read_checkpoint": (C.c_int, [C.c_void_p, C.c_char_p]),
        "ocm_get_neighbor_pairs_noalloc": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_int32), C.c_int64,
                                                     P(C.c_int64)]),
        "ocm_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
        "ocm_get_timing": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_int64)]),
        "ocm_get_list_stats": (C.c_int, [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int64)]),
        "ocm_get_inner_list_stats": (C.c_int, [C.c_void_p, P(C.c_int64), P(C.c_int32)]),
        "ocm_get_list_counts": (C.c_int, [C.c_void_p, P(C.c_int64)]),
        "ocm_get_launch_shape": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_int32)]),
        "ocm_get_resident": (C.c_int, [C.c_void_p, P(C.c_int32)]),
        "ocm_batch_system_status": (C.c_int, [C.c_void_p, C.c_int32]),
        "ocm_batch_size": (C.c_int32, [C.c_void_p]),
        "ocm_last_error": (C.c_char_p, []),
        "ocm_version": (C.c_char_p, []),
        "ocm_minimize": (C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
        "ocm_server_options_default": (None, [P(ServerOptions)]),
        "ocm_server_create": (C.c_int, [P(ServerOptions), P(C.c_void_p)]),
        "ocm_server_destroy": (C.c_int, [C.c_void_p]),
        "ocm_server_submit": (C.c_int, [C.c_void_p, P(SystemDesc), P(IntegratorDesc), P(C.c_double),
                                        P(C.c_double), C.c_int32, C.c_int64, C.c_int64, P(C.c_int64)]),
        "ocm_server_next_frame": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_double),
                                            C.c_int32, P(C.c_int32)]),
        "ocm_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, P(C.c_double), P(C.c_double),
                                      C.c_int32, P(JobInfo)]),
        "ocm_server_get_stats": (C.c_int, [C.c_void_p, P(ServerStats)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != OCM_OK:
        msg = _lib.ocm_last_error().decode() if _lib is not None else ""
        raise OcmError(code, msg)
