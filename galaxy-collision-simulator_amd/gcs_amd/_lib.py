"""ctypes binding of libgcs_amd.so (the C ABI declared in include/gcs.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes
import os

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "libgcs_amd.so")

# types.h:5-12 Star (32 bytes) and types.h:14-18 PlaybackStar (16 bytes).
STAR = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"), ("vy", "<f4"),
                 ("vz", "<f4"), ("mass", "<f4"), ("type", "u1"), ("r", "u1"), ("g", "u1"),
                 ("b", "u1")])
PLAYBACK_STAR = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("r", "u1"), ("g", "u1"),
                          ("b", "u1"), ("padding", "u1")])
assert STAR.itemsize == 32 and PLAYBACK_STAR.itemsize == 16

OK = 0
ERR_INVALID = -1
ERR_NO_DEVICE = -2
ERR_OOM = -3
ERR_HIP = -4
ERR_CAPACITY = -5
ERR_STACK = -6
ERR_NONFINITE = -7
ERR_NOT_FOUND = -8

NOALLOC_MAX_KEYS = 1 << 22  # GCS_NOALLOC_MAX_KEYS

MAC_REFERENCE = 0
MAC_GROUP = 1
PREC_PLAIN = 0
PREC_COMPENSATED = 1


class Params(ctypes.Structure):
    _fields_ = [("G", ctypes.c_float), ("eps2", ctypes.c_float), ("dt", ctypes.c_float),
                ("theta", ctypes.c_float), ("mac", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("leaf_size", ctypes.c_int32),
                ("block_size", ctypes.c_int32), ("stack_depth", ctypes.c_int32),
                ("use_bounds", ctypes.c_int32), ("global_min", ctypes.c_float),
                ("global_max", ctypes.c_float), ("colours", ctypes.c_int32),
                ("diagnostics", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32), ("dt_eta", ctypes.c_float),
                ("group_size", ctypes.c_int32), ("open_offset", ctypes.c_int32),
                ("team", ctypes.c_int32), ("mono_far", ctypes.c_float),
                ("local_kappa", ctypes.c_float), ("instance_pass", ctypes.c_int32),
                ("sort", ctypes.c_int32)]


class CollisionSpec(ctypes.Structure):
    _fields_ = [("n_a", ctypes.c_int32), ("n_b", ctypes.c_int32),
                ("offset", ctypes.c_float * 3), ("velocity", ctypes.c_float * 3),
                ("tilt_deg", ctypes.c_float), ("seed", ctypes.c_uint64)]


class Job(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("n", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("arrival_s", ctypes.c_double)]


class JobResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int64), ("n", ctypes.c_int32), ("steps", ctypes.c_int32),
                ("arrival_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("finish_s", ctypes.c_double)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_resident", ctypes.c_int64),
                ("max_jobs", ctypes.c_int32), ("steps_per_round", ctypes.c_int32),
                ("gen_threads", ctypes.c_int32), ("keep_final", ctypes.c_int32),
                ("params", Params)]


class GcsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gcs error %d: %s" % (code, msg))
        self.code = code


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F = ctypes.c_float

SIGNATURES = {
    "gcs_last_error": ([], ctypes.c_char_p),
    "gcs_params_default": ([_P], None),
    "gcs_device_bytes": ([_I64, _I32], _I64),
    "gcs_create": ([_I64, _I32, _I32, _P], _I32),
    "gcs_destroy": ([_P], _I32),
    "gcs_set_stream": ([_P, _P], _I32),
    "gcs_load": ([_P, _P, _P, _I32], _I32),
    "gcs_instance_add": ([_P, _P, _I32, _P], _I32),
    "gcs_instance_remove": ([_P, _I32], _I32),
    "gcs_instance_count": ([_P, _P, _P], _I32),
    "gcs_instance_info": ([_P, _I32, _P, _P], _I32),
    "gcs_read_stars": ([_P, _P], _I32),
    "gcs_read_instance": ([_P, _I32, _P], _I32),
    "gcs_read_position_lo": ([_P, _P], _I32),
    "gcs_read_playback": ([_P, _P], _I32),
    "gcs_bake": ([_P, _P, _I32, ctypes.c_char_p, _P], _I32),
    "gcs_step": ([_P, _P, _I32], _I32),
    "gcs_sync": ([_P], _I32),
    "gcs_compute_accels": ([_P, _P], _I32),
    "gcs_read_accels": ([_P, _P], _I32),
    "gcs_read_keys": ([_P, _P, _P, _P], _I32),
    "gcs_read_tree": ([_P, _I32, _P, _P, _P, _P, _P, _P, _P], _I32),
    "gcs_read_bounds": ([_P, _I32, _P, _P], _I32),
    "gcs_read_root": ([_P, _I32, _P], _I32),
    "gcs_read_interactions": ([_P, _P], _I32),
    "gcs_read_monopole_interactions": ([_P, _P], _I32),
    "gcs_read_interaction_counts": ([_P, _P], _I32),
    "gcs_read_diagnostics": ([_P, _I32, _P], _I32),
    "gcs_read_stage_ms": ([_P, _P], _I32),
    "gcs_engine_info": ([_P, _P], _I32),
    "gcs_morton_keys": ([_P, _P, _P, _I64, _F, _F, _P], _I32),
    "gcs_sort_pairs": ([_P, _I64, _P, _P], _I32),
    "gcs_karras_tree": ([_P, _I32, _P, _P, _P], _I32),
    "gcs_sort_pairs_noalloc": ([_P, _I64, _P, _P, _P], _I32),
    "gcs_spec_default": ([_I32, ctypes.c_uint64, _P], None),
    "gcs_spec_random": ([ctypes.c_uint64, _I64, _I32, _I32, _P], None),
    "gcs_generate": ([_P, _P], _I32),
    "gcs_load_generated": ([_P, _P, _I32], _I32),
    "gcs_instance_add_generated": ([_P, _P, _P], _I32),
    "gcs_autotune": ([_P, _P, _I32, _P, _P], _I32),
    "gcs_autotune_cache_clear": ([], _I32),
    "gcs_autotune_cache_save": ([ctypes.c_char_p], _I32),
    "gcs_autotune_cache_load": ([ctypes.c_char_p], _I32),
    "gcs_read_max_accel": ([_P, _I32, _P], _I32),
    "gcs_read_sim_time": ([_P, _I32, _P], _I32),
    "gcs_checkpoint_bytes": ([_I32], ctypes.c_int64),
    "gcs_checkpoint": ([_P, _I32, _P, ctypes.c_int64], _I32),
    "gcs_restore": ([_P, _P, ctypes.c_int64, _P], _I32),
    "gcs_server_create": ([_P, _P], _I32),
    "gcs_server_destroy": ([_P], _I32),
    "gcs_server_submit": ([_P, _P, _I32], _I32),
    "gcs_server_run": ([_P, ctypes.c_double], _I32),
    "gcs_server_results": ([_P, _P, _I32, _P], _I32),
    "gcs_server_stats": ([_P, _P, _I32], _I32),
    "gcs_server_final": ([_P, ctypes.c_int64, _P, _I32], _I32),
}


def lib():
    """The loaded library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libgcs_amd.so is not built (run __graft_entry__.build()): %s"
                              % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != OK:
        msg = lib().gcs_last_error()
        raise GcsError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)
