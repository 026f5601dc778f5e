"""ctypes binding of the C ABI in include/sdim_hip.h (libsdim_hip.so, built in-tree).

There is no CPU fallback: if the library is missing or no GPU is visible, the
engine raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdim_hip.so")

# every symbol declared in include/sdim_hip.h
EXPORTS = (
    "sdim_ctx_create", "sdim_ctx_destroy", "sdim_last_error", "sdim_set_option", "sdim_get_option",
    "sdim_submit", "sdim_frame_submit", "sdim_simulate_noalloc", "sdim_wait", "sdim_poll", "sdim_release", "sdim_num_measurements",
    "sdim_fetch_records", "sdim_fetch_tableau", "sdim_fetch_snapshot", "sdim_ticket_stats", "sdim_replay",
    "sdim_ticket_counters", "sdim_ticket_info", "sdim_ticket_poison", "sdim_server_create", "sdim_server_submit", "sdim_server_poll",
    "sdim_server_wait", "sdim_server_result", "sdim_server_stats", "sdim_server_destroy",
)

MODE_BATCHED = 0
MODE_PLAIN = 1


class SdimJob(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("d", ctypes.c_int32),
        ("n_ops", ctypes.c_int64),
        ("ir", ctypes.c_void_p),
        ("noise_channel", ctypes.c_void_p),
        ("noise_prob", ctypes.c_void_p),
        ("shots", ctypes.c_int64),
        ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_uint64),
        ("init", ctypes.c_void_p * 6),
    ]


class EngineError(RuntimeError):
    pass


_lib = None


def load(path=None):
    """Load libsdim_hip.so (raises EngineError when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise EngineError(f"{p} is missing: build it with `python -m sdim_amd.build`")
    L = ctypes.CDLL(p)
    vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    L.sdim_ctx_create.argtypes = [i32, ctypes.POINTER(vp)]
    L.sdim_ctx_destroy.argtypes = [vp]
    L.sdim_last_error.restype = ctypes.c_char_p
    L.sdim_last_error.argtypes = []
    L.sdim_set_option.argtypes = [vp, ctypes.c_char_p, i64]
    L.sdim_get_option.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(i64)]
    L.sdim_submit.argtypes = [vp, ctypes.POINTER(SdimJob), i32, i32, ctypes.POINTER(u64)]
    L.sdim_frame_submit.argtypes = [vp, ctypes.POINTER(SdimJob), i32, ctypes.POINTER(vp), i64, ctypes.POINTER(u64)]
    L.sdim_simulate_noalloc.argtypes = [vp, ctypes.POINTER(SdimJob), i32, vp, u64, ctypes.POINTER(u64)]
    L.sdim_wait.argtypes = [vp, u64]
    L.sdim_poll.argtypes = [vp, u64, ctypes.POINTER(i32)]
    L.sdim_release.argtypes = [vp, u64]
    L.sdim_num_measurements.argtypes = [vp, u64, i32, ctypes.POINTER(i64)]
    L.sdim_fetch_records.argtypes = [vp, u64, i32, vp, i64]
    L.sdim_fetch_tableau.argtypes = [vp, u64, i32, i64, vp, vp, vp, vp, vp, vp]
    L.sdim_fetch_snapshot.argtypes = [vp, u64, i32, i64, i64, vp, vp, vp, vp, vp, vp]
    L.sdim_ticket_stats.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64),
                                    ctypes.POINTER(u64), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.sdim_replay.argtypes = [vp, u64, i32, i32, ctypes.POINTER(ctypes.c_double)]
    L.sdim_ticket_counters.argtypes = [vp, u64, ctypes.POINTER(u64), i32]
    L.sdim_ticket_info.argtypes = [vp, u64, ctypes.c_char_p, ctypes.POINTER(i64)]
    L.sdim_ticket_poison.argtypes = [vp, u64, i32]
    dp = ctypes.POINTER(ctypes.c_double)
    L.sdim_server_create.argtypes = [vp, u64, i32, i32, ctypes.POINTER(vp)]
    L.sdim_server_submit.argtypes = [vp, ctypes.POINTER(SdimJob), ctypes.POINTER(u64)]
    L.sdim_server_poll.argtypes = [vp, u64, ctypes.POINTER(i32)]
    L.sdim_server_wait.argtypes = [vp, u64, ctypes.c_double]
    L.sdim_server_result.argtypes = [vp, u64, vp, i64, ctypes.POINTER(i64), dp]
    L.sdim_server_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64),
                                    ctypes.POINTER(u64), dp, dp, dp]
    L.sdim_server_destroy.argtypes = [vp]
    for name in EXPORTS:
        f = getattr(L, name)
        if name != "sdim_last_error":
            f.restype = ctypes.c_int
    L.sdim_ticket_counters.restype = ctypes.c_int
    if path is None:
        _lib = L
    return L


def check(rc, L=None):
    if rc != 0:
        L = L or load()
        msg = L.sdim_last_error().decode("utf-8", "replace")
        if rc == -1:
            raise ValueError(msg)
        raise EngineError(f"sdim error {rc}: {msg}")
