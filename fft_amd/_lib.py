"""ctypes binding of libsfft.so (the C ABI in include/sfft.h).

The library is built in-tree by `make -C fft_amd/csrc` (or __graft_entry__.build())
and loaded from fft_amd/libsfft.so.  There is no fallback: if the library is
missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfft.so")

SFFT_OK = 0
SFFT_ERR_INVALID_ARG = -1
SFFT_ERR_UNSUPPORTED_SIZE = -2
SFFT_ERR_OUT_OF_MEMORY = -3
SFFT_ERR_DEVICE = -4
SFFT_ERR_ALIASING = -5
SFFT_ERR_NOT_READY = -6
SFFT_ERR_BUDGET = -7

C2C, R2C, R2C_HALFSHIFT = 0, 1, 2
F32, F64, F16, BF16, F32X = 0, 1, 2, 3, 4
FORWARD, INVERSE = 0, 1

# Every symbol include/sfft.h declares (checked by tests/test_capi_exports.py).
EXPORTED = [
    "sfft_init", "sfft_last_error", "sfft_device_info",
    "sfft_size_minimum", "sfft_size_maximum", "sfft_real_size_minimum",
    "sfft_real_size_maximum", "sfft_size_supported",
    "sfft_plan_create", "sfft_plan_destroy", "sfft_plan_get_info",
    "sfft_plan_create_hinted", "sfft_plan_variants",
    "sfft_autotune", "sfft_tune_save", "sfft_tune_load", "sfft_tune_clear",
    "sfft_execute", "sfft_execute_strided", "sfft_plan_workspace_size", "sfft_execute_workspace",
    "sfft_execute_noalloc",
    "sfft_batch_create", "sfft_batch_destroy", "sfft_batch_execute",
    "sfft_batch_execute_per_size",
    "sfft_server_create", "sfft_server_destroy", "sfft_server_submit", "sfft_server_poll",
    "sfft_server_drain", "sfft_server_fetch", "sfft_server_release", "sfft_server_now",
    "sfft_server_run_poisson",
    "sfft_malloc_device", "sfft_free_device", "sfft_malloc_host", "sfft_free_host",
    "sfft_memcpy", "sfft_stream_synchronize", "sfft_host_device_pointer",
    "sfft_stream_create", "sfft_stream_destroy", "sfft_event_create", "sfft_event_destroy",
    "sfft_event_record", "sfft_event_synchronize",
    "sfft_fill",
]


class SfftError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sfft error {code}: {msg}")
        self.code = code


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_size_t),
        ("kind", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("passes", ctypes.c_int),
        ("nradices", ctypes.c_int),
        ("radices", ctypes.c_int * 16),
        ("threads_per_transform", ctypes.c_int),
        ("transforms_per_block", ctypes.c_int),
        ("workspace_bytes", ctypes.c_size_t),
        ("table_bytes", ctypes.c_size_t),
        ("mfma_stages", ctypes.c_int),
        ("chirp_length", ctypes.c_size_t),
        ("in_place", ctypes.c_int),
    ]


class PlanHints(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int),
        ("nradices", ctypes.c_int),
        ("radices", ctypes.c_int * 16),
        ("transforms_per_block", ctypes.c_int),
        ("split", ctypes.c_size_t),
        ("elements_per_thread", ctypes.c_int),
        ("mfma", ctypes.c_int),
    ]


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("hints", PlanHints),
        ("ms", ctypes.c_double),
        ("gbs", ctypes.c_double),
        ("batch", ctypes.c_size_t),
        ("candidates", ctypes.c_int),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_size_t),
        ("kind", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("count", ctypes.c_size_t),
        ("in_offset", ctypes.c_size_t),
        ("out_offset", ctypes.c_size_t),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_size_t),
        ("max_inflight", ctypes.c_int),
        ("streams", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("plain", ctypes.c_int),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("n", ctypes.c_size_t),
        ("kind", ctypes.c_int),
        ("direction", ctypes.c_int),
        ("batch", ctypes.c_size_t),
        ("seed", ctypes.c_uint64),
        ("retain", ctypes.c_int),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("status", ctypes.c_int),
        ("t_arrive", ctypes.c_double),
        ("t_admit", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("bytes", ctypes.c_size_t),
    ]


class PoissonConfig(ctypes.Structure):
    _fields_ = [
        ("rate", ctypes.c_double),
        ("duration", ctypes.c_double),
        ("n_min", ctypes.c_size_t),
        ("n_max", ctypes.c_size_t),
        ("batch_max", ctypes.c_size_t),
        ("seed", ctypes.c_uint64),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded libsfft.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libsfft.so not built ({LIB_PATH}); run `make -C fft_amd/csrc` "
                "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        sz, vp, ip = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int
        L.sfft_init.restype = ip
        L.sfft_last_error.restype = ctypes.c_char_p
        L.sfft_device_info.argtypes = [ctypes.POINTER(ip), ctypes.POINTER(sz), ctypes.c_char_p, sz]
        for name in ("sfft_size_minimum", "sfft_size_maximum", "sfft_real_size_minimum",
                     "sfft_real_size_maximum"):
            getattr(L, name).argtypes = [sz]
            getattr(L, name).restype = sz
        L.sfft_size_supported.argtypes = [sz, ip]
        L.sfft_plan_create.argtypes = [ctypes.POINTER(vp), sz, ip, ip]
        L.sfft_plan_destroy.argtypes = [vp]
        L.sfft_plan_create_hinted.argtypes = [ctypes.POINTER(vp), sz, ip, ip,
                                              ctypes.POINTER(PlanHints)]
        L.sfft_plan_variants.argtypes = [sz, ip, ip]
        L.sfft_autotune.argtypes = [sz, ip, ip, sz, ctypes.POINTER(TuneResult)]
        L.sfft_tune_save.argtypes = [ctypes.c_char_p]
        L.sfft_tune_load.argtypes = [ctypes.c_char_p, ip, ctypes.POINTER(sz)]
        L.sfft_plan_get_info.argtypes = [vp, ctypes.POINTER(PlanInfo)]
        L.sfft_execute.argtypes = [vp, vp, vp, sz, ip, vp]
        L.sfft_execute_strided.argtypes = [vp, vp, sz, vp, sz, sz, ip, vp]
        L.sfft_plan_workspace_size.argtypes = [vp, sz, ip, ctypes.POINTER(sz)]
        L.sfft_execute_workspace.argtypes = [vp, vp, vp, sz, ip, vp, sz, vp]
        L.sfft_execute_noalloc.argtypes = [vp, vp, vp, sz, ip, vp]
        L.sfft_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BatchItem), sz, ip]
        L.sfft_batch_destroy.argtypes = [vp]
        L.sfft_batch_execute.argtypes = [vp, vp, vp, vp]
        L.sfft_batch_execute_per_size.argtypes = [vp, vp, vp, vp]
        L.sfft_fill.argtypes = [vp, sz, ip, ip, ctypes.c_uint64, vp]
        L.sfft_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServerConfig)]
        L.sfft_server_destroy.argtypes = [vp]
        L.sfft_server_submit.argtypes = [vp, ctypes.POINTER(Request)]
        L.sfft_server_poll.argtypes = [vp, ctypes.POINTER(Completion), sz, ctypes.POINTER(sz)]
        L.sfft_server_drain.argtypes = [vp]
        L.sfft_server_fetch.argtypes = [vp, ctypes.c_uint64, vp, sz]
        L.sfft_server_release.argtypes = [vp, ctypes.c_uint64]
        L.sfft_server_now.argtypes = [vp]
        L.sfft_server_now.restype = ctypes.c_double
        L.sfft_server_run_poisson.argtypes = [vp, ctypes.POINTER(PoissonConfig),
                                              ctypes.POINTER(Completion), sz, ctypes.POINTER(sz)]
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != SFFT_OK:
        msg = lib().sfft_last_error()
        raise SfftError(rc, msg.decode() if msg else "")
