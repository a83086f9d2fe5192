"""ctypes binding of libctd_hip.so, the C ABI declared in include/ctd_hip.h.

The library is built in-tree (cutrialdive_amd/libctd_hip.so) by
cutrialdive_amd/build.py.  Loading fails loudly when it is missing: there is
no CPU fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctd_hip.so")

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int32

CTD_OK = 0
CTD_E_INVALID = -1
CTD_E_DEVICE = -2
CTD_E_NOMEM = -3
CTD_E_UNSUPPORTED = -4
CTD_E_BUSY = -5

CTD_RUN_GENERIC_ONLY = 1
CTD_RUN_NO_FP64 = 2
CTD_RUN_SM_FULL_STEP = 4
CTD_RUN_SM_NO_SHORT = 16
CTD_RUN_SM_NO_SHIFT = 32
CTD_RUN_SM_NO_PAIR = 64
CTD_RUN_STEPPED = 8
CTD_BATCH_HOST_BUFFERS = 1
CTD_BATCH_NO_GRAPH = 2
CTD_BATCH_NO_AUTOTUNE = 4
CTD_BATCH_NO_FP64 = 8
CTD_BATCH_STEPPED = 16
CTD_SERVE_FRESH_BUFFERS = 1
CTD_SERVE_FIFO = 2
CTD_SERVE_OWN_QUEUES = 4


class CtdJob(ctypes.Structure):
    _fields_ = [
        ("seq_id", u32),
        ("max_factors", u32),
        ("param", u64),
        ("n0", u64),
        ("n1", u64),
        ("f0", u64),
        ("f1", u64),
    ]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, u64, u64)
DUE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
CHECKPOINT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, u64)


class CtdCallbacks(ctypes.Structure):
    _fields_ = [
        ("user", ctypes.c_void_p),
        ("progress", PROGRESS_FN),
        ("checkpoint_due", DUE_FN),
        ("checkpoint", CHECKPOINT_FN),
    ]


class CtdRunOpts(ctypes.Structure):
    _fields_ = [
        ("batch_length", u64),
        ("grid_size", i32),
        ("block_size", i32),
        ("chunk_length", u32),
        ("flags", u32),
    ]


class CtdRunStats(ctypes.Structure):
    _fields_ = [
        ("primes", u64),
        ("updates", u64),
        ("tf_ms", ctypes.c_double),
        ("sieve_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double),
        ("tf_launches", u32),
        ("sieve_launches", u32),
    ]


class CtdBatchOpts(ctypes.Structure):
    _fields_ = [("tile_length", u64), ("streams", u32), ("flags", u32)]


class CtdBatchStats(ctypes.Structure):
    _fields_ = [
        ("primes", u64),
        ("updates", u64),
        ("total_ms", ctypes.c_double),
        ("build_ms", ctypes.c_double),
        ("autotune_ms", ctypes.c_double),
        ("tiles", u32),
        ("graphs", u32),
        ("tuned", u32),
        ("reserved", u32),
    ]


class CtdServerOpts(ctypes.Structure):
    _fields_ = [("memory_budget", u64), ("tile_length", u64), ("streams", u32), ("queue_limit", u32),
                ("flags", u32), ("depth", u32)]


class CtdServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", u64),
        ("admitted", u64),
        ("completed", u64),
        ("failed", u64),
        ("tiles", u64),
        ("primes", u64),
        ("updates", u64),
        ("active_max", u32),
        ("pending_max", u32),
        ("pending", u32),
        ("reserved", u32),
        ("device_bytes_peak", u64),
        ("pinned_bytes_peak", u64),
        ("scheduler_cpu_ms", ctypes.c_double),
        ("scheduler_passes", u64),
    ]


# every symbol include/ctd_hip.h declares, with its ctypes signature
SIGNATURES = {
    "ctd_hip_abi_version": (ctypes.c_int, []),
    "ctd_hip_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "ctd_hip_close": (None, [ctypes.c_void_p]),
    "ctd_hip_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ctd_hip_device_info": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(u64), ctypes.POINTER(i32)],
    ),
    "ctd_hip_sieve": (
        ctypes.c_int,
        [ctypes.c_void_p, u64, u64, ctypes.POINTER(u64), u64, ctypes.POINTER(u64)],
    ),
    "ctd_hip_trial_factor": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(CtdJob),
            ctypes.POINTER(CtdRunOpts),
            ctypes.POINTER(CtdCallbacks),
            ctypes.c_void_p,
            ctypes.c_void_p,
            ctypes.POINTER(u64),
        ],
    ),
    "ctd_hip_trial_factor_device": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(CtdJob),
            ctypes.POINTER(CtdRunOpts),
            ctypes.c_void_p,
            ctypes.c_void_p,
            ctypes.POINTER(u64),
            ctypes.c_void_p,
        ],
    ),
    "ctd_hip_trial_factor_device_noalloc": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(CtdJob),
            ctypes.POINTER(CtdRunOpts),
            ctypes.c_void_p,
            ctypes.c_void_p,
            ctypes.POINTER(u64),
            ctypes.c_void_p,
        ],
    ),
    "ctd_hip_last_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CtdRunStats)]),
    "ctd_hip_residues": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CtdJob), ctypes.POINTER(u64), u64, ctypes.POINTER(u64), u32],
    ),
    "ctd_hip_valu_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, u32, ctypes.POINTER(ctypes.c_double)]),
    "ctd_hip_run_batch": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            ctypes.POINTER(CtdJob),
            u32,
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(ctypes.c_void_p),
            ctypes.POINTER(CtdBatchOpts),
            ctypes.POINTER(CtdBatchStats),
            ctypes.c_void_p,
        ],
    ),
    "ctd_hip_server_open": (
        ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CtdServerOpts), ctypes.POINTER(ctypes.c_void_p)]
    ),
    "ctd_hip_server_submit": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CtdJob), ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(u64)],
    ),
    "ctd_hip_server_poll": (
        ctypes.c_int, [ctypes.c_void_p, u64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    ),
    "ctd_hip_server_wait": (ctypes.c_int, [ctypes.c_void_p, u64, ctypes.POINTER(ctypes.c_double)]),
    "ctd_hip_server_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(CtdServerStats)]),
    "ctd_hip_server_close": (ctypes.c_int, [ctypes.c_void_p]),
    "ctd_hip_batch_shape": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(CtdJob), ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32)],
    ),
}

PROBE_OPS = [
    "v_add_u32",
    "v_mul_lo_u32",
    "v_mul_hi_u32",
    "v_mad_u64_u32",
    "v_mad_u32_u24",
    "v_lshl_add_u64",
    "v_cndmask_b32",
    "v_alignbit_b32",
    "v_min_u32",
    "v_lshlrev_b64",
    "v_add_f64",
    "v_fma_f64",
    "v_mul_f64",
    "v_max_f64",
    "v_cmp_eq_u32",
    "v_cmp_eq_f64",
    "v_cmp_eq_u64",
    "v_xor_b32",
    "v_sub_u32",
    "v_add_f32",
]

_lib = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Loads libctd_hip.so and binds every C ABI symbol; raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build the HIP engine first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
