"""ctypes binding of libpfcsieve.so (include/pfc.h).

The library must have been built in-tree (``python -m pfcsieve_amd.build``);
loading fails loudly otherwise.  There is no Python or CPU fallback for any
device computation.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB

PFC_OK = 0
ERRORS = {
    -1: "PFC_ERR_INVALID", -2: "PFC_ERR_HIP", -3: "PFC_ERR_SIEVE", -4: "PFC_ERR_FACTOR_OVERFLOW",
    -5: "PFC_ERR_VALIDATION", -6: "PFC_ERR_VERIFY", -7: "PFC_ERR_NOMEM", -8: "PFC_ERR_IO",
    -9: "PFC_ERR_CAPACITY",
}


class PfcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Factor(C.Structure):
    _fields_ = [("p", C.c_uint64), ("nc", C.c_int32), ("type", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("modes", C.c_uint32), ("nmin", C.c_uint32), ("nmax", C.c_uint32),
                ("pmin", C.c_uint64), ("pmax", C.c_uint64), ("factor_capacity", C.c_uint64)]


VERIFY_WORDS = 2 + 256  # PFC_VERIFY_WORDS


class Tuning(C.Structure):
    _fields_ = [("primes_per_lane", C.c_uint32), ("block_size", C.c_uint32),
                ("walk_chunk", C.c_uint32), ("seg_candidates", C.c_uint32), ("use_graph", C.c_uint32),
                ("skip_table_check", C.c_uint32), ("lds_segment", C.c_uint32),
                ("walk_forms_off", C.c_uint32), ("batch_walk_blocks", C.c_uint32)]


class JobStats(C.Structure):
    _fields_ = [("dev_checksum", C.c_uint64), ("primecount", C.c_uint64),
                ("raw_factors", C.c_uint64), ("segments", C.c_uint64), ("device_ms", C.c_double),
                ("walk_ms", C.c_double), ("walk_launches", C.c_uint64), ("walk_updates", C.c_uint64)]


class Ctrl(C.Structure):
    _fields_ = [("ncand", C.c_uint32), ("nfactor", C.c_uint32), ("flags", C.c_uint32),
                ("pad", C.c_uint32), ("checksum", C.c_uint64), ("primecount", C.c_uint64),
                ("walked", C.c_uint64), ("reserved", C.c_uint64)]


CTRL_BYTES = 48  # PFC_CTRL_BYTES
assert C.sizeof(Ctrl) == CTRL_BYTES


class SetupTable(C.Structure):
    _fields_ = [("chunks", C.c_void_p), ("groups", C.c_void_p), ("npos", C.c_uint32),
                ("digit_bits", C.c_uint32)]


# Every symbol include/pfc.h declares, with (restype, argtypes).
_u64, _u32, _i32, _sz, _vp, _int = C.c_uint64, C.c_uint32, C.c_int32, C.c_size_t, C.c_void_p, C.c_int
SIGNATURES = {
    "pfc_last_error": (C.c_char_p, []),
    "pfc_device_open": (_int, [_int, C.POINTER(_vp)]),
    "pfc_device_close": (_int, [_vp]),
    "pfc_device_info": (_int, [_vp, C.c_char_p, _sz, C.POINTER(_int), C.POINTER(_u64)]),
    "pfc_set_tuning": (_int, [_vp, C.POINTER(Tuning)]),
    "pfc_get_tuning": (_int, [_vp, C.POINTER(Tuning)]),
    "pfc_autotune": (_int, [_vp, C.POINTER(JobDesc), C.POINTER(Tuning)]),
    "pfc_job_footprint": (_int, [_vp, C.POINTER(JobDesc), C.POINTER(_u64)]),
    "pfc_job_create": (_int, [_vp, C.POINTER(JobDesc), C.POINTER(_vp)]),
    "pfc_job_run": (_int, [_vp, _u64]),
    "pfc_job_position": (_int, [_vp, C.POINTER(_u64)]),
    "pfc_job_seek": (_int, [_vp, _u64, _u64, _u64]),
    "pfc_job_stats_get": (_int, [_vp, C.POINTER(JobStats)]),
    "pfc_job_factors": (_int, [_vp, C.POINTER(Factor), _sz, C.POINTER(_sz)]),
    "pfc_job_clear_factors": (_int, [_vp]),
    "pfc_job_sync": (_int, [_vp]),
    "pfc_job_result": (_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), _vp, _sz, C.POINTER(_sz)]),
    "pfc_job_destroy": (_int, [_vp]),
    "pfc_job_debug_shorten_walk": (_int, [_vp, _u32]),
    "pfc_batch_run": (_int, [C.POINTER(_vp), _sz]),
    "pfc_batcher_create": (_int, [_vp, C.POINTER(_vp)]),
    "pfc_batcher_add": (_int, [_vp, _vp]),
    "pfc_batcher_pending": (_int, [_vp, C.POINTER(_sz)]),
    "pfc_batcher_round": (_int, [_vp, _u32, C.POINTER(_vp), _sz, C.POINTER(_sz)]),
    "pfc_batcher_round_work": (_int, [_vp, _u64, C.POINTER(_vp), _sz, C.POINTER(_sz)]),
    "pfc_batcher_destroy": (_int, [_vp]),
    "pfc_harvest": (_int, [C.POINTER(Factor), _sz, _int, C.POINTER(_u64), C.POINTER(_sz)]),
    "pfc_format_factor": (_int, [C.POINTER(Factor), C.c_char_p, _sz]),
    "pfc_is_prime": (_int, [_u64]),
    "pfc_verify_factor": (_int, [_u64, _u32, _i32, _i32]),
    "pfc_parse_uint64": (_int, [C.c_char_p, _u64, _u64, C.POINTER(_u64)]),
    "pfc_device_malloc": (_int, [_vp, _sz, C.POINTER(_vp)]),
    "pfc_device_free": (_int, [_vp, _vp]),
    "pfc_memcpy": (_int, [_vp, _vp, _vp, _sz, _int]),
    "pfc_device_synchronize": (_int, [_vp]),
    "pfc_stage_generate": (_int, [_u64, _u64, _vp, _u32, _vp, _vp]),
    "pfc_stage_setup": (_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(SetupTable),
                               C.POINTER(SetupTable), _u32, _vp]),
    "pfc_stage_walk": (_int, [_u32, _u32, _u32, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                              _u32, _u32, _vp]),
    "pfc_stage_check": (_int, [_u32, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "pfc_stage_verify_table": (_int, [_u32, _u32, C.POINTER(SetupTable), _vp, _vp]),
    "pfc_stage_verify_table_noalloc": (_int, [_u32, _u32, C.POINTER(SetupTable), _vp, _vp]),
    "pfc_build_setup_table": (_int, [_u32, _u32, C.POINTER(_u64), _sz, C.POINTER(_sz),
                                     C.POINTER(_u32), _sz, C.POINTER(_sz), C.POINTER(_u32), C.POINTER(_u32)]),
    "pfc_probe_mul_rates": (_int, [_vp, C.POINTER(C.c_double * 3)]),
    "pfc_probe_copy": (_int, [_vp, _u64, C.POINTER(C.c_double)]),
}

_lib = None


def lib():
    """Load libpfcsieve.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with `python -m pfcsieve_amd.build` "
                               "(there is no fallback path)")
        L = C.CDLL(LIB)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != PFC_OK:
        raise PfcError(rc, lib().pfc_last_error().decode(errors="replace"))
