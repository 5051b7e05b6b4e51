"""ctypes binding of libzipgpu.so (the C ABI of include/zipgpu.h).

The library is loaded from the package's parent directory, where the build
puts it.  There is no fallback: if the HIP library is missing, every call
raises, so a GPU run can never silently route around the device code.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ZIPGPU_LIB names another build of the same library (an instrumented one).
LIB_PATH = os.environ.get("ZIPGPU_LIB") or os.path.join(os.path.dirname(_HERE), "libzipgpu.so")

ZP_OK = 0
ZP_ERR_INVALID_PARAM = 1
ZP_ERR_OVERFLOW = 2
ZP_ERR_SHAPE = 3
ZP_ERR_DEVICE = 4
ZP_ERR_OOM = 5
ZP_ERR_TIMEOUT = 6

ZP_CODE_IPRS = 0
ZP_CODE_RAA = 1
ZP_CODE_RAA_SIGNFLIP = 2

ZP_EVAL_I32 = 0
ZP_EVAL_BINPOLY = 1
ZP_EVAL_I64 = 2
ZP_EVAL_DENSE_I64 = 3

ZP_BATCH_GRAPH = 1
ZP_BATCH_AUTOTUNE = 2

_c_void_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_int = ctypes.c_int


class CommitDesc(ctypes.Structure):
    """zp_commit_desc."""

    _fields_ = [("code", ctypes.c_void_p), ("eval_type", ctypes.c_int32),
                ("degree_plus_one", ctypes.c_uint32), ("num_rows", ctypes.c_uint64),
                ("batch", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    """zp_server_config."""

    _fields_ = [("hbm_budget", ctypes.c_uint64), ("slots", ctypes.c_uint32)]


class ServerEvent(ctypes.Structure):
    """zp_server_event."""

    _fields_ = [("request", ctypes.c_uint64), ("kind", ctypes.c_int32),
                ("status", ctypes.c_int32), ("root", ctypes.c_uint8 * 32),
                ("queued_us", ctypes.c_double), ("service_us", ctypes.c_double),
                ("latency_us", ctypes.c_double), ("message", ctypes.c_char * 128),
                ("proof_bytes", ctypes.c_uint64)]


# name -> (restype, argtypes); mirrors include/zipgpu.h one for one.
SIGNATURES = {
    "zp_init": (_int, [_int]),
    "zp_status_string": (ctypes.c_char_p, [_int]),
    "zp_last_error": (ctypes.c_char_p, []),
    "zp_set_option": (_int, [ctypes.c_char_p, ctypes.c_int64]),
    "zp_encode_timing_start": (_int, [_u32, _u32]),
    "zp_encode_timing_stop": (_int, [ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_u64)]),
    "zp_code_create": (_int, [_int, _u32, _u64, _i32, ctypes.POINTER(_c_void_p)]),
    "zp_code_destroy": (None, [_c_void_p]),
    "zp_code_set_checked": (_int, [_c_void_p, _int]),
    "zp_code_is_checked": (_int, [_c_void_p]),
    "zp_code_row_len": (_u64, [_c_void_p]),
    "zp_code_codeword_len": (_u64, [_c_void_p]),
    "zp_code_depth": (_u32, [_c_void_p]),
    "zp_code_kind_of": (_int, [_c_void_p]),
    "zp_cw_elem_bytes": (_u64, [_c_void_p, _int, _u32]),
    "zp_encode_rows": (_int, [_c_void_p, _int, _u32, _c_void_p, _u64, _c_void_p, _c_void_p]),
    "zp_encode_wide": (_int, [_c_void_p, _u32, _c_void_p, _u64, _c_void_p, _c_void_p]),
    "zp_merkle_root": (_int, [_c_void_p, _u64, _u64, _u32, _c_void_p, _c_void_p]),
    "zp_commit": (_int, [_c_void_p, _int, _u32, _c_void_p, _u64, _u32,
                         ctypes.POINTER(_c_void_p), _c_void_p, _c_void_p]),
    "zp_commit_host": (_int, [_c_void_p, _int, _u32, _c_void_p, _u64, _u32,
                              ctypes.POINTER(_c_void_p), _c_void_p, _c_void_p]),
    "zp_commit_noalloc": (_int, [_c_void_p, _int, _u32, _c_void_p, _u64, _u32, _c_void_p,
                                 _c_void_p, _c_void_p]),
    "zp_hint_info": (_int, [_c_void_p, ctypes.POINTER(_u64), ctypes.POINTER(_u32),
                            ctypes.POINTER(_u64), ctypes.POINTER(_u32)]),
    "zp_hint_codewords": (_c_void_p, [_c_void_p]),
    "zp_hint_copy_codewords": (_int, [_c_void_p, _u32, _c_void_p, _c_void_p]),
    "zp_hint_copy_layer": (_int, [_c_void_p, _u32, _c_void_p, _c_void_p]),
    "zp_merkle_prove": (_int, [_c_void_p, _u64, _c_void_p, ctypes.POINTER(_u32), _c_void_p]),
    "zp_hint_open_column": (_int, [_c_void_p, _u64, _c_void_p, _c_void_p]),
    "zp_hint_free": (None, [_c_void_p]),
    "zp_hint_record_stream": (_int, [_c_void_p, _c_void_p]),
    "zp_validate_input": (_int, [_u32, _u64, _int, _u32, _u32, _u32]),
    "zp_batch_create": (_int, [_c_void_p, _u32, _u32, _u32, ctypes.POINTER(_c_void_p)]),
    "zp_batch_commit": (_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "zp_batch_hint": (_c_void_p, [_c_void_p, _u32]),
    "zp_batch_info": (_int, [_c_void_p, ctypes.POINTER(_u64), ctypes.POINTER(_u64),
                             ctypes.POINTER(_u32)]),
    "zp_batch_schedule": (_int, [_c_void_p, ctypes.POINTER(_u32), ctypes.POINTER(_u32)]),
    "zp_batch_set_schedule": (_int, [_c_void_p, _u32, _u32]),
    "zp_batch_destroy": (None, [_c_void_p]),
    "zp_autotune": (_int, [_c_void_p, _int, _u32, _u64, _u32, _c_void_p]),
    "zp_code_tuning": (_int, [_c_void_p, ctypes.POINTER(_i32), ctypes.POINTER(_i32),
                              ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "zp_server_create": (_int, [_c_void_p, ctypes.POINTER(_c_void_p)]),
    "zp_server_commit": (_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p,
                                ctypes.POINTER(_u64)]),
    "zp_server_open_sizes": (_int, [_c_void_p, _u64, _u32, _u32, ctypes.POINTER(_u64),
                                    ctypes.POINTER(_u64), ctypes.POINTER(_u64)]),
    "zp_server_open": (_int, [_c_void_p, _u64, _u32, _c_void_p, _c_void_p, _c_void_p, _u32,
                              _c_void_p, _c_void_p, _c_void_p]),
    "zp_server_prove": (_int, [_c_void_p, _u64, _c_void_p, _u32, _u32, _u32]),
    "zp_server_take_proof": (_int, [_c_void_p, _u64, _c_void_p, _u64, ctypes.POINTER(_u64),
                                    _c_void_p, _c_void_p]),
    "zp_server_poll": (_int, [_c_void_p, ctypes.c_int64, _c_void_p]),
    "zp_server_stats": (_int, [_c_void_p] + [ctypes.POINTER(_u64)] * 5),
    "zp_server_destroy": (None, [_c_void_p]),
    "zp_combined_row": (_int, [_u32, _u32, _u64, _u64, _u32, _c_void_p, _c_void_p, _c_void_p,
                               _c_void_p, _c_void_p]),
    "zp_transcript_create": (_int, [ctypes.POINTER(_c_void_p)]),
    "zp_transcript_clone": (_int, [_c_void_p, ctypes.POINTER(_c_void_p)]),
    "zp_transcript_destroy": (None, [_c_void_p]),
    "zp_transcript_absorb_slice": (_int, [_c_void_p, _c_void_p, _u64]),
    "zp_transcript_absorb_entries": (_int, [_c_void_p, _c_void_p, _u64, _u64]),
    "zp_transcript_absorb_entries_device": (_int, [_c_void_p, _c_void_p, _u64, _u64, _c_void_p]),
    "zp_transcript_challenge": (_int, [_c_void_p, _c_void_p, _u32]),
    "zp_transcript_squeeze_idx": (_int, [_c_void_p, _u64, ctypes.POINTER(_u64)]),
    "zp_transcript_random_field_cfg": (_int, [_c_void_p, _c_void_p]),
    "zp_transcript_write": (_int, [_c_void_p, _c_void_p, _u64]),
    "zp_transcript_stream_len": (_u64, [_c_void_p]),
    "zp_transcript_stream": (_int, [_c_void_p, _c_void_p, _u64]),
    "zp_prove": (_int, [_c_void_p, _int, _u32, _u32, _c_void_p, _u64, _u32, _c_void_p,
                        _c_void_p, _u32, _c_void_p, _u32, _c_void_p, _c_void_p, _c_void_p]),
}

_lib = None


def load():
    """Load libzipgpu.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libzipgpu.so not found at {LIB_PATH}: build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().zp_last_error().decode(errors="replace")
