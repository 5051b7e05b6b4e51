"""The C-ABI library loads and exports every entry point of include/zipgpu.h
(CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zipgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["zp_code_create", "zp_encode_rows", "zp_commit", "zp_merkle_root",
                 "zp_merkle_prove", "zp_combined_row", "zp_hint_free"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from zipgpu import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from zipgpu import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_status_strings_without_device():
    from zipgpu import _lib
    lib = _lib.load()
    assert lib.zp_status_string(0) == b"ok"
    assert lib.zp_status_string(1) == b"invalid parameter"


def test_commit_host_null_arguments():
    # Rejected before any device call.
    from zipgpu import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    root = (ctypes.c_uint8 * 32)()
    assert lib.zp_commit_host(None, 1, 32, None, 1, 1, None, root, None) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_commit_host(None, 1, 32, None, 1, 1, ctypes.byref(h), None,
                              None) == _lib.ZP_ERR_INVALID_PARAM


def test_encode_timing_argument_errors():
    # Checked before any device call: a pool size outside 1..2^20, a zero
    # sampling interval, and stop without start.
    from zipgpu import _lib
    lib = _lib.load()
    assert lib.zp_encode_timing_start(0, 1) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_encode_timing_start((1 << 20) + 1, 1) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_encode_timing_start(4, 0) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_encode_timing_stop(None, None, None) == _lib.ZP_ERR_INVALID_PARAM


def test_setup_validation():
    # pcs/structs.rs:74-88 asserts; checked on the host before any device work.
    import zipgpu

    class Code:
        def row_len(self):
            return 256

    assert zipgpu.ZipPlus.setup(1 << 16, Code()).num_rows == 256
    with pytest.raises(zipgpu.ShapeError):
        zipgpu.ZipPlus.setup(3 << 10, Code())
    with pytest.raises(zipgpu.ShapeError):
        zipgpu.ZipPlus.setup(128, Code())


def test_null_buffers_rejected_before_device_work():
    # Argument checks run before any HIP call, so they hold without a GPU.
    from zipgpu import _lib
    lib = _lib.load()
    root = (ctypes.c_uint8 * 32)()
    h = ctypes.c_void_p()
    assert lib.zp_merkle_root(None, 4, 8, 32, root, None) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_merkle_root(ctypes.c_void_p(64), 4, 8, 32, None, None) == \
        _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_merkle_root(ctypes.c_void_p(64), 4, 8, 0, root, None) == _lib.ZP_ERR_SHAPE
    assert lib.zp_commit(ctypes.c_void_p(64), 0, 32, None, 4, 1, ctypes.byref(h), root,
                         None) == _lib.ZP_ERR_INVALID_PARAM
    assert lib.zp_combined_row(32, 5, 4, 16, 1, None, None, None, None, None) == \
        _lib.ZP_ERR_INVALID_PARAM
    assert b"null" in lib.zp_last_error()
