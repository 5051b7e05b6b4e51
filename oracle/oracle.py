"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of the reference's commitment path
(oracle/zip_oracle.cpp).  Only tests/, the smoke check and bench.py's
cpu_baseline leg load it, as the checker or the timed CPU baseline; the
product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(_vp)


def set_threads(n):
    return load().or_set_threads(ctypes.c_int(n))


def octet_reversal(x, num_octets):
    o = ctypes.c_uint64()
    rc = load().or_octet_reversal(_u64(x), _u64(num_octets), ctypes.byref(o))
    if rc:
        raise ValueError("octet_reversal assert")
    return o.value


def iprs_params(row_len, depth=-1, rep=4):
    """-> dict(row_len, codeword_len, depth, base_len, base_dim, base_dim_log2) or None."""
    out = np.zeros(6, dtype=np.uint64)
    rc = load().or_iprs_params(_u64(row_len), _i64(depth), _u64(rep), _p(out))
    if rc:
        return None
    keys = ["row_len", "codeword_len", "depth", "base_len", "base_dim", "base_dim_log2"]
    return dict(zip(keys, (int(v) for v in out)))


def base_matrix(row_len, depth, rep=4):
    p = iprs_params(row_len, depth, rep)
    out = np.zeros((p["base_dim"], p["base_len"]), dtype=np.int64)
    assert load().or_iprs_base_matrix(_u64(row_len), _u64(depth), _u64(rep), _p(out)) == 0
    return out


def twiddles(row_len, depth, stage, rep=4):
    p = iprs_params(row_len, depth, rep)
    sub = p["base_dim"] << (3 * stage)
    out = np.zeros((sub, 7, 8), dtype=np.int64)
    assert load().or_iprs_twiddles(_u64(row_len), _u64(depth), _u64(rep), _u64(stage),
                                   _p(out)) == 0
    return out


def roots_of_unity(n):
    out = np.zeros(n, dtype=np.int64)
    assert load().or_roots_of_unity(_u64(n), _p(out)) == 0
    return out


def iprs_encode_binary(evals, num_rows, row_len, depth=-1, D=32, rep=4):
    """[num_rows*row_len] u64 bit patterns -> int64 [num_rows, N, D]."""
    evals = np.ascontiguousarray(evals, dtype=np.uint64)
    N = row_len * rep
    out = np.zeros((num_rows, N, D), dtype=np.int64)
    rc = load().or_iprs_encode_rows_binary(_u64(row_len), _i64(depth), _u64(rep), _u64(D),
                                           _p(evals), _u64(num_rows), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def iprs_encode_i32(evals, num_rows, row_len, depth=-1, rep=4):
    """[num_rows*row_len] i32 -> int64 [num_rows, N, 2] (i128 LE words)."""
    evals = np.ascontiguousarray(evals, dtype=np.int32)
    N = row_len * rep
    out = np.zeros((num_rows, N, 2), dtype=np.int64)
    rc = load().or_iprs_encode_rows_i32(_u64(row_len), _i64(depth), _u64(rep), _p(evals),
                                        _u64(num_rows), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def raa_perm(n, seed):
    out = np.zeros(n, dtype=np.uint64)
    load().or_raa_perm(_u64(n), _u64(seed), _p(out))
    return out


def shuffle(values, seed):
    a = np.ascontiguousarray(values, dtype=np.uint64).copy()
    load().or_shuffle_u64(_p(a), _u64(len(a)), _u64(seed))
    return a


def chacha12_words(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    load().or_chacha12_words(_u64(seed), _p(out), _u64(n))
    return out


def raa_encode_i64(row, rep=4, sign_flip=False):
    row = np.ascontiguousarray(row, dtype=np.int64)
    out = np.zeros(len(row) * rep, dtype=np.int64)
    rc = load().or_raa_encode_i64(_u64(len(row)), _u64(rep), ctypes.c_int(int(sign_flip)),
                                  _p(row), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def raa_encode_binary(evals, num_rows, row_len, D=32, rep=4, sign_flip=False):
    evals = np.ascontiguousarray(evals, dtype=np.uint64)
    out = np.zeros((num_rows, row_len * rep, D), dtype=np.int32)
    rc = load().or_raa_encode_rows_binary(_u64(row_len), _u64(rep), _u64(D),
                                          ctypes.c_int(int(sign_flip)), _p(evals),
                                          _u64(num_rows), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def blake3(data):
    b = np.frombuffer(bytes(data), dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    load().or_blake3(_p(b), _u64(len(data)), _p(out))
    return out.tobytes()


def blake3_parent(left, right, root):
    l = np.frombuffer(left, dtype=np.uint8)
    r = np.frombuffer(right, dtype=np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    load().or_blake3_parent(_p(l), _p(r), ctypes.c_int(int(root)), _p(out))
    return out.tobytes()


def blake3_chunk_cv(data, chunk_idx):
    b = np.frombuffer(bytes(data), dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    assert load().or_blake3_chunk_cv(_p(b), _u64(len(data)), _u64(chunk_idx), _p(out)) == 0
    return out.tobytes()


def merkle_leaves(rows_bytes, num_rows, width, elem_bytes):
    rows_bytes = np.ascontiguousarray(rows_bytes).view(np.uint8).reshape(-1)
    out = np.zeros((width, 32), dtype=np.uint8)
    rc = load().or_merkle_leaves(_p(rows_bytes), _u64(num_rows), _u64(width), _u64(elem_bytes),
                                 _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def merkle_layers(leaves):
    leaves = np.ascontiguousarray(leaves, dtype=np.uint8)
    n = leaves.shape[0]
    out = np.zeros((max(1, 2 * n - 1), 32), dtype=np.uint8)
    rc = load().or_merkle_layers(_p(leaves), _u64(n), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def merkle_root(rows_bytes, num_rows, width, elem_bytes):
    return merkle_layers(merkle_leaves(rows_bytes, num_rows, width, elem_bytes))[-1].tobytes()


def merkle_verify(root, column_bytes, leaf, leaf_count, siblings):
    col = np.frombuffer(bytes(column_bytes), dtype=np.uint8)
    sib = np.frombuffer(b"".join(siblings) or b"\0", dtype=np.uint8)
    r = np.frombuffer(root, dtype=np.uint8)
    return load().or_merkle_verify(_p(r), _p(col), _u64(len(column_bytes)), _u64(leaf),
                                   _u64(leaf_count), _p(sib), _u64(len(siblings))) == 0


KIND = {"iprs": 0, "raa": 1, "raa_signflip": 2}


def commit(kind, evals, num_rows, row_len, batch=1, depth=-1, D=32, eval_type="binpoly", rep=4,
           want_cw=False, want_layers=False):
    """ZipPlus::commit over `batch` polys stored contiguously -> dict(root, cw, layers)."""
    ev = np.ascontiguousarray(evals, dtype=np.uint64 if eval_type == "binpoly" else np.int32)
    N = row_len * rep
    if kind == "iprs":
        elem = 8 * D if eval_type == "binpoly" else 16
    else:
        elem = 4 * D
    cw = np.zeros(batch * num_rows * N * elem, dtype=np.uint8) if want_cw else None
    layers = np.zeros((2 * N - 1, 32), dtype=np.uint8) if want_layers else None
    root = np.zeros(32, dtype=np.uint8)
    rc = load().or_commit(ctypes.c_int(KIND[kind]), _u64(row_len), _i64(depth), _u64(rep),
                          ctypes.c_int(1 if eval_type == "binpoly" else 0), _u64(D), _p(ev),
                          _u64(num_rows), _u64(batch), _p(cw) if want_cw else None,
                          _p(layers) if want_layers else None, _p(root))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return {"root": root.tobytes(), "cw": cw, "layers": layers}


def combined_row(evals, alphas, coeffs, num_rows, row_len, batch=1, D=64, L=5):
    """-> uint64 [row_len, L] limbs of Int<L>."""
    ev = np.ascontiguousarray(evals, dtype=np.uint64)
    al = np.ascontiguousarray(alphas, dtype=np.int64)
    co = np.ascontiguousarray(coeffs, dtype=np.int64)
    out = np.zeros((row_len, L), dtype=np.uint64)
    rc = load().or_combined_row_binary(_u64(D), _u64(L), _u64(num_rows), _u64(row_len),
                                       _u64(batch), _p(ev), _p(al), _p(co), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def alpha_projection(evals, alphas, D=64, L=5):
    ev = np.ascontiguousarray(evals, dtype=np.uint64)
    al = np.ascontiguousarray(alphas, dtype=np.int64)
    out = np.zeros((len(ev), L), dtype=np.uint64)
    assert load().or_alpha_projection(_u64(D), _u64(L), _u64(len(ev)), _p(ev), _p(al),
                                      _p(out)) == 0
    return out


def iprs_encode_wide(rows, row_len, depth=-1, L=5, rep=4):
    """LinearCode::encode_wide for IPRS: uint64 [num_rows, row_len, L] limbs of
    Int<L> -> uint64 [num_rows, N, L] (arithmetic mod 2^(64L))."""
    x = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1, row_len, L)
    out = np.zeros((x.shape[0], row_len * rep, L), dtype=np.uint64)
    rc = load().or_iprs_encode_wide(_u64(row_len), _i64(depth), _u64(rep), _u64(L), _p(x),
                                    _u64(x.shape[0]), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def raa_encode_wide(rows, row_len, L=5, rep=4, sign_flip=False):
    """LinearCode::encode_wide for RAA / RaaSignFlipping over Int<L> limbs."""
    x = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1, row_len, L)
    out = np.zeros((x.shape[0], row_len * rep, L), dtype=np.uint64)
    rc = load().or_raa_encode_wide(_u64(row_len), _u64(rep), ctypes.c_int(int(sign_flip)),
                                   _u64(L), _p(x), _u64(x.shape[0]), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def limbs_add(a, b):
    """Limb-wise a + b mod 2^(64L) over the last axis (uint64 limbs)."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    out = np.empty_like(a)
    carry = np.zeros(a.shape[:-1], dtype=np.uint64)
    for i in range(a.shape[-1]):
        s = a[..., i] + b[..., i]
        c1 = (s < a[..., i]).astype(np.uint64)
        s2 = s + carry
        c2 = (s2 < s).astype(np.uint64)
        out[..., i] = s2
        carry = c1 | c2
    return out


def sign_extend_limbs(words, L):
    """int64 [..., k] little-endian words of a two's complement value -> uint64 [..., L]."""
    w = np.asarray(words, dtype=np.int64)
    ext = np.where(w[..., -1:] < 0, np.int64(-1), np.int64(0))
    pad = np.repeat(ext, L - w.shape[-1], axis=-1)
    return np.concatenate([w, pad], axis=-1).view(np.uint64)
