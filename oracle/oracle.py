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
# ZP_ORACLE_LIB names another build of the oracle (an instrumented one).
LIB_PATH = os.environ.get("ZP_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
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


def baseline_commit_iprs(evals, num_rows, row_len, D=32, threads=1, simd=True, depth=-1, rep=4,
                         want_cw=False):
    """The timed CPU baseline (oracle/baseline.cpp): ZipPlus::commit of one
    BinaryPoly<D> poly under IPRS on the reference's parallel axes, with its
    AVX-512 widening and 16-lane BLAKE3 where the CPU has AVX-512.
    -> dict(root, cw)."""
    ev = np.ascontiguousarray(evals, dtype=np.uint64)
    N = row_len * rep
    cw = np.zeros(num_rows * N * D, dtype=np.int64) if want_cw else None
    root = np.zeros(32, dtype=np.uint8)
    rc = load().or_baseline_commit_iprs_binary(_u64(row_len), _i64(depth), _u64(rep), _u64(D),
                                               _p(ev), _u64(num_rows), ctypes.c_int(threads),
                                               ctypes.c_int(1 if simd else 0),
                                               _p(cw) if want_cw else None, _p(root))
    if rc:
        raise ValueError(f"baseline rc {rc}")
    return {"root": root.tobytes(), "cw": cw}


def baseline_has_avx512():
    return bool(load().or_baseline_has_avx512())


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


def iprs_encode_wide_checked(rows, row_len, depth=-1, L=5, rep=4):
    """The checked build's IPRS encode_wide (IprsCode<.., CHECKED>): returns
    (out, overflowed); overflowed is True where the reference panics, and
    out then holds no meaningful values."""
    x = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1, row_len, L)
    out = np.zeros((x.shape[0], row_len * rep, L), dtype=np.uint64)
    rc = load().or_iprs_encode_wide_checked(_u64(row_len), _i64(depth), _u64(rep), _u64(L),
                                            _p(x), _u64(x.shape[0]), _p(out))
    if rc not in (0, 2):
        raise ValueError(f"oracle rc {rc}")
    return out, rc == 2


def raa_encode_wide_checked(rows, row_len, L=5, rep=4, sign_flip=False):
    """The checked build's RAA encode_wide (CHECK_FOR_OVERFLOWS): (out, overflowed)."""
    x = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1, row_len, L)
    out = np.zeros((x.shape[0], row_len * rep, L), dtype=np.uint64)
    rc = load().or_raa_encode_wide_checked(_u64(row_len), _u64(rep), ctypes.c_int(int(sign_flip)),
                                           _u64(L), _p(x), _u64(x.shape[0]), _p(out))
    if rc not in (0, 2):
        raise ValueError(f"oracle rc {rc}")
    return out, rc == 2


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


# ---------------------------------------------------------------------------
# Opening path: transcript, field, prove and verify (test infrastructure).
# ---------------------------------------------------------------------------

def blake3_xof(data, n):
    """BLAKE3 finalize_xof().fill of n bytes over `data`."""
    b = np.frombuffer(bytes(data), dtype=np.uint8) if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(max(1, n), dtype=np.uint8)
    load().or_blake3_xof(_p(b), _u64(len(data)), _p(out), _u64(n))
    return out[:n].tobytes()


def sprp2(n):
    """Strong probable prime to base 2 (primality/src/lib.rs:9-20)."""
    if n < 5 or n % 2 == 0:
        return n in (2, 3)
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    x = pow(2, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(s - 1):
        x = x * x % n
        if x == n - 1:
            return True
        if x == 1:
            return False
    return False


class Transcript:
    """Blake3Transcript (transcript/src/lib.rs:10-72) with the proof stream of
    PcsProverTranscript (zip-plus/src/pcs_transcript.rs:40-221).  The hash
    state is the whole absorbed byte string, rehashed at each challenge."""

    def __init__(self, stream=b""):
        self.absorbed = bytearray()
        self.stream = bytearray(stream)
        self.pos = 0

    def absorb(self, b):
        self.absorbed += bytes(b)

    def absorb_slice(self, b):  # traits.rs:388-392
        self.absorb(b"\x06" + bytes(b) + b"\x07")

    def challenge(self, n):  # lib.rs:48-55
        out = blake3_xof(self.absorbed, n)
        self.absorb(b"\x12" + out + b"\x34")
        return out

    def challenge_i128(self):
        return int.from_bytes(self.challenge(16), "little", signed=True)

    def squeeze_idx(self, cap):  # pcs_transcript.rs:31-36
        return int.from_bytes(self.challenge(4), "little") % cap

    def random_field_cfg(self):  # traits.rs:371-382, lib.rs:57-72 (gen_random)
        while True:
            buf = blake3_xof(self.absorbed, 32)
            self.absorb(buf)
            c = int.from_bytes(buf, "little")
            if c == 0:
                continue
            if c % 2 == 0:
                c -= 1
            if sprp2(c):
                return c

    def absorb_field(self, p, inner):  # traits.rs:394-412
        self.absorb(b"\x03" + p.to_bytes(32, "little") + b"\x05" + b"\x01" +
                    inner.to_bytes(32, "little") + b"\x03")

    def write_field_elements(self, p, values):  # pcs_transcript.rs:99-142
        for v in values:
            inner = (v << 256) % p
            self.absorb_field(p, inner)
            self.stream += p.to_bytes(32, "little") + inner.to_bytes(32, "little")

    def read(self, n):
        if self.pos + n > len(self.stream):
            raise ValueError("proof stream too short")
        b = bytes(self.stream[self.pos:self.pos + n])
        self.pos += n
        return b

    def read_field_elements(self, n):  # pcs_transcript.rs:238-283
        out = []
        for _ in range(n):
            p = int.from_bytes(self.read(32), "little")
            inner = int.from_bytes(self.read(32), "little")
            self.absorb_field(p, inner)
            out.append((p, inner * pow(1 << 256, -1, p) % p))
        return out


def _limbs4(v):
    return np.array([(v >> (64 * i)) & ((1 << 64) - 1) for i in range(4)], dtype=np.uint64)


def _int4(a):
    return sum(int(a[i]) << (64 * i) for i in range(4))


def _i128_words(values):
    out = np.zeros((len(values), 2), dtype=np.int64)
    for i, v in enumerate(values):
        u = v & ((1 << 128) - 1)
        out[i, 0] = np.uint64(u & ((1 << 64) - 1)).view(np.int64)
        out[i, 1] = np.uint64(u >> 64).view(np.int64)
    return out


def eq_table(r, p):
    """build_eq_x_r (poly/src/utils.rs:140-190): t[i] = prod_k (bit k of i ?
    r_k : 1 - r_k) mod p."""
    t = [1]
    for rk in r:
        t = [x * (1 - rk) % p for x in t] + [x * rk % p for x in t]
    return t


def point_to_tensor(num_rows, point, p):
    """pcs/utils.rs:93-114: q_1 over the first coordinates (columns), q_0
    over the last log2(num_rows) (rows)."""
    k = len(point) - (num_rows.bit_length() - 1)
    return eq_table(point[k:], p), eq_table(point[:k], p)


def prove_b(p, evals, alphas, q1, num_rows, row_len, batch, D, eval_type):
    """b_j of prove_f (phase_prove.rs:176-198), canonical ints."""
    binpoly = eval_type == "binpoly"
    ev = np.ascontiguousarray(evals, dtype=np.uint64 if binpoly else np.int32)
    al = _i128_words([a for poly in alphas for a in poly]) if binpoly else np.zeros(2, np.int64)
    q = np.stack([_limbs4(v) for v in q1])
    out = np.zeros((num_rows, 4), dtype=np.uint64)
    rc = load().or_prove_b(ctypes.c_int(int(binpoly)), _u64(D), _u64(num_rows), _u64(row_len),
                           _u64(batch), _p(ev), _p(al), _p(q), _p(_limbs4(p)), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return [_int4(out[j]) for j in range(num_rows)]


def combined_row_i32(evals, coeffs, num_rows, row_len, batch=1, L=3):
    ev = np.ascontiguousarray(evals, dtype=np.int32)
    co = np.ascontiguousarray(coeffs, dtype=np.int64)
    out = np.zeros((row_len, L), dtype=np.uint64)
    rc = load().or_combined_row_i32(_u64(L), _u64(num_rows), _u64(row_len), _u64(batch), _p(ev),
                                    _p(co), _p(out))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return out


def merkle_prove(layers, n, leaf):
    levels = n.bit_length() - 1
    sib = np.zeros((max(1, levels), 32), dtype=np.uint8)
    count = ctypes.c_uint64()
    lay = np.ascontiguousarray(layers, dtype=np.uint8)
    rc = load().or_merkle_prove(_p(lay), _u64(n), _u64(leaf), _p(sib), ctypes.byref(count))
    if rc:
        raise ValueError(f"oracle rc {rc}")
    return [sib[i].tobytes() for i in range(count.value)]


def prove(ts, p, evals, num_rows, row_len, N, elem, cw, layers, point, batch=1, D=32, L=5,
          eval_type="binpoly", num_openings=147):
    """ZipPlus::prove_f (zip-plus/src/pcs/phase_prove.rs:126-253) on a
    transcript that already holds the commitment and the field config.
    Returns the evaluation (canonical int); the proof bytes are ts.stream."""
    binpoly = eval_type == "binpoly"
    nv = (num_rows * row_len).bit_length() - 1
    if len(point) != nv:
        raise ValueError("Invalid point")
    pt = [v % p for v in point]
    q0, q1 = point_to_tensor(num_rows, pt, p)
    alphas = [[ts.challenge_i128() for _ in range(D)] if binpoly and D > 1 else [1]
              for _ in range(batch)]
    b = prove_b(p, evals, alphas, q1, num_rows, row_len, batch, D, eval_type)
    ts.write_field_elements(p, b)
    ev = sum(x * y for x, y in zip(q0, b)) % p
    coeffs = [1] if num_rows == 1 else [ts.challenge_i128() for _ in range(num_rows)]
    cw_words = _i128_words(coeffs)
    if binpoly:
        comb = combined_row(evals, _i128_words([a for poly in alphas for a in poly]), cw_words,
                            num_rows, row_len, batch, D, L)
    else:
        comb = combined_row_i32(evals, cw_words, num_rows, row_len, batch, L)
    ts.stream += comb.tobytes()
    cwb = np.ascontiguousarray(cw, dtype=np.uint8).reshape(batch * num_rows, N, elem)
    levels = N.bit_length() - 1
    for _ in range(num_openings):
        idx = ts.squeeze_idx(N)
        ts.stream += cwb[:, idx, :].tobytes()
        ts.stream += np.array([idx, N, levels], dtype=np.uint64).tobytes()
        ts.stream += b"".join(merkle_prove(layers, N, idx))
    return ev


def _signed(words):
    v = 0
    for i, w in enumerate(words):
        v |= int(w) << (64 * i)
    n = 64 * len(words)
    return v - (1 << n) if v >> (n - 1) else v


def verify(ts, p, root, point, ev, encode_wide, num_rows, row_len, N, elem, batch=1, D=32, L=5,
           eval_type="binpoly", coeff_bytes=8, num_openings=147):
    """ZipPlus::verify (zip-plus/src/pcs/phase_verify.rs:100-305) on a
    verifier transcript (root absorbed, field config drawn, stream at the
    proof).  encode_wide(uint64 [row_len, L]) -> uint64 [N, L].  Raises
    ValueError naming the failed check."""
    binpoly = eval_type == "binpoly"
    alphas = [[ts.challenge_i128() for _ in range(D)] if binpoly and D > 1 else [1]
              for _ in range(batch)]
    pt = [v % p for v in point]
    q0, q1 = point_to_tensor(num_rows, pt, p)
    b = ts.read_field_elements(num_rows)
    if any(m != p for m, _ in b):
        raise ValueError("modulus mismatch")
    b = [v for _, v in b]
    if sum(x * y for x, y in zip(q0, b)) % p != ev % p:
        raise ValueError("Evaluation consistency failure")
    coeffs = [1] if num_rows == 1 else [ts.challenge_i128() for _ in range(num_rows)]
    comb = np.frombuffer(ts.read(row_len * L * 8), dtype=np.uint64).reshape(row_len, L)
    enc = encode_wide(comb)
    lhs = sum(_signed(comb[c]) * q1[c] for c in range(row_len)) % p
    rhs = sum(s * x for s, x in zip(coeffs, b)) % p
    if lhs != rhs:
        raise ValueError("Coherence failure")
    mask = (1 << (64 * L)) - 1
    levels = N.bit_length() - 1
    for _ in range(num_openings):
        idx = ts.squeeze_idx(N)
        col = ts.read(batch * num_rows * elem)
        leaf, count, nsib = np.frombuffer(ts.read(24), dtype=np.uint64).tolist()
        sib = [ts.read(32) for _ in range(nsib)]
        if leaf != idx or count != N or nsib != levels:
            raise ValueError("Merkle proof header mismatch")
        acc = 0
        for pb in range(batch):
            for j in range(num_rows):
                e = col[(pb * num_rows + j) * elem:(pb * num_rows + j + 1) * elem]
                ncoef = elem // coeff_bytes
                coefs = [int.from_bytes(e[k * coeff_bytes:(k + 1) * coeff_bytes], "little",
                                        signed=True) for k in range(ncoef)]
                acc += coeffs[j] * sum(c * a for c, a in zip(coefs, alphas[pb]))
        if acc & mask != _signed(enc[idx]) & mask:
            raise ValueError("Proximity failure")
        if not merkle_verify(root, col, idx, N, sib):
            raise ValueError("Column opening verification failed")
    if ts.pos != len(ts.stream):
        raise ValueError("trailing proof bytes")
    return True


# ---------------------------------------------------------------------------
# i64 and DensePolynomial<i64, D> evaluations (protocol/src/lib.rs:490-559):
# the code applied to each coefficient plane in the codeword type
# (Int<L> arithmetic mod 2^(64L)): IPRS i64 -> i128 (L = 2), IPRS
# DensePolynomial<i64, D> -> [i64; D] (L = 1), RAA i64 -> i128 (L = 2), RAA
# DensePolynomial<i64, D> -> [Int<4>; D] (L = 4).
# ---------------------------------------------------------------------------

def i64_plane_limbs(kind, eval_type):
    if eval_type == "i64":
        return 2
    return 1 if kind == "iprs" else 4


def encode_i64_planes(kind, evals, num_rows, row_len, W, L, depth=-1, rep=4):
    """evals: int64 [num_rows, row_len, W] -> uint64 [num_rows, N, W, L]."""
    x = np.ascontiguousarray(evals, dtype=np.int64).reshape(num_rows, row_len, W)
    N = row_len * rep
    out = np.zeros((num_rows, N, W, L), dtype=np.uint64)
    for c in range(W):
        limbs = sign_extend_limbs(x[:, :, c:c + 1], L)
        if kind == "iprs":
            out[:, :, c, :] = iprs_encode_wide(limbs, row_len, depth, L, rep)
        else:
            out[:, :, c, :] = raa_encode_wide(limbs, row_len, L, rep,
                                              sign_flip=kind == "raa_signflip")
    return out


def commit_i64(kind, evals, num_rows, row_len, batch=1, D=1, eval_type="i64", depth=-1, rep=4):
    """ZipPlus::commit of i64 / DensePolynomial<i64, D> polys -> dict(root, cw, layers)."""
    W = 1 if eval_type == "i64" else D
    L = i64_plane_limbs(kind, eval_type)
    cw = encode_i64_planes(kind, evals, batch * num_rows, row_len, W, L, depth, rep)
    N = row_len * rep
    elem = W * L * 8
    layers = merkle_layers(merkle_leaves(cw.view(np.uint8).reshape(-1), batch * num_rows, N, elem))
    return {"root": layers[-1].tobytes(), "cw": cw, "layers": layers, "elem": elem}
