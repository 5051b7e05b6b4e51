"""zipgpu -- the Zip/Zip+ commitment hot path on MI355X (gfx950).

A host-side mirror of the reference crate zip-plus over the C ABI of
include/zipgpu.h.  Names, argument meaning and error behaviour follow the
reference:

* ``IprsCode(row_len, depth)`` / ``IprsCode.new_with_optimal_depth(row_len)``
  -- zip-plus/src/code/iprs.rs:30-51
* ``RaaCode(row_len)`` -- zip-plus/src/code/raa.rs:38-90;
  ``RaaSignFlippingCode(row_len)`` -- zip-plus/src/code/raa_sign_flip.rs
* ``LinearCode.encode / row_len / codeword_len / params_string``
  -- zip-plus/src/code.rs:10-72
* ``ZipPlus.setup / commit / commit_single / encode_rows``
  -- zip-plus/src/pcs/structs.rs:74-88, pcs/phase_commit.rs:61-176
* ``ZipPlus.combined_row`` -- the combined-row block of ``prove_f``,
  zip-plus/src/pcs/phase_prove.rs:153-244
* ``MerkleTree`` root / ``prove`` / ``MerkleProof`` -- zip-plus/src/merkle.rs
* ``BatchCommit`` -- many independent ``ZipPlus::commit`` calls (as the
  reference's callers issue them, protocol/src/prover.rs:710-726) run as one
  device plan, optionally one HIP graph; ``LinearCode.autotune`` picks the
  fastest kernel variant for a shape.  Neither has a reference counterpart.

Device memory and streams come from torch (ROCm build); every computation
runs in libzipgpu.so.  ``ZipError`` subclasses stand in for the reference's
``Result::Err(ZipError::InvalidPcsParam)`` and its assert panics.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

__all__ = [
    "ZipError", "InvalidPcsParam", "ShapeError", "DeviceError", "OutOfMemory", "Overflow",
    "ZipTypes", "BinPolyTypes", "IntTypes", "I64Types", "DensePolyTypes", "IprsCode", "RaaCode", "RaaSignFlippingCode",
    "ZipPlus", "ZipPlusParams", "ZipPlusHint", "ZipPlusCommitment", "MerkleProof",
    "merkle_root", "lib", "set_option", "BatchCommit", "CommitServer", "PcsProverTranscript",
    "step0_commit", "validate_input",
]


class ZipError(Exception):
    """A non-zero zp_status from the library."""

    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status
        self.message = message


class InvalidPcsParam(ZipError):
    """ZipError::InvalidPcsParam (zip-plus/src/lib.rs:13-29)."""


class ShapeError(ZipError):
    """The reference's assert! panics on shapes."""


class DeviceError(ZipError):
    """A HIP runtime failure."""


class OutOfMemory(ZipError):
    """Device allocation failed."""


class Overflow(ZipError):
    """A checked code's encoding left its Cw type: where the reference's
    checked build panics ("Multiplication by twiddle should not overflow",
    add!'s "Addition overflow", "Accumulation overflow", neg!)."""


_ERRORS = {L.ZP_ERR_INVALID_PARAM: InvalidPcsParam, L.ZP_ERR_SHAPE: ShapeError,
           L.ZP_ERR_DEVICE: DeviceError, L.ZP_ERR_OOM: OutOfMemory,
           L.ZP_ERR_OVERFLOW: Overflow}


def lib():
    return L.load()


def set_option(name, value):
    """zp_set_option: library-wide kernel-variant overrides, -1 = automatic.

    "iprs_encoder" (0 matrix-core with r-block rows spread, 1 VALU, 2
    matrix-core with consecutive rows), "raa_kernel" (0 full-row,
    1 / 2 plane groups), "commit_overlap" (1 overlap encode and hash, 0 not:
    the default), "wide_base" (Int<L>, i64 and i32 IPRS base layer: 0
    matrix cores where the code has the int8 tables, 1 per-element), "merkle_tree"
    (0 a 4-lane quad per parent node, 1 a thread per parent), "chunk_tree" (0
    register folds and quad merges for 4 to 2048 chunks per column, 1 the LDS
    tree for every count), "combined_row" (0 slice sums on the int8 matrix
    cores, 1 64-bit VALU sums).
    """
    _check(L.load().zp_set_option(name.encode(), int(value)))


class EncodeTiming:
    """zp_encode_timing_start/stop as a context: the encoder's device time
    inside one zp_commit in `every` run in the block (the first, the
    every+1-th, ...), up to `max_commits` of them, by HIP events on the
    commit stream (created before the block).  After the block, `launches`
    is the number of timed commits, `total_ms` the sum of their whole
    encoder spans and `last_ms` the sum of their last-kernel spans (the
    matrix-core IPRS encoder after its bit-plane pre-pass); `mean_ms` and
    `mean_last_ms` are the averages."""

    def __init__(self, max_commits=1024, every=1):
        self.max_commits, self.every = int(max_commits), int(every)
        self.total_ms, self.last_ms, self.launches = 0.0, 0.0, 0

    def __enter__(self):
        _check(L.load().zp_encode_timing_start(self.max_commits, self.every))
        return self

    def __exit__(self, *exc):
        ms, last, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        _check(L.load().zp_encode_timing_stop(ctypes.byref(ms), ctypes.byref(last),
                                              ctypes.byref(n)))
        self.total_ms, self.last_ms, self.launches = ms.value, last.value, n.value
        return False

    @property
    def mean_ms(self):
        return self.total_ms / self.launches if self.launches else float("nan")

    @property
    def mean_last_ms(self):
        return self.last_ms / self.launches if self.launches else float("nan")


def _check(rc):
    if rc != L.ZP_OK:
        raise _ERRORS.get(rc, ZipError)(rc, L.last_error())


def validate_input(rep, row_len, types, batch_size):
    """validate_input's bit budget (zip-plus/src/pcs/utils.rs:24-54), the
    assert ZipPlus::commit runs first (phase_commit.rs:70-77): raises
    ShapeError where the reference panics.  Host only (zp_validate_input)."""
    _check(L.load().zp_validate_input(int(rep), int(row_len), types.code,
                                      types.degree_plus_one, types.comb_limbs,
                                      int(batch_size)))


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dev_ptr(t):
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


class ZipTypes:
    """The Eval / Cw / CombR choice of a ZipTypes impl (pcs/structs.rs:15-60).

    ``eval_type`` is "binpoly" (Eval = BinaryPoly<D>, u64 bit-packed, Cw =
    [i64;D] under IPRS and [i32;D] under RAA), "i32" (Eval = i32, Cw =
    i128), "i64" (Eval = i64, Cw = i128) or "dense_i64" (Eval =
    DensePolynomial<i64, D>, Cw = [i64;D] under IPRS and [Int<4>;D] under
    RAA).  ``comb_limbs`` is L of CombR = Int<L>.
    """

    _CODES = {"binpoly": L.ZP_EVAL_BINPOLY, "i32": L.ZP_EVAL_I32, "i64": L.ZP_EVAL_I64,
              "dense_i64": L.ZP_EVAL_DENSE_I64}

    def __init__(self, eval_type, degree_plus_one=1, comb_limbs=5, num_column_openings=147):
        if eval_type not in self._CODES:
            raise ValueError(eval_type)
        self.eval_type = eval_type
        self.degree_plus_one = degree_plus_one if eval_type in ("binpoly", "dense_i64") else 1
        self.comb_limbs = comb_limbs
        self.num_column_openings = num_column_openings

    @property
    def code(self):
        return self._CODES[self.eval_type]

    @property
    def eval_dtype(self):
        return torch.int32 if self.eval_type == "i32" else torch.int64

    @property
    def eval_words(self):
        """Elements of eval_dtype per evaluation (D coefficients for dense_i64)."""
        return self.degree_plus_one if self.eval_type == "dense_i64" else 1


def BinPolyTypes(degree_plus_one, comb_limbs=5):
    """BenchZipPlusTypes<_, D> (zip-plus/benches/zip_plus_benches.rs:37-65)."""
    return ZipTypes("binpoly", degree_plus_one, comb_limbs)


def IntTypes(comb_limbs=3):
    """BenchZipTypes (zip-plus/benches/zip_benches.rs:25-40): i32 -> i128."""
    return ZipTypes("i32", 1, comb_limbs)


def I64Types(comb_limbs=8):
    """IntZipTypes of the protocol (protocol/src/lib.rs:546-559): i64 -> i128."""
    return ZipTypes("i64", 1, comb_limbs)


def DensePolyTypes(degree_plus_one, comb_limbs=8):
    """ArbitraryPolyZipTypesIprs / ...Raa (protocol/src/lib.rs:490-541):
    DensePolynomial<i64, D> -> [i64; D] (IPRS) or [Int<4>; D] (RAA)."""
    return ZipTypes("dense_i64", degree_plus_one, comb_limbs)


class LinearCode:
    """A code with device-resident tables (trait LinearCode, code.rs:10-72)."""

    KIND = None

    def __init__(self, kind, rep, row_len, depth, checked=False):
        self._h = None
        lib_ = L.load()
        h = ctypes.c_void_p()
        _check(lib_.zp_code_create(kind, rep, row_len, depth, ctypes.byref(h)))
        self._h = h
        self._lib = lib_
        self.REPETITION_FACTOR = rep
        if checked:
            _check(lib_.zp_code_set_checked(h, 1))

    @property
    def checked(self):
        """The CHECK const (IprsCode<.., CHECK>, RaaConfig::CHECK_FOR_OVERFLOWS):
        a checked code's encodings raise Overflow where the reference panics."""
        return bool(self._lib.zp_code_is_checked(self._h))

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zp_code_destroy(self._h)
            self._h = None

    def row_len(self):
        return int(self._lib.zp_code_row_len(self._h))

    def codeword_len(self):
        return int(self._lib.zp_code_codeword_len(self._h))

    def cw_shape(self, types):
        """Trailing shape and dtype of one codeword element."""
        if types.eval_type in ("i32", "i64"):
            return (2,), torch.int64  # i128 as two LE words
        if types.eval_type == "dense_i64":
            if self.KIND == L.ZP_CODE_IPRS:
                return (types.degree_plus_one,), torch.int64
            return (types.degree_plus_one, 4), torch.int64  # Int<4> limbs
        if self.KIND == L.ZP_CODE_IPRS:
            return (types.degree_plus_one,), torch.int64
        return (types.degree_plus_one,), torch.int32

    def cw_elem_bytes(self, types):
        return int(self._lib.zp_cw_elem_bytes(self._h, types.code, types.degree_plus_one))

    def encode_rows(self, evals, num_rows, types, out=None, stream=None):
        """Encode num_rows consecutive rows of a flat device eval tensor."""
        if evals.dtype != types.eval_dtype:
            raise ValueError(f"evals must be {types.eval_dtype}")
        if evals.numel() != num_rows * self.row_len() * types.eval_words:
            raise ShapeError(L.ZP_ERR_SHAPE,
                             f"Input length {evals.numel()} does not match "
                             f"{num_rows} rows of {self.row_len()}")
        tail, dtype = self.cw_shape(types)
        if out is None:
            out = torch.empty((num_rows, self.codeword_len()) + tail, dtype=dtype,
                              device=evals.device)
        _check(self._lib.zp_encode_rows(self._h, types.code, types.degree_plus_one,
                                        _dev_ptr(evals), num_rows, _dev_ptr(out),
                                        _stream(stream)))
        return out

    def encode_wide(self, rows, stream=None):
        """LinearCode::encode_wide (code.rs:55) of rows of CombR = Int<L>.

        rows: device int64 [num_rows, row_len, L] or [row_len, L] (u64 limbs,
        little-endian, two's complement); returns [.., codeword_len, L].
        """
        if rows.dtype != torch.int64:
            raise ValueError("CombR rows must be int64 limbs")
        single = rows.dim() == 2
        r = rows.unsqueeze(0) if single else rows
        if r.dim() != 3 or r.shape[1] != self.row_len():
            raise ShapeError(L.ZP_ERR_SHAPE,
                             f"Input length {r.shape[1] if r.dim() == 3 else -1} does not match "
                             f"expected row length {self.row_len()}")
        r = r.contiguous()
        out = torch.empty((r.shape[0], self.codeword_len(), r.shape[2]), dtype=torch.int64,
                          device=r.device)
        _check(self._lib.zp_encode_wide(self._h, r.shape[2], _dev_ptr(r), r.shape[0],
                                        _dev_ptr(out), _stream(stream)))
        return out[0] if single else out

    def autotune(self, types, num_rows, batch=1, stream=None):
        """zp_autotune: time this code's kernel variants at a shape, keep the fastest."""
        _check(self._lib.zp_autotune(self._h, types.code, types.degree_plus_one, num_rows, batch,
                                     _stream(stream)))
        return self.tuning()

    def tuning(self):
        """The tuned variants (-1 = not tuned), numbered as in set_option."""
        v = [ctypes.c_int32() for _ in range(4)]
        _check(self._lib.zp_code_tuning(self._h, *[ctypes.byref(x) for x in v]))
        return {"iprs_encoder": v[0].value, "raa_kernel": v[1].value,
                "commit_overlap": v[2].value, "chunk_tree_fold": v[3].value}

    def encode(self, row, types):
        """LinearCode::encode of one row."""
        if row.numel() != self.row_len():
            raise ShapeError(L.ZP_ERR_SHAPE,
                             f"Input length {row.numel()} does not match expected row length "
                             f"{self.row_len()}")
        return self.encode_rows(row.reshape(-1), 1, types)[0]


class IprsCode(LinearCode):
    """IprsCode<Zt, PnttConfigF65537, REP, _> (zip-plus/src/code/iprs.rs)."""

    KIND = L.ZP_CODE_IPRS

    def __init__(self, row_len, depth, rep=4, checked=False):
        super().__init__(L.ZP_CODE_IPRS, rep, row_len, depth, checked)
        self.depth = int(self._lib.zp_code_depth(self._h))

    @classmethod
    def new_with_optimal_depth(cls, row_len, rep=4, checked=False):
        return cls(row_len, -1, rep, checked)

    def params_string(self):
        return f"row_len={self.row_len()}, rate=1/{self.REPETITION_FACTOR}, depth={self.depth}"


class RaaCode(LinearCode):
    """RaaCode<Zt, _, REP> with permutation seeds 1 and 2 (raa.rs:38-90)."""

    KIND = L.ZP_CODE_RAA

    def __init__(self, row_len, rep=4, checked=False):
        super().__init__(self.KIND, rep, row_len, 0, checked)

    def params_string(self):
        return f"row_len={self.row_len()}, rate=1/{self.REPETITION_FACTOR}"


class RaaSignFlippingCode(RaaCode):
    """RaaSignFlippingCode (zip-plus/src/code/raa_sign_flip.rs)."""

    KIND = L.ZP_CODE_RAA_SIGNFLIP


class ZipPlusParams:
    """ZipPlusParams (pcs/structs.rs:91-109)."""

    def __init__(self, num_vars, num_rows, linear_code):
        self.num_vars = num_vars
        self.num_rows = num_rows
        self.linear_code = linear_code


class ZipPlusCommitment:
    """ZipPlusCommitment { root, batch_size } (pcs/structs.rs:131-157)."""

    def __init__(self, root, batch_size):
        self.root = root
        self.batch_size = batch_size

    def __eq__(self, other):
        return self.root == other.root and self.batch_size == other.batch_size

    def to_bytes(self):
        """Transcription bytes: root then batch_size as u64 LE (structs.rs:140-153)."""
        return self.root + int(self.batch_size).to_bytes(8, "little")


class MerkleProof:
    """MerkleProof { leaf_index, leaf_count, siblings } (merkle.rs:259-278)."""

    def __init__(self, leaf_index, leaf_count, siblings):
        self.leaf_index = leaf_index
        self.leaf_count = leaf_count
        self.siblings = siblings


class ZipPlusHint:
    """ZipPlusHint kept on the device: codewords plus every Merkle layer."""

    def __init__(self, handle, types, code, num_rows, owner=None):
        self._owner = owner  # a BatchCommit whose pool holds the buffers
        self._h = handle
        self._lib = L.load()
        self.types = types
        self.code = code
        self.num_rows = num_rows
        per_poly, batch, leaves, height = ctypes.c_uint64(), ctypes.c_uint32(), \
            ctypes.c_uint64(), ctypes.c_uint32()
        _check(self._lib.zp_hint_info(handle, ctypes.byref(per_poly), ctypes.byref(batch),
                                      ctypes.byref(leaves), ctypes.byref(height)))
        self.cw_bytes_per_poly = per_poly.value
        self.batch = batch.value
        self.num_leaves = leaves.value
        self.height = height.value

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None):
            self._lib.zp_hint_free(self._h)
            self._h = None

    def _cw_np(self):
        tail, dtype = self.code.cw_shape(self.types)
        return (self.num_rows, self.code.codeword_len()) + tail, \
            np.int64 if dtype == torch.int64 else np.int32

    def codewords(self, poly=0, stream=None):
        """The poly's DenseRowMatrix<Cw> as a host array [rows, N, ...]."""
        shape, dt = self._cw_np()
        out = np.empty(shape, dtype=dt)
        _check(self._lib.zp_hint_copy_codewords(self._h, poly, out.ctypes.data_as(ctypes.c_void_p),
                                                _stream(stream)))
        return out

    def device_codewords_ptr(self):
        """The device address of the codewords (zp_hint_codewords)."""
        return int(self._lib.zp_hint_codewords(self._h) or 0)

    def record_stream(self, stream):
        """zp_hint_record_stream: work on `stream` reads the hint, so its
        memory is not handed out again until that work has finished."""
        _check(self._lib.zp_hint_record_stream(self._h, ctypes.c_void_p(stream.cuda_stream)))

    def layer(self, level, stream=None):
        n = self.num_leaves >> level
        out = np.empty((n, 32), dtype=np.uint8)
        _check(self._lib.zp_hint_copy_layer(self._h, level, out.ctypes.data_as(ctypes.c_void_p),
                                            _stream(stream)))
        return out

    def prove(self, leaf_index, stream=None):
        """MerkleTree::prove (merkle.rs:105-120)."""
        buf = np.empty((64, 32), dtype=np.uint8)
        n = ctypes.c_uint32()
        _check(self._lib.zp_merkle_prove(self._h, leaf_index, buf.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.byref(n), _stream(stream)))
        return MerkleProof(leaf_index, self.num_leaves, [bytes(buf[i]) for i in range(n.value)])

    def open_column(self, col, stream=None):
        """The bytes of column `col` over all rows of all polys."""
        nbytes = self.batch * self.num_rows * self.code.cw_elem_bytes(self.types)
        out = np.empty(nbytes, dtype=np.uint8)
        _check(self._lib.zp_hint_open_column(self._h, col, out.ctypes.data_as(ctypes.c_void_p),
                                             _stream(stream)))
        return out.tobytes()


def merkle_root(rows, elem_bytes, stream=None):
    """MerkleTree::new(rows).root() over a device tensor [num_rows, width, ...]."""
    if rows.dim() < 2:
        raise ShapeError(L.ZP_ERR_SHAPE, "expected [num_rows, width, ...]")
    out = (ctypes.c_uint8 * 32)()
    _check(L.load().zp_merkle_root(_dev_ptr(rows), rows.shape[0], rows.shape[1], elem_bytes, out,
                                   _stream(stream)))
    return bytes(out)


class PcsProverTranscript:
    """PcsProverTranscript (zip-plus/src/pcs_transcript.rs:40-221): the
    Fiat-Shamir Blake3Transcript (transcript/src/lib.rs:10-72) and the proof
    byte stream, held by the library."""

    def __init__(self, _handle=None):
        self._lib = L.load()
        if _handle is None:
            h = ctypes.c_void_p()
            _check(self._lib.zp_transcript_create(ctypes.byref(h)))
            _handle = h
        self._h = _handle

    @classmethod
    def new_from_commitment(cls, comm):
        return cls.new_from_commitments([comm])

    @classmethod
    def new_from_commitments(cls, comms):
        t = cls()
        for c in comms:
            t.absorb_slice(c.root)
        return t

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zp_transcript_destroy(self._h)
            self._h = None

    def clone(self):
        h = ctypes.c_void_p()
        _check(self._lib.zp_transcript_clone(self._h, ctypes.byref(h)))
        return PcsProverTranscript(h)

    def absorb_slice(self, data):
        data = bytes(data)
        buf = (ctypes.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
        _check(self._lib.zp_transcript_absorb_slice(self._h, buf, len(data)))

    def absorb_entries(self, data, entry_bytes, stream=None):
        """absorb_slice of each consecutive entry of entry_bytes bytes.

        `data` is bytes-like, or a device tensor: then its full BLAKE3 chunks
        are hashed on the GPU (zp_transcript_absorb_entries_device), on
        `stream` or the current stream."""
        if isinstance(data, torch.Tensor) and data.is_cuda:
            t = data.contiguous()
            nbytes = t.numel() * t.element_size()
            count = nbytes // entry_bytes if entry_bytes else 0
            if count * entry_bytes != nbytes:
                raise ShapeError(L.ZP_ERR_SHAPE, "data is not a whole number of entries")
            _check(self._lib.zp_transcript_absorb_entries_device(
                self._h, ctypes.c_void_p(t.data_ptr()), entry_bytes, count, _stream(stream)))
            return
        data = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8))
        count = len(data) // entry_bytes if entry_bytes else 0
        if count * entry_bytes != len(data):
            raise ShapeError(L.ZP_ERR_SHAPE, "data is not a whole number of entries")
        _check(self._lib.zp_transcript_absorb_entries(
            self._h, data.ctypes.data_as(ctypes.c_void_p), entry_bytes, count))

    def get_challenge(self, num_bytes):
        """get_challenge::<T> for T of num_bytes bytes: its little-endian bytes."""
        buf = (ctypes.c_uint8 * max(1, num_bytes))()
        _check(self._lib.zp_transcript_challenge(self._h, buf, num_bytes))
        return bytes(buf)[:num_bytes]

    def get_challenge_i128(self):
        return int.from_bytes(self.get_challenge(16), "little", signed=True)

    def squeeze_challenge_idx(self, cap):
        out = ctypes.c_uint64()
        _check(self._lib.zp_transcript_squeeze_idx(self._h, cap, ctypes.byref(out)))
        return out.value

    def get_random_field_cfg(self):
        """get_random_field_cfg::<MontyField<4>, Uint<4>, MillerRabin>: the modulus."""
        buf = (ctypes.c_uint8 * 32)()
        _check(self._lib.zp_transcript_random_field_cfg(self._h, buf))
        return int.from_bytes(bytes(buf), "little")

    def write(self, data):
        data = bytes(data)
        buf = (ctypes.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
        _check(self._lib.zp_transcript_write(self._h, buf, len(data)))

    def stream(self):
        n = self._lib.zp_transcript_stream_len(self._h)
        buf = (ctypes.c_uint8 * max(1, n))()
        _check(self._lib.zp_transcript_stream(self._h, buf, n))
        return bytes(buf)[:n]


def step0_commit(params, witness, public):
    """ZincPlusPiop::step0_commit (protocol/src/prover.rs:218-262).

    params: (pp, types) for the binary, arbitrary and int traces;
    witness / public: three lists of device tensors (columns, each one poly
    of that trace's ZipTypes).  Each witness trace is committed as one batch
    (commit_optionally, prover.rs:710-726: an empty trace gives no hint and
    a zero root), the transcript starts from the three roots
    (PcsProverTranscript::new_from_commitments) and absorbs every public
    entry's transcription bytes (absorb_public_columns,
    protocol/src/lib.rs:305-316).  -> (hints, commitments, transcript)."""
    hints, comms = [], []
    for (pp, types), cols in zip(params, witness):
        if len(cols) == 0:
            hints.append(None)
            comms.append(ZipPlusCommitment(bytes(32), 0))
            continue
        h, c = ZipPlus(types).commit(pp, cols)
        hints.append(h)
        comms.append(c)
    t = PcsProverTranscript.new_from_commitments(comms)
    for (pp, types), cols in zip(params, public):
        entry = 8 * types.eval_words if types.eval_type != "i32" else 4
        for col in cols:
            t.absorb_entries(col, entry)
    return hints, comms, t


class ZipPlus:
    """ZipPlus<Zt, Lc> (pcs/structs.rs:66-89, phase_commit.rs, phase_prove.rs)."""

    def __init__(self, types):
        self.types = types

    @staticmethod
    def setup(poly_size, linear_code):
        """ZipPlus::setup (pcs/structs.rs:74-88)."""
        if poly_size <= 0 or poly_size & (poly_size - 1):
            raise ShapeError(L.ZP_ERR_SHAPE, "poly_size must be a power of two")
        row_len = linear_code.row_len()
        if row_len <= 0 or poly_size % row_len:
            raise ShapeError(L.ZP_ERR_SHAPE,
                             f"poly_size ({poly_size}) must be divisible by row_len ({row_len})")
        num_rows = poly_size // row_len
        if num_rows & (num_rows - 1):
            raise ShapeError(L.ZP_ERR_SHAPE, f"num_rows ({num_rows}) must be a power of two")
        return ZipPlusParams(poly_size.bit_length() - 1, num_rows, linear_code)

    def _stack(self, pp, polys):
        if isinstance(polys, torch.Tensor):
            t = polys if polys.dim() == 2 else polys.reshape(1, -1)
        else:
            if len(polys) == 0:
                raise ShapeError(L.ZP_ERR_SHAPE, "Batch must contain at least one polynomial")
            # One poly: a view (no device copy); several: one contiguous batch.
            t = (polys[0].reshape(1, -1) if len(polys) == 1
                 else torch.stack([p.reshape(-1) for p in polys]))
        if t.shape[0] == 0:
            raise ShapeError(L.ZP_ERR_SHAPE, "Batch must contain at least one polynomial")
        expected = pp.num_rows * pp.linear_code.row_len() * self.types.eval_words
        if t.shape[1] != expected:
            raise ShapeError(L.ZP_ERR_SHAPE,
                             f"Polynomial has an incorrect number of evaluations ({t.shape[1]}) "
                             f"for the expected matrix size ({expected})")
        if t.dtype != self.types.eval_dtype:
            raise ValueError(f"evals must be {self.types.eval_dtype}")
        return t.contiguous()

    def encode_rows(self, pp, evals, stream=None):
        """ZipPlus::encode_rows (phase_commit.rs:159-176)."""
        return pp.linear_code.encode_rows(evals.reshape(-1), pp.num_rows, self.types,
                                          stream=stream)

    def commit(self, pp, polys, stream=None):
        """ZipPlus::commit (phase_commit.rs:61-101) -> (hint, commitment).

        Device tensors commit in place (zp_commit); host tensors, as the
        reference's callers hold them, go through zp_commit_host, which
        overlaps their upload with the encoding."""
        t = self._stack(pp, polys)
        validate_input(pp.linear_code.REPETITION_FACTOR, pp.linear_code.row_len(), self.types,
                       t.shape[0])
        h = ctypes.c_void_p()
        root = (ctypes.c_uint8 * 32)()
        if t.is_cuda:
            fn, ptr = L.load().zp_commit, _dev_ptr(t)
        else:
            fn, ptr = L.load().zp_commit_host, ctypes.c_void_p(t.data_ptr())
        _check(fn(pp.linear_code._h, self.types.code, self.types.degree_plus_one, ptr,
                  pp.num_rows, t.shape[0], ctypes.byref(h), root, _stream(stream)))
        hint = ZipPlusHint(h, self.types, pp.linear_code, pp.num_rows)
        return hint, ZipPlusCommitment(bytes(root), t.shape[0])

    def commit_single(self, pp, poly, stream=None):
        return self.commit(pp, [poly], stream)

    def prove(self, transcript, pp, polys, point, hint, modulus, stream=None):
        """ZipPlus::prove::<MontyField<4>, _> (phase_prove.rs:85-253).

        point: num_vars Python ints (Pt = i128); modulus: the field config
        drawn from the transcript.  Appends the proof to the transcript and
        returns the evaluation <q_0, b> as a canonical int."""
        t = self._stack(pp, polys)
        pt = np.zeros((max(1, len(point)), 2), dtype=np.int64)
        for i, v in enumerate(point):
            if not -(1 << 127) <= v < (1 << 127):
                raise ValueError("point coordinates are i128")
            u = v & ((1 << 128) - 1)
            pt[i, 0] = np.uint64(u & ((1 << 64) - 1)).view(np.int64)
            pt[i, 1] = np.uint64(u >> 64).view(np.int64)
        mod = (ctypes.c_uint8 * 32).from_buffer_copy(int(modulus).to_bytes(32, "little"))
        ev = (ctypes.c_uint8 * 32)()
        _check(L.load().zp_prove(pp.linear_code._h, self.types.code, self.types.degree_plus_one,
                                 self.types.comb_limbs, _dev_ptr(t), pp.num_rows, t.shape[0],
                                 hint._h, pt.ctypes.data_as(ctypes.c_void_p), len(point), mod,
                                 self.types.num_column_openings, transcript._h, ev,
                                 _stream(stream)))
        return int.from_bytes(bytes(ev), "little")

    def combined_row(self, pp, polys, alphas, coeffs, stream=None):
        """The combined row of prove_f (phase_prove.rs:153-244).

        alphas: device int64 [batch, D, 2] (i128 LE words), coeffs: device
        int64 [num_rows, 2]; returns int64 [row_len, L] (Int<L> limbs).
        """
        if self.types.eval_type != "binpoly":
            raise InvalidPcsParam(L.ZP_ERR_INVALID_PARAM, "combined row needs binary polys")
        t = self._stack(pp, polys)
        D, Lm = self.types.degree_plus_one, self.types.comb_limbs
        row_len = pp.linear_code.row_len()
        if alphas.numel() != t.shape[0] * D * 2 or coeffs.numel() != pp.num_rows * 2:
            raise ShapeError(L.ZP_ERR_SHAPE, "alphas/coeffs have the wrong length")
        out = torch.empty((row_len, Lm), dtype=torch.int64, device=t.device)
        al, co = alphas.contiguous(), coeffs.contiguous()
        _check(L.load().zp_combined_row(D, Lm, pp.num_rows, row_len, t.shape[0], _dev_ptr(t),
                                        _dev_ptr(al), _dev_ptr(co), _dev_ptr(out),
                                        _stream(stream)))
        # The call is stream-ordered: tensors made here (a stacked batch,
        # contiguous copies) stay reserved until `stream` has used them.
        if stream is not None:
            for x in (t, al, co, out):
                x.record_stream(stream)
        return out


class BatchCommit:
    """A plan of independent commitments run together (zp_batch_*).

    ``items`` is a list of ``(pp, types)`` or ``(pp, types, batch)``: each
    item is ``ZipPlus(types).commit(pp, polys)`` for ``batch`` polys (default
    1, i.e. commit_single).  ``graph=True`` replays one HIP graph of the whole
    batch from the second commit on; ``autotune=True`` tunes each distinct
    code once.  Hints returned by ``hint(i)`` are valid until the next
    ``commit`` of the plan.
    """

    def __init__(self, items, graph=True, autotune=False, n_streams=0):
        self._h = None
        self._lib = L.load()
        self.items = []
        descs = (L.CommitDesc * len(items))()
        for i, it in enumerate(items):
            pp, types = it[0], it[1]
            batch = it[2] if len(it) > 2 else 1
            self.items.append((pp, types, batch))
            descs[i] = L.CommitDesc(pp.linear_code._h.value, types.code, types.degree_plus_one,
                                    pp.num_rows, batch)
        flags = (L.ZP_BATCH_GRAPH if graph else 0) | (L.ZP_BATCH_AUTOTUNE if autotune else 0)
        h = ctypes.c_void_p()
        _check(self._lib.zp_batch_create(descs, len(items), flags, n_streams, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.zp_batch_destroy(self._h)
            self._h = None

    def _check_input(self, i, t):
        pp, types, batch = self.items[i]
        n = batch * pp.num_rows * pp.linear_code.row_len() * types.eval_words
        if t.dtype != types.eval_dtype or t.numel() != n:
            raise ShapeError(L.ZP_ERR_SHAPE, f"batch item {i}: expected {n} evals of "
                                             f"{types.eval_dtype}")
        return _dev_ptr(t).value

    def commit(self, evals, stream=None):
        """Commit every item; evals[i] holds item i's polys. -> [ZipPlusCommitment]."""
        if len(evals) != len(self.items):
            raise ShapeError(L.ZP_ERR_SHAPE, "one input per batch item")
        ptrs = (ctypes.c_void_p * len(evals))(*[self._check_input(i, t)
                                                for i, t in enumerate(evals)])
        roots = (ctypes.c_uint8 * (32 * len(evals)))()
        _check(self._lib.zp_batch_commit(self._h, ptrs, roots, _stream(stream)))
        raw = bytes(roots)
        return [ZipPlusCommitment(raw[32 * i:32 * i + 32], self.items[i][2])
                for i in range(len(self.items))]

    def hint(self, i):
        pp, types, _ = self.items[i]
        h = self._lib.zp_batch_hint(self._h, i)
        if not h:
            _check(L.ZP_ERR_INVALID_PARAM)
        return ZipPlusHint(ctypes.c_void_p(h), types, pp.linear_code, pp.num_rows, owner=self)

    def info(self):
        hb, sb, g = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint32()
        _check(self._lib.zp_batch_info(self._h, ctypes.byref(hb), ctypes.byref(sb),
                                       ctypes.byref(g)))
        return {"hint_bytes": hb.value, "scratch_bytes": sb.value, "graph_ready": bool(g.value)}

    @property
    def schedule(self):
        """(grouped, streams): whether the leaf hashing of all items runs in
        grouped launches after the encoders, and the branch streams in use."""
        g, k = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.zp_batch_schedule(self._h, ctypes.byref(g), ctypes.byref(k)))
        return bool(g.value), k.value

    def set_schedule(self, grouped, streams):
        """Fixes the schedule; the next commit re-measures its scratch."""
        _check(self._lib.zp_batch_set_schedule(self._h, 1 if grouped else 0, streams))


class CommitServer:
    """Continuous-batching commit/open service on one device (zp_server_*).

    ``commit(pp, types, evals)`` queues ZipPlus::commit of the polys in
    ``evals`` (a host numpy array, copied to the device by the server, or a
    device tensor) and returns a request id at once.  When the request's
    commit event has arrived, ``open(req, ...)`` queues its opening: the
    combined row for (alphas, coeffs) and the listed columns with their
    Merkle paths.  ``poll`` returns finished jobs as dicts.  Inputs and output
    buffers are kept alive here until their job's event has been polled.
    """

    def __init__(self, hbm_budget=0, slots=4):
        self._h = None
        self._lib = L.load()
        cfg = L.ServerConfig(hbm_budget, slots)
        h = ctypes.c_void_p()
        _check(self._lib.zp_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._keep = {}     # request -> (pp, types, batch, input)
        self._opening = {}  # request -> output buffers

    def close(self):
        if getattr(self, "_h", None):
            self._lib.zp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def commit(self, pp, types, evals, batch=1):
        n = batch * pp.num_rows * pp.linear_code.row_len() * types.eval_words
        desc = L.CommitDesc(pp.linear_code._h.value, types.code, types.degree_plus_one,
                            pp.num_rows, batch)
        if isinstance(evals, torch.Tensor):
            if evals.dtype != types.eval_dtype or evals.numel() != n:
                raise ShapeError(L.ZP_ERR_SHAPE, f"expected {n} evals of {types.eval_dtype}")
            h_ptr, d_ptr = None, _dev_ptr(evals)
        else:
            want = np.int32 if types.eval_type == "i32" else np.int64
            evals = np.ascontiguousarray(evals)
            if evals.dtype.itemsize != np.dtype(want).itemsize or evals.size != n:
                raise ShapeError(L.ZP_ERR_SHAPE, f"expected {n} evals of {np.dtype(want)}")
            h_ptr, d_ptr = evals.ctypes.data_as(ctypes.c_void_p), None
        rid = ctypes.c_uint64()
        _check(self._lib.zp_server_commit(self._h, ctypes.byref(desc), h_ptr, d_ptr,
                                          ctypes.byref(rid)))
        self._keep[rid.value] = (pp, types, batch, evals)
        return rid.value

    def open(self, request, alphas=None, coeffs=None, cols=(), comb_limbs=None):
        """alphas: int64 [batch, D, 2] (i128 words), coeffs: int64 [num_rows, 2];
        cols: column indices.  comb_limbs defaults to the types' CombR limbs."""
        if request not in self._keep:
            raise InvalidPcsParam(L.ZP_ERR_INVALID_PARAM, "unknown request")
        pp, types, batch, _ = self._keep[request]
        limbs = 0 if alphas is None else (comb_limbs or types.comb_limbs)
        cols = np.ascontiguousarray(np.asarray(cols, dtype=np.uint64))
        sizes = [ctypes.c_uint64() for _ in range(3)]
        _check(self._lib.zp_server_open_sizes(self._h, request, limbs, len(cols),
                                              *[ctypes.byref(x) for x in sizes]))
        comb = np.zeros(max(1, sizes[0].value // 8), dtype=np.uint64)
        columns = np.zeros(max(1, sizes[1].value), dtype=np.uint8)
        sib = np.zeros(max(1, sizes[2].value), dtype=np.uint8)
        al = co = None
        if limbs:
            al = np.ascontiguousarray(alphas, dtype=np.int64)
            co = np.ascontiguousarray(coeffs, dtype=np.int64)
            if al.size != 2 * batch * types.degree_plus_one or co.size != 2 * pp.num_rows:
                raise ShapeError(L.ZP_ERR_SHAPE, "alphas/coeffs have the wrong length")
        vp = ctypes.c_void_p
        _check(self._lib.zp_server_open(
            self._h, request, limbs, al.ctypes.data_as(vp) if limbs else None,
            co.ctypes.data_as(vp) if limbs else None, cols.ctypes.data_as(vp), len(cols),
            comb.ctypes.data_as(vp), columns.ctypes.data_as(vp), sib.ctypes.data_as(vp)))
        self._opening[request] = (limbs, cols, comb, columns, sib, al, co,
                                  sizes[1].value, sizes[2].value)

    def prove(self, request, point, comb_limbs=None, num_openings=None):
        """Queue ZipPlus::prove of a committed request at an i128 point
        (the server draws the transcript from the root); the event's
        "proof" is (modulus, eval, proof bytes)."""
        if request not in self._keep:
            raise InvalidPcsParam(L.ZP_ERR_INVALID_PARAM, "unknown request")
        _, types, _, _ = self._keep[request]
        pt = np.zeros((max(1, len(point)), 2), dtype=np.int64)
        for i, v in enumerate(point):
            u = int(v) & ((1 << 128) - 1)
            pt[i, 0] = np.uint64(u & ((1 << 64) - 1)).view(np.int64)
            pt[i, 1] = np.uint64(u >> 64).view(np.int64)
        _check(self._lib.zp_server_prove(
            self._h, request, pt.ctypes.data_as(ctypes.c_void_p), len(point),
            comb_limbs or types.comb_limbs,
            types.num_column_openings if num_openings is None else num_openings))
        self._opening[request] = "prove"

    def poll(self, timeout=None):
        """Next finished job as a dict, or None after `timeout` seconds."""
        ev = L.ServerEvent()
        t = -1 if timeout is None else int(timeout * 1e6)
        rc = self._lib.zp_server_poll(self._h, t, ctypes.byref(ev))
        if rc == L.ZP_ERR_TIMEOUT:
            return None
        _check(rc)
        out = {"request": ev.request, "kind": ("commit", "open", "prove")[ev.kind],
               "status": ev.status, "queued_us": ev.queued_us, "service_us": ev.service_us,
               "latency_us": ev.latency_us,
               "message": ev.message.decode(errors="replace")}
        if ev.kind == 0:
            out["root"] = bytes(ev.root)
            if ev.status != L.ZP_OK:
                self._keep.pop(ev.request, None)
        elif ev.kind == 2:
            self._keep.pop(ev.request, None)
            self._opening.pop(ev.request, None)
            if ev.status == L.ZP_OK:
                buf = np.zeros(max(1, ev.proof_bytes), dtype=np.uint8)
                n = ctypes.c_uint64()
                mod, evl = (ctypes.c_uint8 * 32)(), (ctypes.c_uint8 * 32)()
                _check(self._lib.zp_server_take_proof(
                    self._h, ev.request, buf.ctypes.data_as(ctypes.c_void_p), ev.proof_bytes,
                    ctypes.byref(n), mod, evl))
                out["proof"] = (int.from_bytes(bytes(mod), "little"),
                                int.from_bytes(bytes(evl), "little"), buf[:n.value].tobytes())
        else:
            pp, types, batch, _ = self._keep.pop(ev.request)
            limbs, cols, comb, columns, sib, _, _, cb, sb = self._opening.pop(ev.request)
            row_len = pp.linear_code.row_len()
            out["comb"] = comb[:row_len * limbs].reshape(row_len, limbs) if limbs else None
            per = cb // max(1, len(cols))
            out["columns"] = [columns[k * per:(k + 1) * per].tobytes() for k in range(len(cols))]
            h = sb // max(1, len(cols)) // 32
            out["siblings"] = [[sib[(k * h + l) * 32:(k * h + l + 1) * 32].tobytes()
                                for l in range(h)] for k in range(len(cols))]
        return out

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(5)]
        _check(self._lib.zp_server_stats(self._h, *[ctypes.byref(x) for x in v]))
        keys = ["hbm_in_use", "hbm_peak", "budget", "evals_committed", "pending"]
        return dict(zip(keys, (x.value for x in v)))
