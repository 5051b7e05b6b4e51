"""Pins of the oracle's opening path (CPU): BLAKE3 extended output, the
transcript's prime draw, the field, b and the combined row, and the
prove/verify round trip (zip-plus/src/pcs/phase_prove.rs:85-304,
phase_verify.rs:100-305, transcript/src/lib.rs:10-72)."""
import random

import numpy as np
import pytest

from zipgpu import synth

# Published BLAKE3 extended output (131 bytes) of the empty input.
B3_XOF_EMPTY = (
    "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262e00f03e7b69af26b7faaf09fcd33"
    "3050338ddfe085b8cc869ca98b206c08243a26f5487789e8f660afe6c99ef9e0c52b92e7393024a80459cf91f476"
    "f9ffdbda7001c22e159b402631f277ca96f2defdf1078282314e763699a31c5363165421cce14d")


def test_xof_published_vector(oracle):
    assert oracle.blake3_xof(b"", 131).hex() == B3_XOF_EMPTY


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1023, 1024, 1025, 2049, 4096, 8192])
def test_xof_prefix_is_the_hash(oracle, n):
    data = bytes(i % 251 for i in range(n))
    out = oracle.blake3_xof(data, 200)
    assert out[:32] == oracle.blake3(data)
    assert oracle.blake3_xof(data, 77) == out[:77]


def miller_rabin(n, bases=(2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53)):
    if n < 2:
        return False
    for q in bases:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in bases:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_sprp2_is_the_base_two_test(oracle):
    # Strong pseudoprimes to base 2 pass (the test is base 2 only,
    # primality/src/lib.rs:13-19); odd composites that are not, fail.
    for n in (2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633):
        assert oracle.sprp2(n)
        assert not miller_rabin(n)
    for n in (9, 15, 21, 341, 561, 645, 1105, 1387, 1729, 1905):
        assert not oracle.sprp2(n)
    for n in ((1 << 255) - 19, (1 << 256) - 189, (1 << 127) - 1, 65537):
        assert oracle.sprp2(n)


def test_random_field_cfg_is_first_probable_prime(oracle):
    ts = oracle.Transcript()
    ts.absorb_slice(bytes(range(32)))
    p = ts.random_field_cfg()
    assert p % 2 == 1 and miller_rabin(p) and p.bit_length() > 128
    # Replay the candidate stream: every earlier candidate fails the test.
    t2 = oracle.Transcript()
    t2.absorb_slice(bytes(range(32)))
    while True:
        buf = oracle.blake3_xof(t2.absorbed, 32)
        t2.absorb(buf)
        c = int.from_bytes(buf, "little")
        c = c - 1 if c and c % 2 == 0 else c
        if c == p:
            break
        assert not miller_rabin(c)
    assert t2.absorbed == ts.absorbed


def _limbs(v, n=4):
    return np.array([(v >> (64 * i)) & ((1 << 64) - 1) for i in range(n)], dtype=np.uint64)


def _int(a):
    return sum(int(a[i]) << (64 * i) for i in range(len(a)))


def test_field_ops(oracle):
    rng = random.Random(5)
    lib = oracle.load()
    for p in ((1 << 255) - 19, (1 << 256) - 189, (1 << 192) + 133):
        for _ in range(50):
            a, b = rng.randrange(p), rng.randrange(p)
            r = np.zeros(4, np.uint64)
            lib.or_field_mul(oracle._p(_limbs(a)), oracle._p(_limbs(b)), oracle._p(_limbs(p)),
                             oracle._p(r))
            assert _int(r) == a * b % p
            lib.or_field_inner(oracle._p(_limbs(a)), oracle._p(_limbs(p)), oracle._p(r))
            assert _int(r) == (a << 256) % p
            v = rng.randrange(-(1 << 319), 1 << 319)
            lib.or_field_from_signed(oracle._p(_limbs(v % (1 << 320), 5)), oracle._u64(5),
                                     oracle._p(_limbs(p)), oracle._p(r))
            assert _int(r) == v % p


def _alpha_proj(bits, alphas, D):
    return sum(alphas[c] for c in range(D) if (bits >> c) & 1)


@pytest.mark.parametrize("eval_type,D,batch", [("binpoly", 32, 1), ("binpoly", 7, 2),
                                                ("i32", 1, 1), ("binpoly", 64, 1)])
def test_prove_b_by_definition(oracle, eval_type, D, batch):
    rng = random.Random(D + batch)
    p = (1 << 256) - 189
    num_rows, row_len = 4, 32
    n = batch * num_rows * row_len
    if eval_type == "binpoly":
        ev = synth.binary_polys(D, n, D).view(np.uint64)
        alphas = [[rng.randrange(-(1 << 127), 1 << 127) for _ in range(D)] for _ in range(batch)]
    else:
        ev = synth.i32_evals(3, n)
        alphas = [[1]]
    q1 = [rng.randrange(p) for _ in range(row_len)]
    got = oracle.prove_b(p, ev, alphas, q1, num_rows, row_len, batch, D, eval_type)
    for j in range(num_rows):
        want = 0
        for b in range(batch):
            for c in range(row_len):
                x = int(ev[(b * num_rows + j) * row_len + c])
                w = _alpha_proj(x, alphas[b], D) if eval_type == "binpoly" else x
                want += w * q1[c]
        assert got[j] == want % p


def test_combined_row_i32_by_definition(oracle):
    num_rows, row_len, L = 8, 16, 3
    ev = synth.i32_evals(4, num_rows * row_len)
    co = synth.i128_words(5, num_rows)
    got = oracle.combined_row_i32(ev, co, num_rows, row_len, 1, L)
    s = [int(np.uint64(co[j, 0].view(np.uint64))) | (int(co[j, 1]) << 64) for j in range(num_rows)]
    for c in range(row_len):
        want = sum(s[j] * int(ev[j * row_len + c]) for j in range(num_rows)) % (1 << 192)
        assert _int(got[c]) == want


def _setup(oracle, kind, row_len, num_rows, D, batch, eval_type, seed):
    n = num_rows * row_len
    if eval_type == "binpoly":
        ev = synth.binary_polys(seed, batch * n, D).view(np.uint64)
        elem, coeff_bytes = (8 * D, 8) if kind == "iprs" else (4 * D, 4)
    else:
        ev = synth.i32_evals(seed, batch * n)
        elem, coeff_bytes = 16, 16
    c = oracle.commit(kind, ev, num_rows, row_len, batch=batch, D=D, eval_type=eval_type,
                      want_cw=True, want_layers=True)
    return ev, c, elem, coeff_bytes


def expected_proof_bytes(num_rows, row_len, N, elem, L, batch, num_openings):
    """The proof-stream size of phase_commit.rs:741-771 (the reference's
    proof_size_is_correct_for_parameters) for MontyField<4>: b is one vector
    of num_rows field elements summed over the polys (phase_prove.rs:179-193),
    each 32 bytes of modulus and 32 of value, with no length prefix because
    Uint<4> is ConstTranscribable (transcript/src/traits.rs:88-89); the
    combined row is row_len CombR = Int<L>; each opening is the column of
    every poly's rows plus a Merkle proof of three u64 and one 32-byte hash
    per level of the tree over N leaves."""
    depth = (N - 1).bit_length()  # log2 of N.next_power_of_two()
    return (num_rows * 64 + row_len * 8 * L +
            num_openings * (batch * num_rows * elem + 3 * 8 + depth * 32))


def test_proof_size_formula_matches_the_reference_helper():
    # At batch 1 the restatement is the reference helper with a zero-length
    # prefix; the cfg2 proof (2^20 BinaryPoly<32>, IPRS row 2^14 x 64, Int<5>,
    # 147 openings) is 3146696 bytes.
    assert expected_proof_bytes(64, 1 << 14, 1 << 16, 256, 5, 1, 147) == 3146696
    assert expected_proof_bytes(1, 512, 2048, 8, 5, 1, 147) == 64 + 512 * 40 + 147 * (8 + 24 + 11 * 32)


def _encoder(oracle, kind, row_len, L):
    if kind == "iprs":
        return lambda comb: oracle.iprs_encode_wide(comb.reshape(1, row_len, L), row_len, -1, L)[0]
    return lambda comb: oracle.raa_encode_wide(comb.reshape(1, row_len, L), row_len, L,
                                               sign_flip=kind == "raa_signflip")[0]


@pytest.mark.parametrize("kind,row_len,num_rows,D,batch,L,eval_type", [
    ("iprs", 256, 16, 32, 1, 5, "binpoly"), ("raa", 64, 64, 16, 2, 5, "binpoly"),
    ("raa_signflip", 128, 8, 8, 1, 6, "binpoly"), ("iprs", 512, 1, 1, 1, 5, "binpoly"),
    ("iprs", 256, 4, 1, 1, 3, "i32"),
])
def test_prove_verify_roundtrip_and_eval_closed_form(oracle, kind, row_len, num_rows, D, batch, L,
                                                     eval_type):
    ev, c, elem, cb = _setup(oracle, kind, row_len, num_rows, D, batch, eval_type, 11 + D)
    N = 4 * row_len
    nv = (num_rows * row_len).bit_length() - 1
    point = [int(x) for x in synth.i128_words(13, nv)[:, 0]]
    ts = oracle.Transcript()
    ts.absorb_slice(c["root"])
    p = ts.random_field_cfg()
    before = bytes(ts.absorbed)
    e = oracle.prove(ts, p, ev, num_rows, row_len, N, elem, c["cw"], c["layers"], point,
                     batch=batch, D=D, L=L, eval_type=eval_type, num_openings=20)
    # eval = sum over polys of the MLE of the alpha-projected evals at the point.
    t2 = oracle.Transcript()
    t2.absorbed = bytearray(before)
    alphas = [[t2.challenge_i128() for _ in range(D)] if eval_type == "binpoly" and D > 1 else [1]
              for _ in range(batch)]
    eq = oracle.eq_table([v % p for v in point], p)
    want = 0
    for b in range(batch):
        for i in range(num_rows * row_len):
            x = int(ev[b * num_rows * row_len + i])
            w = _alpha_proj(x, alphas[b], D) if eval_type == "binpoly" else x
            want += w * eq[i]
    assert e == want % p
    assert len(ts.stream) == expected_proof_bytes(num_rows, row_len, N, elem, L, batch, 20)
    tv = oracle.Transcript(ts.stream)
    tv.absorb_slice(c["root"])
    assert tv.random_field_cfg() == p
    assert oracle.verify(tv, p, c["root"], point, e, _encoder(oracle, kind, row_len, L), num_rows,
                         row_len, N, elem, batch=batch, D=D, L=L, eval_type=eval_type,
                         coeff_bytes=cb, num_openings=20)


@pytest.mark.parametrize("where", ["b", "comb", "column", "path", "eval"])
def test_verify_rejects_planted_mistakes(oracle, where):
    kind, row_len, num_rows, D, L = "iprs", 256, 8, 16, 5
    ev, c, elem, cb = _setup(oracle, kind, row_len, num_rows, D, 1, "binpoly", 3)
    N = 4 * row_len
    point = [int(x) for x in synth.i128_words(7, 11)[:, 0]]
    ts = oracle.Transcript()
    ts.absorb_slice(c["root"])
    p = ts.random_field_cfg()
    e = oracle.prove(ts, p, ev, num_rows, row_len, N, elem, c["cw"], c["layers"], point, D=D, L=L,
                     num_openings=5)
    stream = bytearray(ts.stream)
    b_end = num_rows * 64
    comb_end = b_end + row_len * L * 8
    off = {"b": 40, "comb": b_end + 3, "column": comb_end + 5,
           "path": comb_end + num_rows * elem + 24 + 7}.get(where)
    if off is not None:
        stream[off] ^= 1
    tv = oracle.Transcript(bytes(stream))
    tv.absorb_slice(c["root"])
    tv.random_field_cfg()
    with pytest.raises(ValueError):
        oracle.verify(tv, p, c["root"], point, e + (where == "eval"), _encoder(oracle, kind,
                      row_len, L), num_rows, row_len, N, elem, D=D, L=L, num_openings=5)


# ---------------------------------------------------------------------------
# i64 and DensePolynomial<i64, D> evaluations (the protocol's commits).
# ---------------------------------------------------------------------------

def test_iprs_i64_small_equals_i32_path(oracle):
    # |x| < 2^31: Eval = i64 -> Cw = i128 is the pinned i32 -> i128 encode.
    row_len, rows = 512, 3
    x = synth.i32_evals(8, rows * row_len)
    got = oracle.encode_i64_planes("iprs", x.astype(np.int64), rows, row_len, 1, 2)
    want = oracle.iprs_encode_i32(x, rows, row_len).view(np.uint64).reshape(rows, -1, 1, 2)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", ["iprs", "raa", "raa_signflip"])
def test_dense_binary_coefficients_equal_binary_path(oracle, kind):
    # DensePolynomial<i64, D> with 0/1 coefficients encodes like BinaryPoly<D>
    # (IPRS: both [i64; D]; RAA: [Int<4>; D] against [i32; D] sign-extended).
    row_len, rows, D = 256, 2, 8
    bits = synth.binary_polys(9, rows * row_len, D).view(np.uint64)
    dense = ((bits[:, None] >> np.arange(D, dtype=np.uint64)) & np.uint64(1)).astype(np.int64)
    L = 1 if kind == "iprs" else 4
    got = oracle.encode_i64_planes(kind, dense, rows, row_len, D, L)
    if kind == "iprs":
        want = oracle.iprs_encode_binary(bits, rows, row_len, -1, D).view(np.uint64)
        assert np.array_equal(got[..., 0], want.reshape(rows, -1, D))
    else:
        want = oracle.raa_encode_binary(bits, rows, row_len, D,
                                        sign_flip=kind == "raa_signflip").astype(np.int64)
        assert np.array_equal(got, oracle.sign_extend_limbs(want.reshape(rows, -1, D, 1), 4))


def test_raa_i64_small_equals_i64_path(oracle):
    row_len, rows = 128, 2
    x = synth.i32_evals(10, rows * row_len).astype(np.int64)
    got = oracle.encode_i64_planes("raa", x, rows, row_len, 1, 2)
    for r in range(rows):
        want = oracle.raa_encode_i64(x[r * row_len:(r + 1) * row_len])
        assert np.array_equal(got[r, :, 0, :], oracle.sign_extend_limbs(
            np.asarray(want, dtype=np.int64).reshape(-1, 1), 2))
