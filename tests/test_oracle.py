"""Pins of the CPU oracle against the reference's known answers, published
test vectors and closed forms (CPU only).

Each test names what fixes the expected value: a known answer from the
reference's own tests, a public test vector, or mathematics (a closed form,
an invariant, an independent computation).
"""
import ctypes

import numpy as np
import pytest

from zipgpu import synth

P = 65537


# ---------------------------------------------------------------------------
# F65537 parameters and octet reversal
# ---------------------------------------------------------------------------

def test_octet_reversal_known_results(oracle):
    # zip-plus/src/code/iprs/pntt/radix8/octet_reversal.rs:105-115
    assert oracle.octet_reversal(0b101_111_010_000_110, 5) == 0b110_000_010_111_101
    assert oracle.octet_reversal(0b101_111_010_000_110_000, 6) == 0b000_110_000_010_111_101
    assert oracle.octet_reversal(0, 0) == 0
    for i in range(8):
        assert oracle.octet_reversal(i, 1) == i


def test_octet_reversal_self_inverse(oracle):
    # octet_reversal.rs:78-91
    for d in range(0, 5):
        for x in range(1 << (3 * d)):
            assert oracle.octet_reversal(oracle.octet_reversal(x, d), d) == x


def test_octet_reversal_zero_octets_asserts(oracle):
    with pytest.raises(ValueError):
        oracle.octet_reversal(1, 0)


def test_twiddles_are_eighth_roots(oracle):
    # params.rs:232-242: BASE_TWIDDLES == precompute_roots_of_unity(8)
    assert list(oracle.roots_of_unity(8)) == [1, 4096, -256, 16, -1, -4096, 256, -16]


def test_roots_closed_form(oracle):
    # w_n = 3^(65536/n) is a primitive n-th root (ark F65537, generator 3).
    for n in [2, 8, 64, 1024, 65536]:
        r = oracle.roots_of_unity(n) % P
        w = pow(3, 65536 // n, P)
        assert int(r[1]) == w
        assert pow(w, n, P) == 1 and (n == 1 or pow(w, n // 2, P) == P - 1)
        assert all(int(r[i]) == pow(w, i, P) for i in range(0, n, max(1, n // 64)))


def test_normalization_rule(oracle):
    # params.rs:176-181: x >= (p-1)/2 ? x - p : x.  Residue 32768 lifts to
    # -32769 and residue 32769 to -32768 (not antisymmetric).
    n = 1024
    r = oracle.roots_of_unity(n)
    res = r % P
    assert r.min() >= -32769 and r.max() <= 32767
    assert np.all(r[res >= 32768] == res[res >= 32768] - P)
    assert np.all(r[res < 32768] == res[res < 32768])
    assert (res == 32768).any() and (res == 32769).any()  # the quirk is live at n=1024
    assert np.all(r[res == 32768] == -32769) and np.all(r[res == 32769] == -32768)


@pytest.mark.parametrize("row_len,depth,ok", [
    (256, 1, True), (1 << 14, 2, True), (1 << 15, 2, False),  # codeword_len >= p
    (12, 1, False),   # not a multiple of 8^depth (params.rs:67-73)
    (1, 1, False),    # iprs.rs:225-235
    (24 * 8, 1, False),  # base_dim = 24*4 not a power of two (params.rs:75-80)
])
def test_params_errors(oracle, row_len, depth, ok):
    assert (oracle.iprs_params(row_len, depth) is not None) == ok


def test_optimal_depth_table(oracle):
    # iprs.rs:43-51
    table = {2**8: (1, 32, 128), 2**9: (1, 64, 256), 2**10: (1, 128, 512),
             2**11: (1, 256, 1024), 2**12: (2, 64, 256), 2**13: (2, 128, 512),
             2**14: (2, 256, 1024)}
    for row_len, (d, bl, bd) in table.items():
        p = oracle.iprs_params(row_len, -1)
        assert (p["depth"], p["base_len"], p["base_dim"]) == (d, bl, bd)
        assert p["codeword_len"] == 4 * row_len
    assert oracle.iprs_params(1 << 15, -1) is None


def test_base_matrix_closed_form(oracle):
    # M[r][k] = lift(w_bd^(r*k)) (params.rs:145-160)
    M = oracle.base_matrix(256, 1)
    bd, bl = M.shape
    w = pow(3, 65536 // bd, P)
    for r in range(0, bd, 7):
        for k in range(bl):
            v = pow(w, r * k, P)
            assert M[r, k] == (v - P if v >= 32768 else v)


def test_stage_twiddles_closed_form(oracle):
    # tw_k[i][j-1][t] = lift(BT[t] * w_N^(8^(d-1-k) * i * j)) (params.rs:114-143)
    row_len, depth = 1 << 12, 2
    N = 4 * row_len
    bt = [1, 4096, -256, 16, -1, -4096, 256, -16]
    w = pow(3, 65536 // N, P)
    for k in range(depth):
        tw = oracle.twiddles(row_len, depth, k)
        cpp = 8 ** (depth - 1 - k)
        for i in range(0, tw.shape[0], 37):
            for j in range(1, 8):
                for t in range(8):
                    v = (bt[t] % P) * pow(w, cpp * i * j, P) % P
                    assert tw[i, j - 1, t] == (v - P if v >= 32768 else v)


# ---------------------------------------------------------------------------
# pNTT (IPRS) outputs
# ---------------------------------------------------------------------------

def dft_mod_p(x, N):
    """sum_k x_k w_N^(i*k) mod p for the zero-padded row (exact, int64)."""
    w = pow(3, 65536 // N, P)
    pw = np.array([pow(w, e, P) for e in range(N)], dtype=np.int64)
    x = np.asarray(x, dtype=np.int64) % P
    i = np.arange(N, dtype=np.int64)[:, None]
    k = np.arange(len(x), dtype=np.int64)[None, :]
    terms = pw[(i * k) % N] * x[None, :] % P
    return terms.sum(axis=1) % P


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_pntt_mod_p_equals_dft(oracle, depth):
    # radix8.rs:224-284 pntt_against_arkworks: 0..row_len, REP 2, base_len 32.
    row_len = 32 * 8 ** depth
    if depth == 3:
        row_len = 4 * 8 ** depth  # keep N*row_len small; base_len 4
    x = np.arange(row_len, dtype=np.int32)
    out = oracle.iprs_encode_i32(x, 1, row_len, depth, rep=2)[0]
    vals = [(int(lo) & (2**64 - 1)) | (int(hi) << 64) for lo, hi in out]
    got = np.array([v % P for v in vals], dtype=np.int64)
    assert np.array_equal(got, dft_mod_p(x, 2 * row_len))


@pytest.mark.parametrize("row_len,depth", [(256, 1), (512, 2), (4096, 2)])
def test_binary_pntt_mod_p_per_plane(oracle, row_len, depth):
    # Every coefficient plane of [i64;D] is the pNTT of that bit plane, so
    # mod p it is the DFT of the bits.
    D = 8
    x = synth.binary_polys(11, row_len, D).view(np.uint64)
    out = oracle.iprs_encode_binary(x, 1, row_len, depth, D)[0]
    for c in [0, 3, 7]:
        bits = ((x >> np.uint64(c)) & np.uint64(1)).astype(np.int64)
        assert np.array_equal(out[:, c] % P, dft_mod_p(bits, 4 * row_len))


def test_binary_planes_equal_scalar_path(oracle):
    # Independent code paths: [i64;D] planes vs the i32 -> i128 pNTT of each
    # plane's 0/1 values.
    row_len, depth, D = 4096, 2, 6
    x = synth.binary_polys(5, row_len, D).view(np.uint64)
    out = oracle.iprs_encode_binary(x, 1, row_len, depth, D)[0]
    for c in range(D):
        bits = ((x >> np.uint64(c)) & np.uint64(1)).astype(np.int32)
        ref = oracle.iprs_encode_i32(bits, 1, row_len, depth)[0]
        assert np.array_equal(ref[:, 0], out[:, c])
        assert np.array_equal(ref[:, 1], out[:, c] >> 63)  # sign extension fits i64


def test_pntt_integer_linearity(oracle):
    # phase_commit.rs:578-600 (linearity of encoding over the integers)
    row_len = 512
    a = synth.i32_evals(1, row_len) >> 8
    b = synth.i32_evals(2, row_len) >> 8
    ea = oracle.iprs_encode_i32(a, 1, row_len)[0]
    eb = oracle.iprs_encode_i32(b, 1, row_len)[0]
    es = oracle.iprs_encode_i32(a + b, 1, row_len)[0]

    def i128(w):
        return [(int(lo) & (2**64 - 1)) | (int(hi) << 64) for lo, hi in w]

    s = [(x + y) & (2**128 - 1) for x, y in zip(i128(ea), i128(eb))]
    assert s == [v & (2**128 - 1) for v in i128(es)]


def test_pntt_brute_force_tiny(oracle):
    # An independent Python restatement of radix8.rs on a 32-entry row
    # (depth 1), with exact integers.
    row_len, depth, rep = 32, 1, 4
    N, bl, bd = 128, 4, 16
    x = [int(v) for v in synth.i32_evals(9, row_len)]
    wb = pow(3, 65536 // bd, P)
    wn = pow(3, 65536 // N, P)

    def lift(v):
        return v - P if v >= 32768 else v

    base = []
    for i in range(N):
        c, r = i // bd, i % bd
        base.append(sum(x[c + 8 * k] * lift(pow(wb, r * k, P)) for k in range(bl)))
    bt = [1, 4096, -256, 16, -1, -4096, 256, -16]
    table = [[(t * j) % 8 for j in range(1, 8)] for t in range(8)]
    out = list(base)
    for i in range(bd):
        xs = [base[j * bd + i] for j in range(8)]
        for t in range(8):
            out[t * bd + i] = xs[0] + sum(
                xs[j] * lift((bt[table[t][j - 1]] % P) * pow(wn, i * j, P) % P)
                for j in range(1, 8))
    got = oracle.iprs_encode_i32(np.array(x, dtype=np.int32), 1, row_len, depth, rep)[0]
    vals = [(int(lo) & (2**64 - 1)) | (int(hi) << 64) for lo, hi in got]
    assert vals == out  # signed i128 values


# ---------------------------------------------------------------------------
# RAA and the rand 0.9 restatement
# ---------------------------------------------------------------------------

RAA_KAT = [0x1E, 0x36, 0x39, 0x5A, 0x70, 0x7E, 0xA5, 0xC1, 0xCB, 0xDC, 0xF9, 0x11E, 0x124,
           0x14C, 0x14D, 0x160]


def test_raa_known_answer(oracle):
    # zip-plus/src/code/raa.rs:396-413 (row [1,2,3,4], REP 4, seeds 1 and 2).
    assert list(oracle.raa_encode_i64([1, 2, 3, 4])) == RAA_KAT


def test_raa_linearity_and_zero(oracle):
    # raa.rs:374-394, :415-428
    a = np.arange(1, 5)
    b = np.arange(5, 9)
    assert np.array_equal(oracle.raa_encode_i64(a) + oracle.raa_encode_i64(b),
                          oracle.raa_encode_i64(a + b))
    assert not oracle.raa_encode_i64(np.zeros(64, dtype=np.int64)).any()


def test_shuffle_properties(oracle):
    # raa.rs:341-372: deterministic per seed, differs across seeds, permutes.
    v = np.arange(1, 11)
    s1, s2, s3 = oracle.shuffle(v, 12345), oracle.shuffle(v, 12345), oracle.shuffle(v, 54321)
    assert np.array_equal(s1, s2) and not np.array_equal(s1, s3)
    assert not np.array_equal(s1, v) and sorted(s1) == list(v)
    for n in [1, 2, 16, 1000, 8192]:
        p = oracle.raa_perm(n, 1)
        assert sorted(p) == list(range(n))


def test_in_place_equals_precomputed(oracle):
    # raa.rs:430-455: shuffling data in place == indexing by shuffled 0..N.
    data = np.arange(1, 4097, dtype=np.uint64) * np.uint64(7)
    p = oracle.raa_perm(4096, 1)
    assert np.array_equal(oracle.shuffle(data, 1), data[p.astype(np.int64)])


def test_raa_repeat_accumulate_brute_force(oracle):
    # Independent restatement of raa.rs:93-126 using the oracle's permutations.
    row = np.array([3, -1, 4, 1, -5, 9, 2, 6], dtype=np.int64)
    N = 32
    p1, p2 = oracle.raa_perm(N, 1).astype(int), oracle.raa_perm(N, 2).astype(int)
    y = np.tile(row, 4)
    y = np.cumsum(y[p1])
    y = np.cumsum(y[p2])
    assert np.array_equal(oracle.raa_encode_i64(row), y)
    z = np.tile(row, 4)
    z[1::2] *= -1
    z = np.cumsum(z[p1])
    z[1::2] *= -1
    z = np.cumsum(z[p2])
    assert np.array_equal(oracle.raa_encode_i64(row, sign_flip=True), z)


def test_raa_binary_planes(oracle):
    row_len, D = 64, 32
    x = synth.binary_polys(3, row_len, D).view(np.uint64)
    out = oracle.raa_encode_binary(x, 1, row_len, D)[0]
    for c in [0, 17, 31]:
        bits = ((x >> np.uint64(c)) & np.uint64(1)).astype(np.int64)
        assert np.array_equal(out[:, c], oracle.raa_encode_i64(bits))


# ---------------------------------------------------------------------------
# BLAKE3 and the Merkle tree
# ---------------------------------------------------------------------------

# Published BLAKE3 test vectors (input byte i = i mod 251) and common KATs.
B3_VECTORS = {
    0: "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262",
    1: "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213",
    1023: "10108970eeda3eb932baac1428c7a2163b0e924c9a9e25b35bba72b28f70bd11",
    1024: "42214739f095a406f3fc83deb889744ac00df831c10daa55189b5d121c855af7",
    1025: "d00278ae47eb27b34faecf67b4fe263f82d5412916c1ffd97c8cb7fb814b8444",
    2048: "e776b6028c7cd22a4d0ba182a8bf62205d2ef576467e838ed6f2529b85fba24a",
    2049: "5f4d72f40d7a5f82b15ca2b2e44b1de3c2ef86c426c95c1af0b6879522563030",
    3072: "b98cb0ff3623be03326b373de6b9095218513e64f1ee2edd2525c7ad1e5cffd2",
    4096: "015094013f57a5277b59d8475c0501042c0b642e531b0a1c8f58d2163229e969",
    8192: "aae792484c8efe4f19e2ca7d371d8c467ffb10748d8a5a1ae579948f718a2a63",
}


@pytest.mark.parametrize("n", sorted(B3_VECTORS))
def test_blake3_vectors(oracle, n):
    data = bytes(i % 251 for i in range(n))
    assert oracle.blake3(data).hex() == B3_VECTORS[n]


def test_blake3_abc(oracle):
    assert oracle.blake3(b"abc").hex() == \
        "6437b3ac38465133ffb63b75273a8db548c558465d79db03fd359c6cd5bd9d85"


def test_hazmat_merges_rebuild_hash(oracle):
    # A 4-chunk input hashed in one go equals root(nonroot(cv0,cv1),
    # nonroot(cv2,cv3)) of its chunk CVs (merkle.rs relies on exactly this).
    data = bytes((7 * i + 3) % 256 for i in range(4096))
    cvs = [oracle.blake3_chunk_cv(data[1024 * i:1024 * (i + 1)], i) for i in range(4)]
    left = oracle.blake3_parent(cvs[0], cvs[1], False)
    right = oracle.blake3_parent(cvs[2], cvs[3], False)
    assert oracle.blake3_parent(left, right, True) == oracle.blake3(data)


def test_merkle_empty_and_single(oracle):
    # merkle.rs:175-189: n = 0 -> blake3(""), n = 1 -> root = leaf.
    assert oracle.merkle_layers(np.zeros((0, 32), np.uint8))[0].tobytes().hex() == B3_VECTORS[0]
    leaf = np.frombuffer(oracle.blake3(b"x"), dtype=np.uint8).reshape(1, 32)
    assert np.array_equal(oracle.merkle_layers(leaf)[-1], leaf[0])


def test_merkle_leaves_are_column_hashes(oracle):
    rows, width, elem = 3, 8, 16
    m = synth.splitmix64(4, rows * width * elem // 8).view(np.uint8).reshape(rows, width, elem)
    leaves = oracle.merkle_leaves(m, rows, width, elem)
    for j in range(width):
        assert leaves[j].tobytes() == oracle.blake3(m[:, j, :].tobytes())
    layers = oracle.merkle_layers(leaves)
    assert layers[8].tobytes() == oracle.blake3_parent(leaves[0].tobytes(), leaves[1].tobytes(),
                                                       False)
    assert layers[-1].tobytes() == oracle.blake3_parent(layers[12].tobytes(),
                                                        layers[13].tobytes(), True)


def test_merkle_prove_verify_roundtrip(oracle):
    # merkle.rs:443-469 round trip, plus tamper detection.
    import ctypes
    rows, width, elem = 4, 16, 8
    m = synth.splitmix64(6, rows * width).view(np.uint8).reshape(rows, width, elem)
    leaves = oracle.merkle_leaves(m, rows, width, elem)
    layers = oracle.merkle_layers(leaves)
    root = layers[-1].tobytes()
    for leaf in [0, 5, 15]:
        sib = np.zeros((8, 32), np.uint8)
        n = ctypes.c_uint64()
        oracle.load().or_merkle_prove(layers.ctypes.data_as(ctypes.c_void_p), ctypes.c_uint64(width),
                                      ctypes.c_uint64(leaf), sib.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.byref(n))
        sibs = [sib[i].tobytes() for i in range(n.value)]
        col = m[:, leaf, :].tobytes()
        assert oracle.merkle_verify(root, col, leaf, width, sibs)
        bad = bytearray(col)
        bad[0] ^= 1
        assert not oracle.merkle_verify(root, bytes(bad), leaf, width, sibs)


# ---------------------------------------------------------------------------
# Commit and the combined row
# ---------------------------------------------------------------------------

def test_commit_equals_encode_then_merkle(oracle):
    # phase_commit.rs:61-101: root over the poly-major concatenation of rows.
    row_len, rows, D, batch = 256, 4, 8, 2
    x = synth.binary_polys(21, batch * rows * row_len, D).view(np.uint64)
    res = oracle.commit("iprs", x, rows, row_len, batch=batch, D=D, want_cw=True)
    cw = np.concatenate([oracle.iprs_encode_binary(x[b * rows * row_len:(b + 1) * rows * row_len],
                                                   rows, row_len, -1, D) for b in range(batch)])
    assert cw.tobytes() == res["cw"].tobytes()
    assert res["root"] == oracle.merkle_root(cw, batch * rows, 4 * row_len, 8 * D)


def big(limbs):
    v = 0
    for i, w in enumerate(limbs):
        v |= (int(w) & (2**64 - 1)) << (64 * i)
    return v


def test_combined_row_bigint(oracle):
    # phase_prove.rs:153-244 with exact Python integers, summed in the
    # reference's order (per eval alpha projection, then rows).
    D, L, rows, row_len, batch = 8, 5, 4, 16, 2
    x = synth.binary_polys(31, batch * rows * row_len, D).view(np.uint64)
    alphas = synth.i128_words(32, batch * D)
    coeffs = synth.i128_words(33, rows)

    def i128(w):
        v = big(w)
        return v - 2**128 if v >= 2**127 else v

    al = [i128(a) for a in alphas]
    co = [i128(c) for c in coeffs]
    got = oracle.combined_row(x, alphas, coeffs, rows, row_len, batch, D, L)
    for col in range(row_len):
        acc = 0
        for b in range(batch):
            for j in range(rows):
                e = int(x[(b * rows + j) * row_len + col])
                w = sum(al[b * D + c] for c in range(D) if (e >> c) & 1)
                acc += co[j] * w
        assert big(got[col]) == acc % 2**(64 * L)


def test_alpha_projection_matches_definition(oracle):
    # binary_u64.rs:393-432: sum of alpha_c over set bits (incl. negatives).
    D, L = 64, 5
    ev = synth.binary_polys(41, 32, D).view(np.uint64)
    alphas = synth.i128_words(42, D)
    got = oracle.alpha_projection(ev, alphas, D, L)
    for k in range(32):
        want = 0
        for c in range(D):
            if (int(ev[k]) >> c) & 1:
                v = big(alphas[c])
                want += v - 2**128 if v >= 2**127 else v
        assert big(got[k]) == want % 2**320


# ---------------------------------------------------------------------------
# encode_wide over Int<L> (code.rs:55; iprs.rs:148-150; raa.rs:151-153)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("row_len,depth,L", [(256, 1, 3), (512, 2, 5), (2048, 3, 8), (64, 1, 4)])
def test_iprs_wide_equals_i32_path_on_small_inputs(oracle, row_len, depth, L):
    # i32 rows sign-extended to Int<L>: the wide encoder must give the i128
    # codewords of the (pinned) i32 path, sign-extended to L limbs.
    from zipgpu import synth
    rows = 2
    x = synth.i32_evals(31 + row_len, rows * row_len)
    want = oracle.iprs_encode_i32(x, rows, row_len, depth)           # [rows, N, 2]
    wide_in = oracle.sign_extend_limbs(x.astype(np.int64)[:, None], L).reshape(rows, row_len, L)
    got = oracle.iprs_encode_wide(wide_in, row_len, depth, L)
    assert np.array_equal(got, oracle.sign_extend_limbs(want, L))


@pytest.mark.parametrize("kind,L", [("iprs", 5), ("raa", 5), ("raa_flip", 3), ("iprs", 8)])
def test_wide_linear_mod_2_64L(oracle, kind, L):
    # Full-range limbs: encode(x + y) = encode(x) + encode(y) modulo 2^(64L).
    from zipgpu import synth
    row_len = 256
    x = synth.splitmix64(5, row_len * L).reshape(1, row_len, L)
    y = synth.splitmix64(6, row_len * L).reshape(1, row_len, L)
    if kind == "iprs":
        enc = lambda v: oracle.iprs_encode_wide(v, row_len, -1, L)  # noqa: E731
    else:
        enc = lambda v: oracle.raa_encode_wide(v, row_len, L, sign_flip=kind == "raa_flip")  # noqa: E731
    lhs = enc(oracle.limbs_add(x, y))
    rhs = oracle.limbs_add(enc(x), enc(y))
    assert np.array_equal(lhs, rhs)
    assert not np.array_equal(enc(x), enc(y))


@pytest.mark.parametrize("flip", [False, True])
def test_raa_wide_equals_i64_path(oracle, flip):
    # Small integer rows: the wide RAA encoder agrees with the i64 restatement
    # that reproduces the raa.rs:396-413 known answer.
    row_len, L = 64, 4
    rng = np.random.default_rng(3)
    row = rng.integers(-1000, 1000, row_len).astype(np.int64)
    want = np.zeros(row_len * 4, dtype=np.int64)
    assert oracle.load().or_raa_encode_i64(ctypes.c_uint64(row_len), ctypes.c_uint64(4),
                                          ctypes.c_int(int(flip)), oracle._p(row),
                                          oracle._p(want)) == 0
    got = oracle.raa_encode_wide(oracle.sign_extend_limbs(row[:, None], L)[None], row_len, L,
                                 sign_flip=flip)
    assert np.array_equal(got[0], oracle.sign_extend_limbs(want[:, None], L))
