"""Pins of the oracle's checked encodings (the reference's CHECKED build:
IprsCode<.., CHECKED> and RaaConfig::CHECK_FOR_OVERFLOWS), CPU only.

The checked build panics at the first product or partial sum that leaves
the Cw type (code/iprs.rs:76-84, pntt/radix8.rs:150-157,
butterfly.rs:34-42, raa.rs:205-218, raa_sign_flip.rs:146-155) and
otherwise returns the unchecked values.  Each test below fixes its verdict
from a closed form on the code's base matrix (itself pinned against the
reference's tables), not from a second copy of the fold.
"""
import numpy as np
import pytest

from zipgpu import synth

MAX = (1 << 63) - 1
MIN = -(1 << 63)


def words(vals, L=1):
    """Python ints -> uint64 limbs [n, L] (two's complement)."""
    mask = (1 << 64) - 1
    out = np.zeros((len(vals), L), dtype=np.uint64)
    for i, v in enumerate(vals):
        for l in range(L):
            out[i, l] = (v >> (64 * l)) & mask
    return out


def test_checked_partial_sum_panics_even_when_every_final_fits(oracle):
    # depth 0, row_len 4, REP 2: output r = x0 + x1 f(r) + x2 f(2r) + x3 f(3r)
    # folded left to right (radix8.rs:150-157).  With x0 = MAX - 3000,
    # x1 = 1, x2 = 5 the partial x0 + f(1) exceeds MAX at row 1 (f(1) =
    # 4096), yet every final sum fits: the reference panics, a check of
    # final sums alone (or a fold from the other end) would not.
    M = np.asarray(oracle.base_matrix(4, 0, 2), dtype=np.int64).reshape(8, 4)
    a, b = 1, 5
    for c, want in ((3000, True), (5000, False)):
        x0 = MAX - c
        finals = [x0 + a * int(M[r, 1]) + b * int(M[r, 2]) for r in range(8)]
        partials = [x0 + a * int(M[r, 1]) for r in range(8)]
        assert all(MIN <= v <= MAX for v in finals)
        assert any(v > MAX for v in partials) == want
        # the fold from the far end never leaves the range here
        assert all(MIN <= x0 + b * int(M[r, 2]) <= MAX for r in range(8))
        x = words([x0, a, b, 0]).reshape(1, 4, 1)
        out, ovf = oracle.iprs_encode_wide_checked(x, 4, 0, 1, 2)
        assert ovf == want
        if not want:
            assert np.array_equal(out, oracle.iprs_encode_wide(x, 4, 0, 1, 2))


@pytest.mark.parametrize("L", [1, 2])
def test_checked_product_threshold(oracle, L):
    # depth 1, row_len 64: the single input v at index 8 (column 1 of chunk
    # 0) reaches base output r of chunk 0 as v * f(r) and nothing else is
    # non-zero, so the butterflies add only zero products.  The reference
    # panics iff some v * f(r) leaves Int<L>.
    lo, hi = -(1 << (64 * L - 1)), (1 << (64 * L - 1)) - 1
    col = [int(v) for v in np.asarray(oracle.base_matrix(64, 1, 4)).reshape(32, 8)[:, 1]]
    fpos, fneg = max(col), -min(col)
    vmax = min(hi // fpos, -lo // fneg)       # largest positive v that fits
    umax = min(-lo // fpos, hi // fneg)       # largest |v| of a negative v that fits
    for v, want in ((vmax, False), (vmax + 1, True), (-umax, False), (-umax - 1, True)):
        vals = [0] * 64
        vals[8] = v
        x = words(vals, L).reshape(1, 64, L)
        out, ovf = oracle.iprs_encode_wide_checked(x, 64, 1, L, 4)
        assert ovf == want, (v, want)
        if not want:
            assert np.array_equal(out, oracle.iprs_encode_wide(x, 64, 1, L, 4))


def test_checked_addition_threshold(oracle):
    # x0 = MAX - K at index 0 (from_ref of chunk 0) and 1 at index 8: base
    # output r of chunk 0 is x0 + f(r), every product is |f| < 2^16, so the
    # reference panics iff max f > K.
    col = [int(v) for v in np.asarray(oracle.base_matrix(64, 1, 4)).reshape(32, 8)[:, 1]]
    fmax = max(col)
    for K, want in ((fmax, False), (fmax - 1, True)):
        vals = [0] * 64
        vals[0], vals[8] = MAX - K, 1
        x = words(vals).reshape(1, 64, 1)
        _, ovf = oracle.iprs_encode_wide_checked(x, 64, 1, 1, 4)
        assert ovf == want


def small_rows(seed, rows, row_len, L, bits):
    """Int<L> rows of values |x| < 2^bits, sign-extended, as uint64 limbs."""
    w = synth.splitmix64(seed, rows * row_len).view(np.int64) >> np.int64(64 - bits)
    x = np.repeat(np.where(w < 0, np.int64(-1), np.int64(0))[:, None], L, axis=1)
    x[:, 0] = w
    return x.reshape(rows, row_len, L).view(np.uint64)


@pytest.mark.parametrize("row_len,depth,L,bits", [(64, 1, 2, 40), (512, 2, 3, 60),
                                                   (512, 3, 2, 40), (256, 0, 4, 60),
                                                   (512, 2, 1, 8)])
def test_checked_equals_unchecked_without_overflow(oracle, row_len, depth, L, bits):
    # Values below the overflow bound of the code (bits + log2 of its growth
    # < 64L - 1): the checked build returns the unchecked codewords.
    x = small_rows(40 + row_len + L, 2, row_len, L, bits)
    out, ovf = oracle.iprs_encode_wide_checked(x, row_len, depth, L)
    assert not ovf
    assert np.array_equal(out, oracle.iprs_encode_wide(x, row_len, depth, L))
    out, ovf = oracle.raa_encode_wide_checked(x, row_len, L)
    assert not ovf
    assert np.array_equal(out, oracle.raa_encode_wide(x, row_len, L))


def test_checked_full_range_rows_overflow(oracle):
    # Uniform Int<L> rows: the first product by an entry |f| >= 2 of a value
    # above 2^(64L-2) already leaves the range.
    for L in (1, 3):
        x = synth.comb_rows(7 + L, 2 * 64, L).reshape(2, 64, L).view(np.uint64)
        assert oracle.iprs_encode_wide_checked(x, 64, 1, L)[1]
        assert oracle.raa_encode_wide_checked(x, 64, L)[1]


@pytest.mark.parametrize("sign_flip", [False, True])
def test_checked_raa_accumulate_threshold(oracle, sign_flip):
    # One non-zero input v > 0.  Every prefix of either accumulate is v times
    # an integer count, and the encoding of v = 1 shows the counts c_i of the
    # second one; the first one's are at most REP in size.  So the reference
    # panics iff v max c > MAX or v min c < MIN.
    row_len, rep = 16, 4
    vals = [0] * row_len
    vals[2] = 1
    one = oracle.raa_encode_wide(words(vals).reshape(1, row_len, 1), row_len, 1, rep, sign_flip)
    c = one.view(np.int64)[0, :, 0].astype(object)
    cmax, cmin = int(max(c)), int(min(c))
    assert rep <= max(cmax, -cmin)
    vthr = MAX // cmax if cmin >= 0 else min(MAX // cmax, -MIN // -cmin)
    for v, want in ((vthr, False), (vthr + 1, True)):
        vals[2] = v
        x = words(vals).reshape(1, row_len, 1)
        out, ovf = oracle.raa_encode_wide_checked(x, row_len, 1, rep, sign_flip)
        assert ovf == want, (v, want)
        if not want:
            assert np.array_equal(out, oracle.raa_encode_wide(x, row_len, 1, rep, sign_flip))


def test_checked_raa_negating_min_panics(oracle):
    # raa_sign_flip.rs:146-155: neg! of Int::MIN panics; MIN at odd index 1
    # is negated before the first accumulate.
    for L in (1, 2):
        vals = [0] * 8
        vals[1] = -(1 << (64 * L - 1))
        x = words(vals, L).reshape(1, 8, L)
        assert oracle.raa_encode_wide_checked(x, 8, L, 1, True)[1]
        assert not oracle.raa_encode_wide_checked(words([0] * 8, L).reshape(1, 8, L), 8, L, 1,
                                                  True)[1]
