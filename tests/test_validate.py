"""validate_input's bit budget (zip-plus/src/pcs/utils.rs:24-54) at the
boundary, on the host (zp_validate_input needs no device).

The reference asserts actual_lc_bits <= CombR::NUM_BITS with
  actual = 2*ceil_log2(max(REP*d, 1)) + ceil_log2(max(d, 1)) + chal_bits
           + (COEFF_BIT_WIDTH - 1) + ceil_log2(batch),   d = ilog2(row_len),
so each case below is worked out by hand and placed exactly on, and one bit
past, the limit."""
import pytest

import zipgpu


def ok(rep, row_len, types, batch):
    zipgpu.validate_input(rep, row_len, types, batch)


def panics(rep, row_len, types, batch, bits, limit):
    with pytest.raises(zipgpu.ShapeError) as e:
        zipgpu.validate_input(rep, row_len, types, batch)
    assert (f"The number of bits used for linear combinations is too large: {bits} bits, "
            f"max allowed is {limit} bits") in str(e.value)


def test_i32_bench_types_batch_boundary():
    # Eval i32 (32 bits), Chal i128, Comb = CombR (no degree): 128 challenge
    # bits.  row_len 2^14: d = 14, REP*d = 56 -> 6, ceil_log2(14) = 4:
    # 12 + 4 + 128 + 31 = 175, plus ceil_log2(batch).  Int<3> = 192 bits.
    t = zipgpu.IntTypes(comb_limbs=3)
    ok(4, 1 << 14, t, 1)
    ok(4, 1 << 14, t, 1 << 17)  # 175 + 17 = 192: allowed (<=)
    panics(4, 1 << 14, t, (1 << 17) + 1, 193, 192)


def test_binary_poly_alphas_double_the_challenge_bits():
    # BinaryPoly<32>: COEFF_BIT_WIDTH 1, DensePolynomial Comb with degree
    # bound 31 > 0 -> 256 challenge bits.  row_len 2^14: 12 + 4 + 256 + 0 =
    # 272 bits: Int<4> (256) panics, Int<5> (320) holds.
    panics(4, 1 << 14, zipgpu.BinPolyTypes(32, comb_limbs=4), 1, 272, 256)
    ok(4, 1 << 14, zipgpu.BinPolyTypes(32, comb_limbs=5), 1)
    # BinaryPoly<1> has degree bound 0: 128 challenge bits, 144 total.
    ok(4, 1 << 14, zipgpu.BinPolyTypes(1, comb_limbs=3), 1)
    panics(4, 1 << 14, zipgpu.BinPolyTypes(1, comb_limbs=2), 1, 144, 128)


def test_rep_and_row_len_enter_through_ceil_logs():
    # RAA row 2^11, REP 4: d = 11, REP*d = 44 -> 6, ceil_log2(11) = 4.
    # i32: 12 + 4 + 128 + 31 = 175; with batch 2^17: 192 = Int<3>.
    t = zipgpu.IntTypes(comb_limbs=3)
    ok(4, 1 << 11, t, 1 << 17)
    panics(4, 1 << 11, t, (1 << 17) + 1, 193, 192)
    # row_len 2^9, REP 2: 2*9 = 18 -> 5 (doubled: 10), ceil_log2(9) = 4:
    # BinaryPoly<64> 10 + 4 + 256 = 270, batch 3 -> +2 = 272 > Int<4>.
    panics(2, 1 << 9, zipgpu.BinPolyTypes(64, comb_limbs=4), 3, 272, 256)
    # row_len 1: d = 0, max(REP*0, 1) = 1 -> 0 bits, max(d, 1) = 1 -> 0:
    # i64 evals (63 + 128 = 191) fit Int<3> exactly.
    ok(4, 1, zipgpu.I64Types(comb_limbs=3), 1)
    ok(4, 1, zipgpu.I64Types(comb_limbs=3), 2)  # + ceil_log2(2) = 1: 192 bits
    panics(4, 1, zipgpu.I64Types(comb_limbs=3), 3, 193, 192)
    # row_len 2: d = 1, REP*d = 4 -> 2 bits (doubled: 4), ceil_log2(1) = 0:
    # 4 + 0 + 128 + 63 = 195.
    panics(4, 2, zipgpu.I64Types(comb_limbs=3), 1, 195, 192)
    # row_len 2^10 under REP 8: 8*10 = 80 -> 7; ceil_log2(10) = 4.
    # DensePolynomial<i64, 4>: 14 + 4 + 256 + 63 = 337 > Int<5>.
    panics(8, 1 << 10, zipgpu.DensePolyTypes(4, comb_limbs=5), 1, 337, 320)
    ok(8, 1 << 10, zipgpu.DensePolyTypes(4, comb_limbs=6), 1)


def test_non_power_of_two_row_len_uses_floor_log2():
    # usize::ilog2 floors: row_len 3 * 2^10 has d = 11 (not 12).
    # i32: 2*ceil_log2(44) = 12, ceil_log2(11) = 4 -> 175 bits.
    ok(4, 3 << 10, zipgpu.IntTypes(comb_limbs=3), 1 << 17)
    panics(4, 3 << 10, zipgpu.IntTypes(comb_limbs=3), (1 << 17) + 1, 193, 192)


def test_argument_errors():
    with pytest.raises(zipgpu.ShapeError):
        zipgpu.validate_input(4, 1 << 14, zipgpu.IntTypes(), 0)
    with pytest.raises(zipgpu.InvalidPcsParam):
        zipgpu.validate_input(4, 1 << 14, zipgpu.BinPolyTypes(65), 1)
