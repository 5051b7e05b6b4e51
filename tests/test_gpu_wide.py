"""LinearCode::encode_wide on the device (zip-plus/src/code.rs:55) against
the oracle: rows of CombR = Int<L>, arithmetic mod 2^(64L), bit-exact."""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def limbs(got):
    return got.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("row_len,depth,L,rows,small", [
    (256, 1, 5, 3, 0), (512, 2, 3, 2, 2), (2048, 3, 8, 1, 0), (64, 0, 5, 2, 2),
    (4096, 1, 2, 1, 0),     # base_len 512: the chunk inputs pass through LDS in two tiles
    (1 << 12, -1, 4, 2, 2), (8, 1, 1, 4, 0), (1 << 14, 2, 5, 1, 2), (32, 1, 6, 2, 0),
])
def test_iprs_encode_wide(zg, dev, oracle, row_len, depth, L, rows, small):
    say("iprs_wide", row_len, depth, L, rows, small)
    x = synth.comb_rows(900 + row_len + L, rows * row_len, L, small).reshape(rows, row_len, L)
    code = zg.IprsCode(row_len, depth)
    got = limbs(code.encode_wide(to_dev(x, dev)))
    want = oracle.iprs_encode_wide(x.view(np.uint64), row_len, depth, L)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("row_len,depth,L", [(1 << 14, 2, 1), (1 << 14, 2, 5), (256, 1, 8),
                                              (2048, 2, 2), (1 << 13, 1, 3)])
def test_iprs_wide_base_layers_agree(zg, dev, oracle, row_len, depth, L):
    # The matrix-core base layer (signed byte digits x int8 slices) and the
    # per-element base layer are separate device paths; rows of extreme
    # limbs (all ones, sign bit only, alternating) and random rows.
    say("wide_base", row_len, depth, L)
    rows = 4
    x = synth.comb_rows(77 + row_len + L, rows * row_len, L, 0).reshape(rows, row_len, L)
    x[0] = np.uint64(0xffffffffffffffff).view(np.int64)
    x[1] = 0
    x[1, :, L - 1] = np.uint64(1 << 63).view(np.int64)
    x[2, ::2] = np.uint64(0x7fffffffffffffff).view(np.int64)
    code = zg.IprsCode(row_len, depth)
    xd = to_dev(x, dev)
    try:
        zg.set_option("wide_base", 1)
        generic = limbs(code.encode_wide(xd))
    finally:
        zg.set_option("wide_base", -1)
    fast = limbs(code.encode_wide(xd))
    want = oracle.iprs_encode_wide(x.view(np.uint64), row_len, depth, L)
    assert np.array_equal(generic, want) and np.array_equal(fast, want)


@pytest.mark.parametrize("row_len,L,flip,rows,small", [
    (2048, 5, False, 2, 2), (2048, 5, True, 2, 2), (256, 3, True, 3, 0), (8, 8, False, 4, 0),
    (1 << 13, 4, True, 1, 0), (16, 1, True, 5, 0), (1 << 11, 8, False, 1, 2),
])
def test_raa_encode_wide(zg, dev, oracle, row_len, L, flip, rows, small):
    say("raa_wide", row_len, L, flip, rows, small)
    x = synth.comb_rows(950 + row_len + L, rows * row_len, L, small).reshape(rows, row_len, L)
    code = zg.RaaSignFlippingCode(row_len) if flip else zg.RaaCode(row_len)
    got = limbs(code.encode_wide(to_dev(x, dev)))
    want = oracle.raa_encode_wide(x.view(np.uint64), row_len, L, sign_flip=flip)
    assert np.array_equal(got, want)


def test_single_row_shape_and_errors(zg, dev, oracle):
    row_len, L = 512, 5
    x = synth.comb_rows(3, row_len, L, 2)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    got = code.encode_wide(to_dev(x, dev))
    assert tuple(got.shape) == (4 * row_len, L)
    assert np.array_equal(limbs(got), oracle.iprs_encode_wide(x.view(np.uint64), row_len, -1, L)[0])
    with pytest.raises(zg.ShapeError):
        code.encode_wide(to_dev(x[:-1], dev))
    with pytest.raises(zg.InvalidPcsParam):
        code.encode_wide(torch.zeros((row_len, 9), dtype=torch.int64, device=dev))
    empty = code.encode_wide(torch.zeros((0, row_len, L), dtype=torch.int64, device=dev))
    assert tuple(empty.shape) == (0, 4 * row_len, L)


@pytest.mark.parametrize("depth", [1, 2])
def test_wide_agrees_with_i32_encode(zg, dev, depth):
    # An i32 row sign-extended to Int<4> encodes to the i128 codeword of the
    # i32 path (iprs.rs:120-130), sign-extended to Int<4>: both apply the same
    # linear map and |codeword| < 2^127.
    row_len, rows, L = 1 << 12, 2, 4
    say("wide_vs_i32", depth)
    x32 = synth.i32_evals(77 + depth, rows * row_len)
    code = zg.IprsCode(row_len, depth)
    cw = code.encode_rows(to_dev(x32, dev), rows, zg.IntTypes()).cpu().numpy().view(np.int64)
    cw = cw.reshape(rows, 4 * row_len, 2)
    ext = np.where(cw[..., 1:] < 0, np.int64(-1), np.int64(0))
    want = np.concatenate([cw, ext, ext], axis=-1).view(np.uint64)
    xw = x32.astype(np.int64).reshape(rows, row_len, 1)
    xw = np.concatenate([xw] + [np.where(xw < 0, np.int64(-1), np.int64(0))] * (L - 1), axis=-1)
    got = limbs(code.encode_wide(to_dev(xw, dev)))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind,row_len", [("iprs", 1 << 14), ("raa", 1 << 11)])
def test_wide_linear_at_bench_size(zg, dev, oracle, kind, row_len):
    # encode_wide(a) + encode_wide(b) == encode_wide(a + b) mod 2^(64L) at the
    # benchmarked shape, with one row also checked against the oracle.
    rows, L = 16, 5
    say("wide_linear", kind, row_len)
    a = synth.comb_rows(11, rows * row_len, L).reshape(rows, row_len, L)
    b = synth.comb_rows(12, rows * row_len, L).reshape(rows, row_len, L)
    s = oracle.limbs_add(a.view(np.uint64), b.view(np.uint64)).view(np.int64)
    code = zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs" else zg.RaaCode(row_len)
    ea, eb, es = (limbs(code.encode_wide(to_dev(v, dev))) for v in (a, b, s))
    assert np.array_equal(oracle.limbs_add(ea, eb), es)
    if kind == "iprs":
        want = oracle.iprs_encode_wide(a[5:6].view(np.uint64), row_len, -1, L)
    else:
        want = oracle.raa_encode_wide(a[5:6].view(np.uint64), row_len, L)
    assert np.array_equal(ea[5:6], want)
