"""Commits of i64 and DensePolynomial<i64, D> polys on the device against
the oracle: codewords and roots bit-exact (IntZipTypes and
ArbitraryPolyZipTypesIprs/Raa, protocol/src/lib.rs:490-559)."""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def make_code(zg, kind, row_len):
    return {"iprs": lambda: zg.IprsCode.new_with_optimal_depth(row_len),
            "raa": lambda: zg.RaaCode(row_len),
            "raa_signflip": lambda: zg.RaaSignFlippingCode(row_len)}[kind]()


def evals_i64(seed, n, small):
    w = synth.splitmix64(seed, n).view(np.int64)
    return (w >> np.int64(40)) if small else w


@pytest.mark.parametrize("kind,row_len,num_rows,eval_type,D,batch,small", [
    ("iprs", 256, 8, "i64", 1, 1, True), ("iprs", 1024, 4, "i64", 1, 2, False),
    ("iprs", 512, 4, "dense_i64", 8, 1, True), ("iprs", 256, 2, "dense_i64", 32, 1, False),
    ("raa", 256, 8, "i64", 1, 1, False), ("raa_signflip", 128, 16, "i64", 1, 1, True),
    ("raa", 256, 4, "dense_i64", 8, 2, True), ("raa_signflip", 64, 8, "dense_i64", 32, 1, False),
    ("iprs", 1 << 12, 4, "dense_i64", 32, 1, True),
])
def test_i64_commit_matches_oracle(zg, dev, oracle, kind, row_len, num_rows, eval_type, D, batch,
                                   small):
    say("i64 commit", kind, row_len, num_rows, eval_type, D, batch, small)
    types = zg.I64Types() if eval_type == "i64" else zg.DensePolyTypes(D)
    W = types.eval_words
    n = num_rows * row_len
    x = evals_i64(70 + row_len + D, batch * n * W, small)
    code = make_code(zg, kind, row_len)
    zp = zg.ZipPlus(types)
    pp = zp.setup(n, code)
    polys = [torch.from_numpy(x[b * n * W:(b + 1) * n * W]).to(dev) for b in range(batch)]
    hint, comm = zp.commit(pp, polys)
    want = oracle.commit_i64(kind, x, num_rows, row_len, batch, D, eval_type)
    assert comm.root == want["root"]
    L = oracle.i64_plane_limbs(kind, eval_type)
    for b in range(batch):
        got = hint.codewords(b).view(np.uint64).reshape(num_rows, -1, W, L)
        assert np.array_equal(got, want["cw"][b * num_rows:(b + 1) * num_rows])
    hint.free()


def test_i64_encode_rows_shapes(zg, dev):
    code = zg.RaaCode(64)
    types = zg.DensePolyTypes(4)
    x = torch.zeros(2 * 64 * 4, dtype=torch.int64, device=dev)
    out = code.encode_rows(x, 2, types)
    assert tuple(out.shape) == (2, 256, 4, 4)
    with pytest.raises(zg.ShapeError):
        code.encode_rows(x[:-1], 2, types)
