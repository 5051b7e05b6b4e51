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


def _sizes_commit(oracle, kind, x, n, cols, types, eval_type):
    if eval_type == "binpoly":
        return oracle.commit(kind, x.view(np.uint64), 1, n, batch=cols, D=types.degree_plus_one)
    return oracle.commit_i64(kind, x, 1, n, cols, types.degree_plus_one, eval_type)


@pytest.mark.parametrize("empty_arb", [False, True])
def test_step0_commit_matches_oracle(zg, dev, oracle, empty_arb):
    # ZincPlusPiop::step0_commit (protocol/src/prover.rs:218-262) with the
    # protocol's IPRS set-up (checked codes, one row per poly, row_len = poly
    # size, protocol/src/lib.rs:574-575): binary, arbitrary and int witness
    # traces committed, roots and public entries absorbed.  The arbitrary
    # trace's coefficients are small, as the checked codes require.
    say("step0", empty_arb)
    n = 1 << 10
    D = 32
    kinds = [("binpoly", zg.BinPolyTypes(D, 8)), ("dense_i64", zg.DensePolyTypes(D)),
             ("i64", zg.I64Types())]
    params, wit, pub, host_w, host_p = [], [], [], [], []
    for g, (et, types) in enumerate(kinds):
        code = zg.IprsCode.new_with_optimal_depth(n, checked=True)
        pp = zg.ZipPlus(types).setup(n, code)
        params.append((pp, types))
        W = types.eval_words
        nw = 0 if (empty_arb and et == "dense_i64") else 2
        if et == "binpoly":
            xs = [synth.binary_polys(200 + 10 * g + k, n, D) for k in range(nw + 1)]
        else:
            xs = [evals_i64(200 + 10 * g + k, n * W, True) >> np.int64(16 if W > 1 else 0)
                  for k in range(nw + 1)]
        host_w.append(xs[1:])
        host_p.append(xs[:1])
        wit.append([torch.from_numpy(v).to(dev) for v in xs[1:]])
        pub.append([torch.from_numpy(v).to(dev) for v in xs[:1]])
    hints, comms, t = zg.step0_commit(params, wit, pub)
    ts = oracle.Transcript()
    for g, (et, types) in enumerate(kinds):
        cols = host_w[g]
        if not cols:
            assert hints[g] is None and comms[g].root == bytes(32) and comms[g].batch_size == 0
            ts.absorb_slice(bytes(32))
            continue
        want = _sizes_commit(oracle, "iprs", np.concatenate(cols), n, len(cols), types, et)
        assert comms[g].root == want["root"] and comms[g].batch_size == len(cols)
        ts.absorb_slice(want["root"])
    for g, (et, types) in enumerate(kinds):
        entry = 8 * types.eval_words
        for col in host_p[g]:
            raw = col.tobytes()
            for i in range(len(raw) // entry):
                ts.absorb_slice(raw[i * entry:(i + 1) * entry])
    assert t.get_challenge(32) == ts.challenge(32)
    for h in hints:
        if h is not None:
            h.free()
