"""Checked codes on the device (the reference's CHECKED build) against the
oracle's checked encodings: Overflow exactly where the reference panics,
and otherwise the unchecked codewords, bit-exact."""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu

MAX = (1 << 63) - 1
MIN = -(1 << 63)


def say(*a):
    print("[case]", *a, flush=True)


def words(vals, L=1):
    mask = (1 << 64) - 1
    out = np.zeros((len(vals), L), dtype=np.uint64)
    for i, v in enumerate(vals):
        for l in range(L):
            out[i, l] = (v >> (64 * l)) & mask
    return out


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).to(dev)


def host(t):
    return t.cpu().numpy().view(np.uint64)


def wide_verdict(zg, code, x, dev):
    """(device codewords or None on Overflow) of encode_wide."""
    try:
        return host(code.encode_wide(to_dev(x, dev)))
    except zg.Overflow:
        return None


def check_wide(zg, oracle, dev, kind, x, row_len, depth, L, rep=4):
    """encode_wide of a checked code against the oracle's checked encoding."""
    if kind == "iprs":
        code = zg.IprsCode(row_len, depth, rep, checked=True)
        want, ovf = oracle.iprs_encode_wide_checked(x, row_len, depth, L, rep)
    else:
        code = (zg.RaaCode if kind == "raa" else zg.RaaSignFlippingCode)(row_len, rep, checked=True)
        want, ovf = oracle.raa_encode_wide_checked(x, row_len, L, rep, kind == "raa_signflip")
    assert code.checked
    got = wide_verdict(zg, code, x, dev)
    assert (got is None) == ovf
    if not ovf:
        assert np.array_equal(got, want)
    return ovf


def test_checked_fold_order_and_thresholds(zg, dev, oracle):
    # The oracle-pinned boundary cases: a partial sum that overflows while
    # every final fits; the largest product and sum that fit and the
    # smallest that do not.
    M = np.asarray(oracle.base_matrix(4, 0, 2), dtype=np.int64).reshape(8, 4)
    for c, want in ((3000, True), (5000, False)):
        say("fold order", c)
        x = words([MAX - c, 1, 5, 0]).reshape(1, 4, 1)
        assert check_wide(zg, oracle, dev, "iprs", x, 4, 0, 1, 2) == want
    col = [int(v) for v in np.asarray(oracle.base_matrix(64, 1, 4)).reshape(32, 8)[:, 1]]
    for L in (1, 2):
        lo, hi = -(1 << (64 * L - 1)), (1 << (64 * L - 1)) - 1
        vmax = min(hi // max(col), -lo // -min(col))
        umax = min(-lo // max(col), hi // -min(col))
        for v, want in ((vmax, False), (vmax + 1, True), (-umax, False), (-umax - 1, True)):
            say("product threshold", L, v)
            vals = [0] * 64
            vals[8] = v
            assert check_wide(zg, oracle, dev, "iprs", words(vals, L).reshape(1, 64, L), 64, 1,
                              L) == want
    for K, want in ((max(col), False), (max(col) - 1, True)):
        say("sum threshold", K)
        vals = [0] * 64
        vals[0], vals[8] = MAX - K, 1
        assert check_wide(zg, oracle, dev, "iprs", words(vals).reshape(1, 64, 1), 64, 1,
                          1) == want


@pytest.mark.parametrize("kind", ["iprs", "raa", "raa_signflip"])
@pytest.mark.parametrize("L", [1, 3, 8])
def test_checked_encode_wide_matches_oracle(zg, dev, oracle, kind, L):
    row_len = 256
    for small in (1, 0):
        say("checked wide", kind, L, small)
        x = synth.comb_rows(300 + L, 3 * row_len, L, small).reshape(3, row_len, L)
        if small:  # below the code's overflow bound for every L
            x[..., 0] >>= np.int64(50 if L == 1 else 30)
        ovf = check_wide(zg, oracle, dev, kind, x.view(np.uint64), row_len, 1, L)
        assert ovf == (small == 0)


@pytest.mark.parametrize("sign_flip", [False, True])
def test_checked_raa_thresholds(zg, dev, oracle, sign_flip):
    kind = "raa_signflip" if sign_flip else "raa"
    row_len, rep = 16, 4
    vals = [0] * row_len
    vals[2] = 1
    one = oracle.raa_encode_wide(words(vals).reshape(1, row_len, 1), row_len, 1, rep, sign_flip)
    c = [int(v) for v in one.view(np.int64)[0, :, 0]]
    vthr = MAX // max(c) if min(c) >= 0 else min(MAX // max(c), -MIN // -min(c))
    for v, want in ((vthr, False), (vthr + 1, True)):
        say("raa threshold", sign_flip, v)
        vals[2] = v
        assert check_wide(zg, oracle, dev, kind, words(vals).reshape(1, row_len, 1), row_len, 0,
                          1, rep) == want
    if sign_flip:
        vals = [0] * 8
        vals[1] = MIN
        assert check_wide(zg, oracle, dev, kind, words(vals).reshape(1, 8, 1), 8, 0, 1, 1)


def bit_planes(w, D):
    """uint64 bit-packed evals -> [D] arrays of 0/1 Int<1> rows."""
    return [((w >> np.uint64(c)) & np.uint64(1)) for c in range(D)]


@pytest.mark.parametrize("row_len,density,want_ovf", [(4096, 1, True), (4096, 5, False),
                                                      (512, 1, False)])
def test_checked_binary_depth3(zg, dev, oracle, row_len, density, want_ovf):
    # A depth-3 IPRS code over BinaryPoly<32>: coefficients of uniform rows
    # leave i64 at row length 4096 (every plane), sparse rows and shorter
    # rows do not.  The verdict is the oracle's over the 32 coefficient
    # planes (the checked build adds per coefficient).
    say("binary depth 3", row_len, density)
    D, rows = 32, 2
    w = synth.splitmix64(5, rows * row_len)
    for k in range(1, density):
        w &= synth.splitmix64(5 + k, rows * row_len)
    ovf = [oracle.iprs_encode_wide_checked(p.reshape(rows, row_len, 1), row_len, 3, 1)[1]
           for p in bit_planes(w, D)]
    assert all(ovf) if want_ovf else not any(ovf)
    types = zg.BinPolyTypes(D)
    checked = zg.IprsCode(row_len, 3, checked=True)
    plain = zg.IprsCode(row_len, 3)
    ev = to_dev(w, dev)
    if want_ovf:
        with pytest.raises(zg.Overflow):
            checked.encode_rows(ev, rows, types)
        return
    got = checked.encode_rows(ev, rows, types).cpu().numpy()
    assert np.array_equal(got, plain.encode_rows(ev, rows, types).cpu().numpy())
    want = oracle.iprs_encode_binary(w, rows, row_len, 3, D)
    assert np.array_equal(got.reshape(want.shape), want)


@pytest.mark.parametrize("eval_type,row_len,depth,shift,want_ovf", [
    ("dense_i64", 1024, 2, 0, True),     # full-range coefficients: checked kernels, panic
    ("dense_i64", 1024, 2, 40, True),    # |x| < 2^24: checked kernels, panic
    ("dense_i64", 1024, 2, 57, False),   # |x| < 2^7: the magnitude scan clears them
    ("dense_i64", 512, 3, 50, False),    # |x| < 2^14 at depth 3: checked kernels, no panic
    ("i64", 4096, 3, 0, False),          # i64 -> i128 at depth 3: checked kernels, no panic
    ("i64", 1024, 2, 0, False),          # provably cannot overflow i128: unchecked kernels
])
def test_checked_integer_rows(zg, dev, oracle, eval_type, row_len, depth, shift, want_ovf):
    say("checked integer rows", eval_type, row_len, depth, shift)
    D = 4 if eval_type == "dense_i64" else 1
    rows = 2
    types = zg.DensePolyTypes(D) if eval_type == "dense_i64" else zg.I64Types()
    x = synth.splitmix64(17 + row_len + shift, rows * row_len * D).view(np.int64) >> np.int64(shift)
    # The oracle's verdict: each coefficient plane as Int<1> (Cw = i64) or
    # the i64 rows sign-extended to Int<2> (Cw = i128).
    if eval_type == "dense_i64":
        planes = x.reshape(rows, row_len, D)
        res = [oracle.iprs_encode_wide_checked(
            np.ascontiguousarray(planes[:, :, c]).view(np.uint64).reshape(rows, row_len, 1),
            row_len, depth, 1) for c in range(D)]
        ovf = any(r[1] for r in res)
    else:
        ext = np.where(x < 0, np.int64(-1), np.int64(0))
        limbs2 = np.stack([x, ext], axis=-1).view(np.uint64).reshape(rows, row_len, 2)
        ovf = oracle.iprs_encode_wide_checked(limbs2, row_len, depth, 2)[1]
    assert ovf == want_ovf
    checked = zg.IprsCode(row_len, depth, checked=True)
    plain = zg.IprsCode(row_len, depth)
    ev = torch.from_numpy(x).to(dev)
    if ovf:
        with pytest.raises(zg.Overflow):
            checked.encode_rows(ev, rows, types)
        return
    got = checked.encode_rows(ev, rows, types).cpu().numpy()
    assert np.array_equal(got, plain.encode_rows(ev, rows, types).cpu().numpy())


def test_checked_commit_and_noalloc(zg, dev, oracle):
    # ZipPlus::commit with a checked code: Overflow where encoding panics,
    # the unchecked commitment otherwise; zp_commit_noalloc refuses a
    # checked code whose binary rows need the host-read check (depth 3) and
    # commits with one that provably cannot overflow (depth 2).
    say("checked commit")
    row_len, rows = 1024, 4
    types = zg.DensePolyTypes(2)
    x = synth.splitmix64(99, rows * row_len * 2).view(np.int64)
    zp = zg.ZipPlus(types)
    pp = zp.setup(rows * row_len, zg.IprsCode(row_len, 2, checked=True))
    with pytest.raises(zg.Overflow):
        zp.commit(pp, [torch.from_numpy(x).to(dev)])
    small = x >> np.int64(57)  # |x| < 2^7: below the bound, cleared by the magnitude scan
    hint, comm = zp.commit(pp, [torch.from_numpy(small).to(dev)])
    want = oracle.commit_i64("iprs", small, rows, row_len, 1, 2, "dense_i64", 2)
    assert comm.root == want["root"]
    hint.free()

    import ctypes
    lib = zg.lib()
    D = 32
    for depth, ok in ((3, False), (2, True)):
        say("noalloc checked", depth)
        code = zg.IprsCode(4096, depth, checked=True)
        ev = to_dev(synth.splitmix64(3, 4096), dev)
        N = code.codeword_len()
        cw = torch.empty((N, D), dtype=torch.int64, device=dev)
        layers = torch.empty(((2 * N - 1) * 32,), dtype=torch.uint8, device=dev)
        rc = lib.zp_commit_noalloc(code._h, 1, D, ctypes.c_void_p(ev.data_ptr()), 1, 1,
                                   ctypes.c_void_p(cw.data_ptr()),
                                   ctypes.c_void_p(layers.data_ptr()), None)
        torch.cuda.synchronize()
        if not ok:
            assert rc == 1
            continue
        assert rc == 0
        plain = zg.IprsCode(4096, depth)
        assert torch.equal(cw, plain.encode_rows(ev, 1, zg.BinPolyTypes(D)).reshape(N, D))


def test_checked_codes_in_batch_plans_and_autotune(zg, dev, oracle):
    # A graph plan holding a checked depth-3 binary code (its check reads a
    # verdict on the host) replays eagerly: the roots equal the oracle's over
    # three commits, and rows that overflow raise Overflow.  The autotuner
    # times a checked dense code's variants on synthetic rows that overflow,
    # without failing, and leaves the code checked.
    say("checked batch plan")
    D, row_len = 32, 512
    code3 = zg.IprsCode(row_len, 3, checked=True)
    code2 = zg.IprsCode(1 << 12, 2, checked=True)
    types = zg.BinPolyTypes(D)
    items = [(zg.ZipPlus(types).setup(row_len * 2, code3), types),
             (zg.ZipPlus(types).setup(1 << 14, code2), types)]
    plan = zg.BatchCommit(items, graph=True)
    xs = [synth.binary_polys(40, row_len * 2, D), synth.binary_polys(41, 1 << 14, D)]
    want = [oracle.commit("iprs", xs[0], 2, row_len, depth=3, D=D)["root"],
            oracle.commit("iprs", xs[1], 4, 1 << 12, depth=2, D=D)["root"]]
    ev = [torch.from_numpy(x).to(dev) for x in xs]
    for rnd in range(3):
        got = plan.commit(ev)
        assert [c.root for c in got] == want, rnd
    big = zg.BatchCommit([(zg.ZipPlus(types).setup(4096 * 2, zg.IprsCode(4096, 3, checked=True)),
                           types)], graph=True)
    wide = to_dev(synth.splitmix64(5, 4096 * 2), dev)
    with pytest.raises(zg.Overflow):
        big.commit([wide])

    say("checked autotune")
    dense = zg.IprsCode(1024, 2, checked=True)
    dense.autotune(zg.DensePolyTypes(4), 4)
    assert dense.checked


def test_checked_eager_plan_scratch_follows_the_inputs(zg, dev, oracle):
    # A plan with a checked dense item: the first commit's rows need the
    # checked kernels (|x| < 2^12 at depth 2 is above the bound but fits),
    # the second's are cleared by the magnitude scan and take the
    # matrix-core base layer, whose scratch the first commit never used.
    # Both commits equal the plain commits' roots.
    say("checked plan, inputs change tier")
    rows, row_len, D = 4, 1024, 2
    types = zg.DensePolyTypes(D)
    pp = zg.ZipPlus(types).setup(rows * row_len, zg.IprsCode(row_len, 2, checked=True))
    plain_pp = zg.ZipPlus(types).setup(rows * row_len, zg.IprsCode(row_len, 2))
    plan = zg.BatchCommit([(pp, types)], graph=True)
    for shift in (52, 57, 52, 57):
        x = torch.from_numpy(synth.splitmix64(77 + shift, rows * row_len * D).view(np.int64)
                             >> np.int64(shift)).to(dev)
        got = plan.commit([x])[0].root
        h, want = zg.ZipPlus(types).commit(plain_pp, [x])
        h.free()
        assert got == want.root, shift
