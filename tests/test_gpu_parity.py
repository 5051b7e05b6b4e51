"""Bit-exact parity of the HIP path (through the C ABI) with the CPU oracle.

Inputs are seeded (zipgpu.synth); every codeword word, Merkle root and
combined-row limb must equal the oracle's.
"""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


# ---------------------------------------------------------------------------
# IPRS encode
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("row_len,depth,D,rows", [
    (256, 1, 32, 3), (512, -1, 64, 2), (1 << 12, 2, 32, 2), (1 << 14, 2, 32, 2),
    (1 << 14, 2, 64, 1), (32, 1, 3, 4), (2048, 3, 32, 1), (64, 0, 16, 2), (8, 1, 1, 8),
    (1 << 13, -1, 7, 1),
])
def test_iprs_binary_encode(zg, dev, oracle, row_len, depth, D, rows):
    say("iprs_binary", row_len, depth, D, rows)
    x = synth.binary_polys(100 + row_len + D, rows * row_len, D)
    code = zg.IprsCode(row_len, depth)
    types = zg.BinPolyTypes(D)
    got = code.encode_rows(to_dev(x, dev), rows, types).cpu().numpy()
    want = oracle.iprs_encode_binary(x.view(np.uint64), rows, row_len, depth, D)
    assert np.array_equal(got, want)


def test_iprs_binary_ignores_bits_above_degree(zg, dev, oracle):
    # BinaryPoly<D> only has D coefficients; higher bits of the u64 lane are
    # not read (binary_u64.rs:445-483, dense.rs:508-525).
    row_len, D = 512, 20
    x = synth.splitmix64(7, row_len).view(np.int64)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    got = code.encode(to_dev(x, dev), zg.BinPolyTypes(D)).cpu().numpy()
    assert np.array_equal(got, oracle.iprs_encode_binary(x.view(np.uint64), 1, row_len, -1, D)[0])


@pytest.mark.parametrize("row_len,depth,D", [
    (256, 1, 32), (2048, 1, 64), (1 << 13, 2, 32), (1 << 14, 2, 64), (1 << 12, 2, 16),
    (1 << 14, 2, 20), (256, 1, 8), (2048, 1, 40), (1 << 13, 2, 1), (1 << 12, 2, 63)])
def test_iprs_device_encoders_agree(zg, dev, oracle, row_len, depth, D):
    # The matrix-core encoder (its r-block rows spread or consecutive) and
    # the VALU encoder are separate device paths; all must equal the oracle.  D not a multiple of 16 pads the plane
    # words and skips the stores of the extra planes.
    say("encoders", row_len, depth, D)
    rows = 2
    x = synth.binary_polys(600 + row_len + D, rows * row_len, D)
    code = zg.IprsCode(row_len, depth)
    types = zg.BinPolyTypes(D)
    xd = to_dev(x, dev)
    try:
        zg.set_option("iprs_encoder", 0)
        fast = code.encode_rows(xd, rows, types).cpu().numpy()
        zg.set_option("iprs_encoder", 2)
        fast_c = code.encode_rows(xd, rows, types).cpu().numpy()
        zg.set_option("iprs_encoder", 1)
        valu = code.encode_rows(xd, rows, types).cpu().numpy()
    finally:
        zg.set_option("iprs_encoder", -1)
    want = oracle.iprs_encode_binary(x.view(np.uint64), rows, row_len, depth, D)
    assert np.array_equal(fast, want) and np.array_equal(fast_c, want)
    assert np.array_equal(valu, want)


@pytest.mark.parametrize("row_len,depth,rows", [
    (256, 1, 2), (1 << 14, 2, 4), (2048, 3, 1), (128, 0, 1), (1 << 12, -1, 3)])
def test_iprs_i32_encode(zg, dev, oracle, row_len, depth, rows):
    # Depths 1 and 2 take the matrix-core base layer (5 signed digits per
    # i32) by default; wide_base = 1 forces the per-element i128 kernel.
    say("iprs_i32", row_len, depth, rows)
    x = synth.i32_evals(200 + row_len, rows * row_len)
    x[0] = -2**31
    x[-1] = 2**31 - 1
    code = zg.IprsCode(row_len, depth)
    want = oracle.iprs_encode_i32(x, rows, row_len, depth)
    try:
        for opt in (-1, 1):
            zg.set_option("wide_base", opt)
            got = code.encode_rows(to_dev(x, dev), rows, zg.IntTypes()).cpu().numpy()
            assert np.array_equal(got, want), opt
    finally:
        zg.set_option("wide_base", -1)


def test_iprs_extreme_inputs(zg, dev, oracle):
    # All-ones binary rows and i32::MIN/MAX rows give the largest magnitudes.
    row_len = 1 << 14
    ones = np.full(row_len, -1, dtype=np.int64)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    got = code.encode(to_dev(ones, dev), zg.BinPolyTypes(64)).cpu().numpy()
    assert np.array_equal(got, oracle.iprs_encode_binary(ones.view(np.uint64), 1, row_len, -1,
                                                         64)[0])
    for v in (-2**31, 2**31 - 1):
        xi = np.full(row_len, v, dtype=np.int32)
        got = code.encode(to_dev(xi, dev), zg.IntTypes()).cpu().numpy()
        assert np.array_equal(got, oracle.iprs_encode_i32(xi, 1, row_len)[0])


@pytest.mark.parametrize("row_len,depth", [(1 << 15, 2), (12, 1), (1, 1), (24 * 8, 1)])
def test_iprs_invalid_params(zg, row_len, depth):
    # params.rs:59-80, iprs.rs:225-235
    with pytest.raises(zg.InvalidPcsParam):
        zg.IprsCode(row_len, depth)


def test_encode_wrong_length(zg, dev):
    code = zg.IprsCode(256, 1)
    with pytest.raises(zg.ShapeError):
        code.encode(torch.zeros(255, dtype=torch.int64, device=dev), zg.BinPolyTypes(32))


# ---------------------------------------------------------------------------
# RAA encode
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("row_len,D,rows,flip", [
    (4, 32, 2, False), (64, 32, 3, False), (2048, 64, 2, False), (2048, 32, 2, True),
    (1 << 13, 3, 1, False), (1, 8, 4, False), (256, 64, 5, True)])
def test_raa_binary_encode(zg, dev, oracle, row_len, D, rows, flip):
    say("raa", row_len, D, rows, flip)
    x = synth.binary_polys(300 + row_len, rows * row_len, D)
    code = zg.RaaSignFlippingCode(row_len) if flip else zg.RaaCode(row_len)
    got = code.encode_rows(to_dev(x, dev), rows, zg.BinPolyTypes(D)).cpu().numpy()
    want = oracle.raa_encode_binary(x.view(np.uint64), rows, row_len, D, sign_flip=flip)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("row_len,D,rows,flip", [
    (2048, 64, 3, False), (1024, 32, 2, False), (256, 64, 2, False), (4096, 32, 1, False),
    (4096, 64, 2, False), (512, 32, 4, False), (2048, 32, 2, False), (2048, 64, 2, True),
    (256, 32, 3, True), (4096, 64, 1, True), (1024, 32, 2, True), (2048, 3, 2, False),
    (1024, 16, 3, False), (2048, 48, 2, True), (256, 1, 2, True), (4096, 40, 1, False),
    (2048, 17, 2, True)])
def test_raa_device_kernels_agree(zg, dev, oracle, row_len, D, rows, flip):
    # The full-row RAA kernel (compact first scan, 32 planes of a row per
    # workgroup; for the sign-flipping code a second scan over the odd perm1
    # entries) and the plane-group kernel are separate device paths; both
    # must equal the oracle.
    say("raa kernels", row_len, D, rows, flip)
    x = synth.binary_polys(800 + row_len + D, rows * row_len, D)
    code = zg.RaaSignFlippingCode(row_len) if flip else zg.RaaCode(row_len)
    types = zg.BinPolyTypes(D)
    xd = to_dev(x, dev)
    outs = []
    try:
        for variant in (0, 1, 2):  # full row, plane groups of 8 / of 4 planes
            zg.set_option("raa_kernel", variant)
            outs.append(code.encode_rows(xd, rows, types).cpu().numpy())
    finally:
        zg.set_option("raa_kernel", -1)
    want = oracle.raa_encode_binary(x.view(np.uint64), rows, row_len, D, sign_flip=flip)
    for got in outs:
        assert np.array_equal(got, want)


@pytest.mark.parametrize("flip", [False, True])
def test_raa_full_size_kernels_and_linearity(zg, dev, flip):
    # 2^22 BinaryPoly<64>, row 2^11 x 2^11 rows: the full-row kernel equals
    # the plane-group kernel, and for bitwise-disjoint a, b the codewords
    # add (RAA is linear over the integers and every entry fits i32).
    say("raa full size", flip)
    n, row_len, D = 1 << 22, 1 << 11, 64
    x = synth.binary_polys(41, n, D)
    m = synth.binary_polys(42, n, D)
    a, b = x & m, x & ~m
    code = zg.RaaSignFlippingCode(row_len) if flip else zg.RaaCode(row_len)
    types = zg.BinPolyTypes(D)
    rows = n // row_len
    try:
        zg.set_option("raa_kernel", 1)
        groups = code.encode_rows(to_dev(x, dev), rows, types)
        zg.set_option("raa_kernel", 0)
        full = code.encode_rows(to_dev(x, dev), rows, types)
        ea = code.encode_rows(to_dev(a, dev), rows, types)
        eb = code.encode_rows(to_dev(b, dev), rows, types)
    finally:
        zg.set_option("raa_kernel", -1)
    assert torch.equal(full, groups)
    assert torch.equal(ea + eb, full)


def test_raa_width_assert(zg, dev):
    # raa.rs:52-69: 1 + 2*14 + 2*2 = 33 bits do not fit i32 codewords.
    code = zg.RaaCode(1 << 14)
    with pytest.raises(zg.ShapeError):
        code.encode(torch.zeros(1 << 14, dtype=torch.int64, device=dev), zg.BinPolyTypes(32))


def test_raa_known_answer_pattern(zg, dev, oracle):
    # The reference's KAT row [1,2,3,4] (raa.rs:396-413), bit plane by plane:
    # sum_c 2^c * plane_c equals the integer encoding of the row.
    row = np.array([1, 2, 3, 4], dtype=np.int64)
    code = zg.RaaCode(4)
    got = code.encode(to_dev(row, dev), zg.BinPolyTypes(3)).cpu().numpy().astype(np.int64)
    combined = (got * (1 << np.arange(3))).sum(axis=1)
    assert list(combined) == [0x1E, 0x36, 0x39, 0x5A, 0x70, 0x7E, 0xA5, 0xC1, 0xCB, 0xDC, 0xF9,
                              0x11E, 0x124, 0x14C, 0x14D, 0x160]


# ---------------------------------------------------------------------------
# Merkle
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("rows,width,elem", [
    (1, 1, 4), (4, 16, 16), (64, 1024, 256), (5, 64, 256), (3, 8, 24), (2048, 64, 256),
    (1, 4, 8), (17, 32, 12), (3, 16, 1024), (12, 8, 512), (7, 4, 320), (4, 2, 64)])
def test_merkle_root(zg, dev, oracle, rows, width, elem):
    say("merkle", rows, width, elem)
    m = synth.splitmix64(rows * 1000 + width, rows * width * elem // 8 + 1).view(np.uint8)
    m = m[:rows * width * elem].reshape(rows, width, elem)
    got = zg.merkle_root(to_dev(m, dev), elem)
    assert got == oracle.merkle_root(m, rows, width, elem)


# Columns of 4 to 2048 chunks (the register-fold chunk tree with 1, 2, 4 or
# 8 chunks per lane and 4 to 256 lanes per column; one with a partial last
# chunk, one with 200-byte elements, one single column) and of 2, 5, 6 and
# 4096 chunks (the LDS tree kernel or the per-level kernel in both
# settings).
@pytest.mark.parametrize("rows,width,elem", [
    (16, 64, 256), (32, 64, 256), (64, 128, 256), (128, 32, 256), (15, 16, 256), (20, 8, 200),
    (64, 1, 256), (256, 8, 256), (2048, 8, 256), (4096, 4, 256), (8192, 2, 256),
    (4, 8192, 1024), (72, 4, 64), (5, 64, 256), (12, 8, 512), (16384, 2, 256)])
def test_chunk_tree_kernels_agree(zg, dev, oracle, rows, width, elem):
    say("chunk trees", rows, width, elem)
    m = synth.splitmix64(rows * 7 + width + elem, rows * width * elem // 8 + 1).view(np.uint8)
    m = m[:rows * width * elem].reshape(rows, width, elem)
    md = to_dev(m, dev)
    roots = {}
    try:
        for opt in (1, -1):
            zg.set_option("chunk_tree", opt)
            roots[opt] = zg.merkle_root(md, elem)
    finally:
        zg.set_option("chunk_tree", -1)
    want = oracle.merkle_root(m, rows, width, elem)
    assert roots[1] == want and roots[-1] == want


@pytest.mark.parametrize("kind,n,row_len,D", [("iprs", 1 << 20, 1 << 14, 32),
                                              ("raa", 1 << 22, 1 << 11, 64)])
def test_chunk_tree_full_size_settings_agree(zg, dev, kind, n, row_len, D):
    # The benchmarked commits: 16 chunks per column (4 per lane, 4 lanes) and
    # 512 chunks per column (8 per lane, 64 lanes) against the LDS tree.
    say("chunk trees full size", kind, n, D)
    x = to_dev(synth.binary_polys(12, n, D), dev)
    code = zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs" else zg.RaaCode(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    roots = {}
    try:
        for opt in (1, -1):
            zg.set_option("chunk_tree", opt)
            hint, comm = zp.commit(pp, [x])
            roots[opt] = comm.root
            hint.free()
    finally:
        zg.set_option("chunk_tree", -1)
    assert roots[1] == roots[-1]


# ---------------------------------------------------------------------------
# Commit (encode + Merkle) and the hint
# ---------------------------------------------------------------------------

def test_commit_plain_zip_cfg1(zg, dev, oracle):
    # n = 2^16 i32 evals, IPRS rate 1/4, row_len 2^14 x 4 rows (depth 2).
    n, row_len = 1 << 16, 1 << 14
    x = synth.i32_evals(1, n)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    zp = zg.ZipPlus(zg.IntTypes())
    pp = zp.setup(n, code)
    hint, comm = zp.commit(pp, [to_dev(x, dev)])
    ref = oracle.commit("iprs", x, pp.num_rows, row_len, eval_type="i32", want_cw=True)
    assert comm.root == ref["root"] and comm.batch_size == 1
    assert hint.codewords(0).tobytes() == ref["cw"].tobytes()


@pytest.mark.parametrize("kind,n,row_len,D,batch", [
    ("iprs", 1 << 16, 1 << 14, 32, 1), ("iprs", 1 << 12, 256, 64, 2), ("raa", 1 << 14, 128, 64, 1),
    ("raa", 1 << 12, 64, 32, 5), ("raa_signflip", 1 << 12, 64, 32, 1),
    ("iprs", 1 << 10, 256, 3, 3)])
def test_commit_binary(zg, dev, oracle, kind, n, row_len, D, batch):
    say("commit", kind, n, row_len, D, batch)
    x = synth.binary_polys(400 + n + D, batch * n, D)
    code = {"iprs": lambda: zg.IprsCode.new_with_optimal_depth(row_len),
            "raa": lambda: zg.RaaCode(row_len),
            "raa_signflip": lambda: zg.RaaSignFlippingCode(row_len)}[kind]()
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    polys = [to_dev(x[b * n:(b + 1) * n], dev) for b in range(batch)]
    hint, comm = zp.commit(pp, polys)
    ref = oracle.commit(kind, x.view(np.uint64), pp.num_rows, row_len, batch=batch, D=D,
                        want_cw=True, want_layers=True)
    assert comm.root == ref["root"]
    per = ref["cw"].reshape(batch, -1)
    for b in range(batch):
        assert hint.codewords(b).tobytes() == per[b].tobytes()
    assert np.array_equal(hint.layer(0), ref["layers"][:code.codeword_len()])


def test_commit_zip_plus_cfg2_full(zg, dev, oracle):
    # The headline workload at full size: 2^20 BinaryPoly<32>, IPRS row 2^14
    # x 64 rows (depth 2); every codeword byte and the root.
    n, row_len, D = 1 << 20, 1 << 14, 32
    x = synth.binary_polys(2, n, D)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    hint, comm = zp.commit(pp, [to_dev(x, dev)])
    ref = oracle.commit("iprs", x.view(np.uint64), pp.num_rows, row_len, D=D, want_cw=True)
    assert comm.root == ref["root"]
    assert hint.codewords(0).tobytes() == ref["cw"].tobytes()


@pytest.mark.parametrize("kind,n,row_len,D", [
    ("iprs", 1 << 18, 1 << 14, 32), ("iprs", 1 << 16, 1 << 12, 64), ("raa", 1 << 16, 256, 32)])
def test_commit_overlap_matches_sequential(zg, dev, oracle, kind, n, row_len, D):
    # The two-stream commit (hash finished row groups while later rows
    # encode) and the sequential one give the same root as the oracle.
    x = synth.binary_polys(700 + n, n, D)
    code = (zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs"
            else zg.RaaCode(row_len))
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    xd = to_dev(x, dev)
    try:
        zg.set_option("commit_overlap", 1)
        _, on = zp.commit(pp, [xd])
        zg.set_option("commit_overlap", 0)
        _, off = zp.commit(pp, [xd])
    finally:
        zg.set_option("commit_overlap", -1)
    ref = oracle.commit(kind, x.view(np.uint64), pp.num_rows, row_len, D=D)
    assert on.root == off.root == ref["root"]


@pytest.mark.parametrize("kind,n,row_len,D,batch", [
    ("iprs", 1 << 16, 1 << 12, 32, 1), ("iprs", 1 << 14, 1 << 10, 64, 2),
    ("raa", 1 << 16, 256, 32, 1)])
def test_encode_timing_inside_commits(zg, dev, oracle, kind, n, row_len, D, batch):
    # Commits recorded with encoder timing events give the oracle's roots,
    # three events per commit with positive spans; commits outside
    # the block are not recorded.
    xs = [synth.binary_polys(900 + n + i, n, D) for i in range(batch)]
    code = (zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs"
            else zg.RaaCode(row_len))
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    xd = [to_dev(x, dev) for x in xs]
    _, before = zp.commit(pp, xd)
    with zg.EncodeTiming(3) as et:
        roots = [zp.commit(pp, xd)[1].root for _ in range(4)]  # the 4th is past the pool
    with zg.EncodeTiming(8, every=3) as es:  # commits 1, 4 and 7 of 7
        roots += [zp.commit(pp, xd)[1].root for _ in range(7)]
    _, after = zp.commit(pp, xd)
    assert et.launches == 3 and et.total_ms >= et.last_ms > 0
    assert es.launches == 3 and es.total_ms >= es.last_ms > 0
    if batch == 1:
        ref = oracle.commit(kind, xs[0].view(np.uint64), pp.num_rows, row_len, D=D)["root"]
    else:
        ref = before.root
    assert roots == [ref] * 11 and before.root == after.root == ref


@pytest.mark.parametrize("kind,n,row_len,D,batch,eval_type,pinned", [
    ("iprs", 1 << 20, 1 << 14, 32, 1, "binpoly", True),   # cfg2: 64 rows, four groups of 16
    ("iprs", 1 << 20, 1 << 14, 32, 1, "binpoly", False),
    ("iprs", 1 << 13, 1 << 10, 64, 2, "binpoly", False),  # 8 rows per poly, two polys
    ("raa", 1 << 16, 256, 32, 1, "binpoly", True),
    ("raa_signflip", 1 << 12, 64, 16, 3, "binpoly", False),
    ("iprs", 1 << 15, 1 << 14, 1, 1, "i32", True),          # 2 rows: one group
    ("iprs", 1 << 12, 1 << 12, 1, 1, "i32", False)])        # 1 row
def test_commit_from_host_memory(zg, dev, oracle, kind, n, row_len, D, batch, eval_type, pinned):
    # zp_commit_host (row groups copied on a second stream while earlier
    # groups encode) equals zp_commit of the same bytes on the device, and
    # the oracle.
    say("commit_host", kind, n, row_len, D, batch, eval_type, pinned)
    if eval_type == "binpoly":
        x = synth.binary_polys(600 + n + D, batch * n, D)
        types, ox = zg.BinPolyTypes(D), x.view(np.uint64)
    else:
        x = synth.i32_evals(600 + n, batch * n)
        types, ox = zg.IntTypes(), x
    code = {"iprs": lambda: zg.IprsCode.new_with_optimal_depth(row_len),
            "raa": lambda: zg.RaaCode(row_len),
            "raa_signflip": lambda: zg.RaaSignFlippingCode(row_len)}[kind]()
    zp = zg.ZipPlus(types)
    pp = zp.setup(n, code)
    host = [torch.from_numpy(np.ascontiguousarray(x[b * n:(b + 1) * n])) for b in range(batch)]
    if pinned:
        host = [t.pin_memory() for t in host]
    assert not host[0].is_cuda
    hh, hc = zp.commit(pp, host)
    dh, dc = zp.commit(pp, [t.to(dev) for t in host])
    ref = oracle.commit(kind, ox, pp.num_rows, row_len, batch=batch, D=D, eval_type=eval_type)
    assert hc.root == dc.root == ref["root"] and hc.batch_size == batch
    for b in range(batch):
        assert hh.codewords(b).tobytes() == dh.codewords(b).tobytes()
    # The host buffers may be reused at once: a second commit of changed
    # inputs gives the changed root.
    host[0][:8] ^= 1
    _, hc2 = zp.commit(pp, host)
    _, dc2 = zp.commit(pp, [t.to(dev) for t in host])
    assert hc2.root == dc2.root != hc.root


def test_hint_openings_verify(zg, dev, oracle):
    n, row_len, D = 1 << 12, 256, 32
    x = synth.binary_polys(77, 2 * n, D)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    hint, comm = zp.commit(pp, [to_dev(x[:n], dev), to_dev(x[n:], dev)])
    N = code.codeword_len()
    for col in [0, 1, 17, N - 1]:
        proof = hint.prove(col)
        assert len(proof.siblings) == N.bit_length() - 1
        column = hint.open_column(col)
        assert oracle.merkle_verify(comm.root, column, col, N, proof.siblings)
        bad = bytearray(column)
        bad[-1] ^= 0x80
        assert not oracle.merkle_verify(comm.root, bytes(bad), col, N, proof.siblings)
    with pytest.raises(zg.InvalidPcsParam):
        hint.prove(N)


def test_commit_shape_errors(zg, dev):
    code = zg.IprsCode(256, 1)
    zp = zg.ZipPlus(zg.BinPolyTypes(32))
    pp = zp.setup(1 << 10, code)
    with pytest.raises(zg.ShapeError):
        zp.commit(pp, [])
    with pytest.raises(zg.ShapeError):
        zp.commit(pp, [torch.zeros(512, dtype=torch.int64, device=dev)])


def test_commit_linearity_full_size(zg, dev):
    # Size-independent property at the headline size: for bitwise-disjoint
    # inputs a, b the IPRS codewords add, enc(a | b) = enc(a) + enc(b).
    n, row_len, D = 1 << 20, 1 << 14, 32
    x = synth.binary_polys(8, n, D)
    m = synth.binary_polys(9, n, D)
    a, b = x & m, x & ~m & ((1 << D) - 1)
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    types = zg.BinPolyTypes(D)
    rows = n // row_len
    ea = code.encode_rows(to_dev(a, dev), rows, types)
    eb = code.encode_rows(to_dev(b, dev), rows, types)
    ex = code.encode_rows(to_dev(a | b, dev), rows, types)
    assert torch.equal(ea + eb, ex)


# ---------------------------------------------------------------------------
# Combined row
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("rows,row_len,D,batch,L", [
    (256, 256, 64, 1, 5), (64, 128, 32, 2, 5), (1, 64, 64, 1, 5), (16, 32, 3, 3, 4),
    (2048, 64, 64, 1, 5)])
def test_combined_row(zg, dev, oracle, rows, row_len, D, batch, L):
    say("combined_row", rows, row_len, D, batch, L)
    n = rows * row_len
    x = synth.binary_polys(500 + rows, batch * n, D)
    alphas = synth.i128_words(501, batch * D)
    coeffs = synth.i128_words(502, rows) if rows > 1 else np.array([[1, 0]], dtype=np.int64)
    code = zg.RaaCode(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D, comb_limbs=L))
    pp = zp.setup(n, code)
    polys = [to_dev(x[b * n:(b + 1) * n], dev) for b in range(batch)]
    got = zp.combined_row(pp, polys, to_dev(alphas, dev), to_dev(coeffs, dev)).cpu().numpy()
    want = oracle.combined_row(x.view(np.uint64), alphas, coeffs, rows, row_len, batch, D, L)
    assert np.array_equal(got.view(np.uint64), want)


@pytest.mark.parametrize("sign", [-1, 1])
def test_combined_row_extreme_coefficients(zg, dev, oracle, sign):
    # Every bit set and every s_j = i128::MIN (or i128::MAX) for 256 rows:
    # each 128-row segment sums 128 extreme words in every plane.
    say("combined_row extreme", sign)
    rows, row_len, D, L = 256, 64, 64, 5
    n = rows * row_len
    x = np.full(n, -1, dtype=np.int64)
    coeffs = np.zeros((rows, 2), dtype=np.int64)
    if sign < 0:
        coeffs[:, 1] = np.iinfo(np.int64).min
    else:
        coeffs[:, 0] = -1
        coeffs[:, 1] = np.iinfo(np.int64).max
    alphas = synth.i128_words(503, D)
    zp = zg.ZipPlus(zg.BinPolyTypes(D, comb_limbs=L))
    pp = zp.setup(n, zg.RaaCode(row_len))
    got = zp.combined_row(pp, [to_dev(x, dev)], to_dev(alphas, dev), to_dev(coeffs, dev))
    want = oracle.combined_row(x.view(np.uint64), alphas, coeffs, rows, row_len, 1, D, L)
    assert np.array_equal(got.cpu().numpy().view(np.uint64), want)


@pytest.mark.parametrize("kind,n,row_len,D", [("iprs", 1 << 6, 8, 32), ("iprs", 1 << 8, 64, 32),
                                              ("raa", 1 << 12, 1 << 11, 32),
                                              ("iprs", 1 << 16, 1 << 14, 32),
                                              ("raa", 1 << 13, 1 << 13, 16)])
def test_merkle_layer_kernels_agree(zg, dev, oracle, kind, n, row_len, D):
    # Every layer of the tree (the proofs read them all), from the
    # quad-per-node kernel and the thread-per-node kernel; 32 to 2^16 leaves.
    say("merkle layers", kind, n, row_len, D)
    x = synth.binary_polys(31 + n, n, D)
    code = zg.IprsCode(row_len, -1) if kind == "iprs" else zg.RaaCode(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    layers = {}
    try:
        for opt in (1, 0):
            zg.set_option("merkle_tree", opt)
            hint, comm = zp.commit(pp, [to_dev(x, dev)])
            layers[opt] = ([hint.layer(lv) for lv in range(hint.height)], comm.root)
            hint.free()
    finally:
        zg.set_option("merkle_tree", -1)
    assert layers[0][1] == layers[1][1]
    assert len(layers[0][0]) == len(layers[1][0]) and len(layers[0][0]) > 1
    for a, b in zip(layers[0][0], layers[1][0]):
        assert np.array_equal(a, b)
    want = oracle.commit(kind, x.view(np.uint64), pp.num_rows, row_len, D=D)["root"]
    assert layers[0][1] == want
