"""Batched commitment plans and the autotuner, against the CPU oracle.

Every root of a plan must equal the oracle's ZipPlus::commit of that item,
on the first (eager, measuring) commit, on the commit that captures the HIP
graph, and on later replays -- with ordinary eager commits run between the
replays -- and again after the inputs move (which recaptures the graph).
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


# (kind, n, D, eval_type, batch): IPRS matrix-core and VALU shapes, i32
# evals, RAA full-row and plane-group shapes, the sign-flipping code, and an
# item with several polys under one root.
ITEMS = [
    ("iprs", 1 << 12, 32, "binpoly", 1),
    ("iprs", 1 << 15, 64, "binpoly", 1),
    ("raa", 1 << 12, 32, "binpoly", 1),
    ("raa", 1 << 14, 64, "binpoly", 1),
    ("iprs", 1 << 11, 0, "i32", 1),
    ("raa_signflip", 1 << 12, 64, "binpoly", 1),
    ("iprs", 1 << 10, 7, "binpoly", 3),
    ("raa", 1 << 16, 32, "binpoly", 1),
]


def make_code(zg, kind, row_len):
    if kind == "iprs":
        return zg.IprsCode.new_with_optimal_depth(row_len)
    return zg.RaaCode(row_len) if kind == "raa" else zg.RaaSignFlippingCode(row_len)


def build(zg, items):
    plan_items, meta = [], []
    for kind, n, D, et, batch in items:
        row_len = synth.row_split("iprs" if kind == "iprs" else "raa", n)
        if kind == "iprs":
            row_len = min(row_len, 1 << 12)  # several rows per tree
        types = zg.IntTypes() if et == "i32" else zg.BinPolyTypes(D)
        code = make_code(zg, kind, row_len)
        pp = zg.ZipPlus(types).setup(n, code)
        plan_items.append((pp, types, batch))
        meta.append((kind, n, D, et, batch, row_len, pp.num_rows))
    return plan_items, meta


def inputs(meta, seed):
    out = []
    for i, (kind, n, D, et, batch, _, _) in enumerate(meta):
        s = seed * 100 + i
        out.append(synth.i32_evals(s, batch * n) if et == "i32"
                   else synth.binary_polys(s, batch * n, D))
    return out


def oracle_roots(oracle, meta, xs):
    roots = []
    for (kind, n, D, et, batch, row_len, rows), x in zip(meta, xs):
        ev = x if et == "i32" else x.view(np.uint64)
        roots.append(oracle.commit(kind, ev, rows, row_len, batch=batch, D=max(D, 1),
                                   eval_type=et)["root"])
    return roots


def eager_noise(zg, dev):
    # Ordinary work between graph replays: a plain commit and a torch op.
    x = synth.binary_polys(5, 1 << 12, 32)
    zp = zg.ZipPlus(zg.BinPolyTypes(32))
    pp = zp.setup(1 << 12, zg.RaaCode(64))
    hint, _ = zp.commit(pp, [to_dev(x, dev)])
    hint.free()
    torch.ones(1 << 20, device=dev).cumsum(0)
    torch.cuda.synchronize()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_commit_matches_oracle(zg, dev, oracle, graph):
    say("batch plan", "graph" if graph else "eager")
    plan_items, meta = build(zg, ITEMS)
    plan = zg.BatchCommit(plan_items, graph=graph)
    xs = inputs(meta, 1)
    want = oracle_roots(oracle, meta, xs)
    xd = [to_dev(x, dev) for x in xs]
    for rnd in range(4):
        say("commit", rnd)
        got = plan.commit(xd)
        assert [c.root for c in got] == want, f"commit {rnd}"
        assert [c.batch_size for c in got] == [m[4] for m in meta]
        eager_noise(zg, dev)
    info = plan.info()
    assert info["graph_ready"] == graph
    assert info["hint_bytes"] > 0
    # New inputs at new addresses: the graph is captured again.
    xs2 = inputs(meta, 2)
    want2 = oracle_roots(oracle, meta, xs2)
    xd2 = [to_dev(x, dev) for x in xs2]
    for rnd in range(2):
        got = plan.commit(xd2)
        assert [c.root for c in got] == want2, f"moved inputs, commit {rnd}"
    # Back to the first inputs, contents changed in place.
    for i, x in enumerate(xs2):
        xd[i].copy_(to_dev(x, dev))
    got = plan.commit(xd)
    assert [c.root for c in got] == want2


def test_batch_hints_match_oracle(zg, dev, oracle):
    items = [("iprs", 1 << 13, 32, "binpoly", 1), ("raa", 1 << 12, 64, "binpoly", 2)]
    plan_items, meta = build(zg, items)
    plan = zg.BatchCommit(plan_items, graph=True)
    xs = inputs(meta, 3)
    xd = [to_dev(x, dev) for x in xs]
    for _ in range(3):
        plan.commit(xd)
    for i, ((kind, n, D, et, batch, row_len, rows), x) in enumerate(zip(meta, xs)):
        say("hint", i)
        ref = oracle.commit(kind, x.view(np.uint64), rows, row_len, batch=batch, D=D,
                            want_cw=True, want_layers=True)
        hint = plan.hint(i)
        per = ref["cw"].reshape(batch, -1)
        for b in range(batch):
            assert hint.codewords(b).tobytes() == per[b].tobytes()
        N = hint.num_leaves
        assert np.array_equal(hint.layer(0), ref["layers"][:N])
        col = N // 3
        proof = hint.prove(col)
        assert len(proof.siblings) == N.bit_length() - 1
        assert oracle.merkle_verify(ref["root"], hint.open_column(col), col, N,
                                    proof.siblings)
        hint.free()  # a plan's hint is not freed by its holder
    got = plan.commit(xd)
    assert got[0].root == oracle_roots(oracle, meta, xs)[0]


def test_batch_plan_errors(zg, dev):
    with pytest.raises(zg.ShapeError):
        zg.BatchCommit([])
    plan_items, meta = build(zg, ITEMS[:2])
    plan = zg.BatchCommit(plan_items)
    with pytest.raises(zg.ShapeError):
        plan.commit([torch.zeros(8, dtype=torch.int64, device=dev)] * 2)
    with pytest.raises(zg.ShapeError):
        plan.commit([])


@pytest.mark.parametrize("kind,row_len,D,rows", [
    ("iprs", 1 << 14, 32, 8), ("iprs", 1 << 12, 64, 4), ("raa", 2048, 64, 16),
    ("raa", 256, 32, 8)])
def test_autotune_keeps_results(zg, dev, oracle, kind, row_len, D, rows):
    say("autotune", kind, row_len, D, rows)
    code = make_code(zg, kind, row_len)
    types = zg.BinPolyTypes(D)
    assert code.tuning() == {"iprs_encoder": -1, "raa_kernel": -1, "commit_overlap": -1,
                             "chunk_tree_fold": -1}
    t = code.autotune(types, rows)
    if kind == "iprs":
        assert t["iprs_encoder"] in (0, 2) and t["raa_kernel"] == -1  # matrix cores
    else:
        assert t["raa_kernel"] in (0, 1, 2) and t["iprs_encoder"] == -1
    assert t["commit_overlap"] in (0, 1)
    assert t["chunk_tree_fold"] in (-1, 1, 2, 4, 8)
    n = rows * row_len
    x = synth.binary_polys(900 + row_len, n, D)
    xd = to_dev(x, dev)
    got = code.encode_rows(xd, rows, types).cpu().numpy()
    if kind == "iprs":
        want = oracle.iprs_encode_binary(x.view(np.uint64), rows, row_len, -1, D)
    else:
        want = oracle.raa_encode_binary(x.view(np.uint64), rows, row_len, D)
    assert np.array_equal(got, want)
    zp = zg.ZipPlus(types)
    _, comm = zp.commit(zp.setup(n, code), [xd])
    assert comm.root == oracle.commit(kind, x.view(np.uint64), rows, row_len, D=D)["root"]


def test_batch_autotune_flag(zg, dev, oracle):
    plan_items, meta = build(zg, ITEMS[:4])
    plan = zg.BatchCommit(plan_items, graph=True, autotune=True)
    for pp, _, _ in plan_items:
        assert pp.linear_code.tuning()["commit_overlap"] in (0, 1)
    xs = inputs(meta, 4)
    xd = [to_dev(x, dev) for x in xs]
    want = oracle_roots(oracle, meta, xs)
    for _ in range(3):
        assert [c.root for c in plan.commit(xd)] == want


def test_batch_i64_and_dense_items_match_plain_commits(zg, dev):
    # The i64 and DensePolynomial<i64, D> items (the protocol's int and
    # arbitrary traces) encode through the Int<L> kernels, whose scratch
    # (digit fragments, the RAA middle row) comes from the capture arena in
    # a graph.  Their roots equal plain commits (themselves checked against
    # the oracle in test_gpu_i64.py) on capture and on every replay.
    say("batch i64/dense")
    rng = np.random.default_rng(41)
    specs = [("iprs", zg.I64Types(), 1 << 14, 2), ("iprs", zg.DensePolyTypes(16), 1 << 12, 1),
             ("raa", zg.I64Types(), 1 << 12, 1), ("raa", zg.DensePolyTypes(8), 1 << 10, 2),
             ("iprs", zg.BinPolyTypes(32), 1 << 12, 1)]
    items, xs = [], []
    for kind, types, n, batch in specs:
        row_len = min(n, 1 << 12) if kind == "iprs" else synth.row_split("raa", n)
        code = make_code(zg, kind, row_len)
        items.append((zg.ZipPlus(types).setup(n, code), types, batch))
        words = batch * n * types.eval_words
        if types.eval_type == "binpoly":
            xs.append(to_dev(synth.binary_polys(7, batch * n, 32), dev))
        else:
            xs.append(to_dev(rng.integers(-2**40, 2**40, words, dtype=np.int64), dev))
    want = []
    for (pp, types, batch), x in zip(items, xs):
        polys = list(x.reshape(batch, -1))
        h, c = zg.ZipPlus(types).commit(pp, polys)
        h.free()
        want.append(c.root)
    plan = zg.BatchCommit(items, graph=True)
    for rep in range(3):
        assert [c.root for c in plan.commit(xs)] == want, rep
        eager_noise(zg, dev)


# Leaf shapes of the grouped hashing: one chunk per column (aligned and
# unaligned elements), two chunks (the pair kernel), 3 chunks (not a power
# of two: the LDS tree kernel), the quad-tree folds, and 8192 chunks (above
# 2048: one launch per level).  (kind, n, D, eval_type, batch, row_len).
GROUPED_ITEMS = [
    ("iprs", 1 << 12, 32, "binpoly", 1, 1 << 12),   # 1 chunk, aligned
    ("iprs", 1 << 10, 7, "binpoly", 3, 1 << 10),    # 1 chunk, 56-byte elements
    ("iprs", 1 << 13, 32, "binpoly", 4, 1 << 12),   # 2 chunks
    ("iprs", 1 << 14, 32, "binpoly", 3, 1 << 12),   # 3 chunks
    ("iprs", 1 << 15, 64, "binpoly", 1, 1 << 12),   # 4 chunks
    ("raa", 1 << 14, 64, "binpoly", 1, 128),        # 32 chunks
    ("raa", 1 << 16, 64, "binpoly", 1, 8),          # 8192 chunks
    ("iprs", 1 << 11, 0, "i32", 1, 1 << 11),        # 16-byte elements
    ("raa_signflip", 1 << 12, 64, "binpoly", 1, 64),
]


def build_rows(zg, items):
    plan_items, meta = [], []
    for kind, n, D, et, batch, row_len in items:
        types = zg.IntTypes() if et == "i32" else zg.BinPolyTypes(D)
        pp = zg.ZipPlus(types).setup(n, make_code(zg, kind, row_len))
        plan_items.append((pp, types, batch))
        meta.append((kind, n, D, et, batch, row_len, pp.num_rows))
    return plan_items, meta


@pytest.mark.parametrize("schedule", [(True, 4), (False, 4), (True, 1), (False, 8), (True, 8)])
def test_batch_schedules_match_oracle(zg, dev, oracle, schedule):
    say("schedule", schedule)
    plan_items, meta = build_rows(zg, GROUPED_ITEMS)
    plan = zg.BatchCommit(plan_items, graph=True, n_streams=8)
    plan.set_schedule(*schedule)
    assert plan.schedule == schedule
    xs = inputs(meta, 11)
    want = oracle_roots(oracle, meta, xs)
    xd = [to_dev(x, dev) for x in xs]
    for rnd in range(3):  # measure, capture, replay
        got = plan.commit(xd)
        assert [c.root for c in got] == want, f"commit {rnd}"
        eager_noise(zg, dev)


def test_batch_many_items_grouped(zg, dev, oracle):
    # 60 items: more than one group (48 items) per grouped launch.
    items = []
    for i in range(60):
        kind = ("iprs", "raa", "raa_signflip")[i % 3]
        n = 1 << (10 + i % 4)
        D = (32, 64, 16)[i % 3]
        row_len = min(n, 1 << 10) if kind == "iprs" else 32
        items.append((kind, n, D, "binpoly", 1 + (i % 5 == 0), row_len))
    plan_items, meta = build_rows(zg, items)
    plan = zg.BatchCommit(plan_items, graph=True)
    assert plan.schedule[0] is True
    xs = inputs(meta, 12)
    want = oracle_roots(oracle, meta, xs)
    xd = [to_dev(x, dev) for x in xs]
    for rnd in range(3):
        got = plan.commit(xd)
        assert [c.root for c in got] == want, f"commit {rnd}"


def test_batch_schedule_autotune(zg, dev, oracle):
    plan_items, meta = build_rows(zg, GROUPED_ITEMS[:6])
    plan = zg.BatchCommit(plan_items, graph=True, autotune=True, n_streams=4)
    xs = inputs(meta, 13)
    want = oracle_roots(oracle, meta, xs)
    xd = [to_dev(x, dev) for x in xs]
    for rnd in range(3):
        got = plan.commit(xd)
        assert [c.root for c in got] == want, f"commit {rnd}"
    grouped, streams = plan.schedule
    assert streams in (2, 4, 6) and isinstance(grouped, bool)
    with pytest.raises(zg.ZipError):
        plan.set_schedule(True, 99)


def test_batch_schedule_change_after_capture(zg, dev, oracle):
    # A plan that has measured its arena and captured its graph takes a new
    # schedule: the next commit measures again and captures anew.
    plan_items, meta = build_rows(zg, GROUPED_ITEMS[:5])
    plan = zg.BatchCommit(plan_items, graph=True, n_streams=4)
    xs = inputs(meta, 14)
    want = oracle_roots(oracle, meta, xs)
    xd = [to_dev(x, dev) for x in xs]
    for sched in [None, (False, 2), (True, 3), (False, 1)]:
        if sched is not None:
            plan.set_schedule(*sched)
            assert plan.schedule == sched
        for rnd in range(3):
            got = plan.commit(xd)
            assert [c.root for c in got] == want, f"schedule {sched}, commit {rnd}"
        assert plan.info()["graph_ready"]
