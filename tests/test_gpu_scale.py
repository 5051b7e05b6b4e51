"""Oracle parity at the workloads' configured scale.

* the batched mix exactly as measured: 64 commits, log2 n uniform in
  12..20, D 32/64, IPRS/RAA (synth.request_mix(4, 64, 12, 20)), through one
  BatchCommit plan with a HIP graph, 8 branch streams and the autotuner;
  every root against the oracle's ZipPlus::commit of that item
  (zip-plus/src/pcs/phase_commit.rs:61-101), on the eager commit, the
  capture and a replay;
* a served mix of 2^14..2^22 requests (two 2^22 BinaryPoly<64> RAA commits
  among them, multi-GiB admission) through CommitServer: each request alone
  equals it in the mix equals the oracle, for the root, the combined row
  (phase_prove.rs:153-244) and the opened columns with their Merkle paths;
* the full RAA workload, 2^22 BinaryPoly<64> with rows of 2^11: the root,
  and the combined row at 2^11 x 2^11 x D = 64 in Int<5>.
"""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture
def oracle16(oracle):
    oracle.set_threads(16)
    yield oracle
    oracle.set_threads(8)


def code_for(zg, codes, kind, row_len):
    key = (kind, row_len)
    if key not in codes:
        codes[key] = (zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs"
                      else zg.RaaCode(row_len))
    return codes[key]


def test_batched_mix_at_scale_matches_oracle(zg, dev, oracle16):
    reqs = synth.request_mix(4, 64, 12, 20)
    codes, items, xs = {}, [], []
    for r in reqs:
        types = zg.BinPolyTypes(r["D"])
        pp = zg.ZipPlus(types).setup(r["n"], code_for(zg, codes, r["kind"], r["row_len"]))
        items.append((pp, types))
        xs.append(synth.binary_polys(r["seed"], r["n"], r["D"]))
    say("batched mix", len(reqs), "items,", sum(r["n"] for r in reqs), "evals")
    want = []
    for r, (pp, _), x in zip(reqs, items, xs):
        want.append(oracle16.commit(r["kind"], x.view(np.uint64), pp.num_rows, r["row_len"],
                                    D=r["D"])["root"])
    xd = [torch.from_numpy(x).to(dev) for x in xs]
    plan = zg.BatchCommit(items, graph=True, autotune=True, n_streams=8)
    for rnd in range(3):  # eager (measuring), graph capture, replay
        say("plan commit", rnd)
        got = [c.root for c in plan.commit(xd)]
        bad = [i for i in range(len(reqs)) if got[i] != want[i]]
        assert not bad, f"commit {rnd}: items {bad} differ from the oracle"
    assert plan.info()["graph_ready"]
    del plan


def served_requests(zg, seed, count, lo, hi, n_cols=16):
    codes, out = {}, []
    for r in synth.request_mix(seed, count, lo, hi):
        types = zg.BinPolyTypes(r["D"], 5)
        code = code_for(zg, codes, r["kind"], r["row_len"])
        pp = zg.ZipPlus(types).setup(r["n"], code)
        N = code.codeword_len()
        out.append(dict(r, pp=pp, types=types, N=N,
                        x=synth.binary_polys(r["seed"], r["n"], r["D"]),
                        al=synth.i128_words(r["seed"] + 1, r["D"]),
                        co=synth.i128_words(r["seed"] + 2, pp.num_rows),
                        cols=(synth.splitmix64(r["seed"] + 3, n_cols) % np.uint64(N))))
    return out


def serve_alone(srv, reqs):
    res = []
    for q in reqs:
        rid = srv.commit(q["pp"], q["types"], q["x"])
        c = srv.poll(120)
        assert c and c["request"] == rid and c["status"] == 0, c
        srv.open(rid, q["al"], q["co"], q["cols"])
        o = srv.poll(120)
        assert o and o["request"] == rid and o["kind"] == "open" and o["status"] == 0, o
        res.append((c["root"], o["comb"].copy(), o["columns"], o["siblings"]))
    return res


def serve_mixed(srv, reqs):
    ids = {srv.commit(q["pp"], q["types"], q["x"]): i for i, q in enumerate(reqs)}
    res, roots, pending = [None] * len(reqs), {}, 2 * len(reqs)
    while pending:
        ev = srv.poll(120)
        assert ev is not None and ev["status"] == 0, ev
        i = ids[ev["request"]]
        if ev["kind"] == "commit":
            roots[i] = ev["root"]
            q = reqs[i]
            srv.open(ev["request"], q["al"], q["co"], q["cols"])
        else:
            res[i] = (roots[i], ev["comb"].copy(), ev["columns"], ev["siblings"])
        pending -= 1
    return res


def test_served_mix_2p14_to_2p22_alone_mixed_oracle(zg, dev, oracle16):
    # Seed 22 of the request mix: RAA 2^22 x D 64 twice, IPRS 2^21 x D 32,
    # IPRS 2^19 x D 64, RAA 2^20, 2^17, 2^15 and IPRS 2^14.
    reqs = served_requests(zg, 22, 8, 14, 22)
    assert max(q["n"] for q in reqs) == 1 << 22
    srv = zg.CommitServer(slots=4)
    say("served alone", [(q["kind"], q["n"].bit_length() - 1, q["D"]) for q in reqs])
    alone = serve_alone(srv, reqs)
    say("served mixed")
    mixed = serve_mixed(srv, reqs)
    st = srv.stats()
    assert st["hbm_in_use"] == 0 and st["pending"] == 0
    assert st["hbm_peak"] > (4 << 30)  # a 2^22 x D 64 RAA request holds over 4 GiB
    srv.close()
    for i, q in enumerate(reqs):
        say("oracle", i, q["kind"], q["n"], q["D"])
        a, m = alone[i], mixed[i]
        assert a[0] == m[0] and np.array_equal(a[1], m[1]) and a[2] == m[2] and a[3] == m[3]
        rows, D = q["pp"].num_rows, q["D"]
        ref = oracle16.commit(q["kind"], q["x"].view(np.uint64), rows, q["row_len"], D=D,
                              want_cw=True)
        assert a[0] == ref["root"]
        comb = oracle16.combined_row(q["x"].view(np.uint64), q["al"], q["co"], rows,
                                     q["row_len"], 1, D, 5)
        assert np.array_equal(a[1], comb)
        elem = (8 if q["kind"] == "iprs" else 4) * D
        cw = ref["cw"].reshape(rows, q["N"], elem)
        for k, col in enumerate(q["cols"]):
            assert a[2][k] == cw[:, int(col), :].tobytes()
            assert oracle16.merkle_verify(a[0], a[2][k], int(col), q["N"], a[3][k])
        del ref, cw


def test_raa_2p22_d64_root_and_combined_row(zg, dev, oracle16):
    n, row_len, D = 1 << 22, 1 << 11, 64
    x = synth.binary_polys(303, n, D)
    code = zg.RaaCode(row_len)
    types = zg.BinPolyTypes(D, 5)
    zp = zg.ZipPlus(types)
    pp = zp.setup(n, code)
    assert pp.num_rows == 1 << 11
    say("raa 2^22 D64 commit")
    xd = torch.from_numpy(x).to(dev)
    hint, comm = zp.commit(pp, [xd])
    hint.free()
    want = oracle16.commit("raa", x.view(np.uint64), pp.num_rows, row_len, D=D)["root"]
    assert comm.root == want
    say("combined row 2^11 x 2^11 x 64, Int<5>")
    al = synth.i128_words(304, D)
    co = synth.i128_words(305, pp.num_rows)
    got = zp.combined_row(pp, [xd], torch.from_numpy(al).to(dev),
                          torch.from_numpy(co).to(dev)).cpu().numpy().view(np.uint64)
    ref = oracle16.combined_row(x.view(np.uint64), al, co, pp.num_rows, row_len, 1, D, 5)
    assert np.array_equal(got, ref)
    # Extreme coefficients: every alpha and s at the i128 bounds.
    ext = np.array([[0, -(1 << 63)], [-1, (1 << 63) - 1]], dtype=np.int64)
    al2 = ext[np.arange(D) % 2]
    co2 = ext[(np.arange(pp.num_rows) // 3) % 2]
    got = zp.combined_row(pp, [xd], torch.from_numpy(al2).to(dev),
                          torch.from_numpy(co2).to(dev)).cpu().numpy().view(np.uint64)
    assert np.array_equal(got, oracle16.combined_row(x.view(np.uint64), al2, co2, pp.num_rows,
                                                     row_len, 1, D, 5))
