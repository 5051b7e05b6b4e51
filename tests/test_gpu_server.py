"""The commit/open service: every result is the same alone and in the mix,
and equal to the oracle's (roots, combined rows, columns, Merkle paths)."""
import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def requests(zg, seed, count, lo, hi):
    out = []
    codes = {}
    for r in synth.request_mix(seed, count, lo, hi):
        key = (r["kind"], r["row_len"])
        if key not in codes:
            codes[key] = (zg.IprsCode.new_with_optimal_depth(r["row_len"]) if r["kind"] == "iprs"
                          else zg.RaaCode(r["row_len"]))
        types = zg.BinPolyTypes(r["D"], 5)
        pp = zg.ZipPlus(types).setup(r["n"], codes[key])
        x = synth.binary_polys(r["seed"], r["n"], r["D"])
        N = codes[key].codeword_len()
        al = synth.i128_words(r["seed"] + 1, r["D"])
        co = synth.i128_words(r["seed"] + 2, pp.num_rows)
        cols = (synth.splitmix64(r["seed"] + 3, 16) % np.uint64(N)).astype(np.uint64)
        out.append(dict(r, pp=pp, types=types, x=x, al=al, co=co, cols=cols, N=N))
    return out


def run_alone(srv, reqs):
    res = []
    for q in reqs:
        rid = srv.commit(q["pp"], q["types"], q["x"])
        c = srv.poll(60)
        assert c and c["request"] == rid and c["kind"] == "commit" and c["status"] == 0, c
        srv.open(rid, q["al"], q["co"], q["cols"])
        o = srv.poll(60)
        assert o and o["request"] == rid and o["kind"] == "open" and o["status"] == 0, o
        res.append((c["root"], o["comb"].copy(), o["columns"], o["siblings"]))
    return res


def run_mixed(srv, reqs):
    ids = {srv.commit(q["pp"], q["types"], q["x"]): i for i, q in enumerate(reqs)}
    res, roots = [None] * len(reqs), {}
    pending = 2 * len(reqs)
    while pending:
        ev = srv.poll(120)
        assert ev is not None, "server stalled"
        assert ev["status"] == 0, ev
        i = ids[ev["request"]]
        if ev["kind"] == "commit":
            roots[i] = ev["root"]
            q = reqs[i]
            srv.open(ev["request"], q["al"], q["co"], q["cols"])
        else:
            res[i] = (roots[i], ev["comb"].copy(), ev["columns"], ev["siblings"])
        pending -= 1
    return res


def test_server_alone_equals_mixed_equals_oracle(zg, dev, oracle):
    reqs = requests(zg, 11, 12, 12, 16)
    srv = zg.CommitServer(slots=4)
    say("alone")
    alone = run_alone(srv, reqs)
    say("mixed")
    mixed = run_mixed(srv, reqs)
    st = srv.stats()
    assert st["hbm_in_use"] == 0 and st["pending"] == 0
    srv.close()
    for i, q in enumerate(reqs):
        say("check", i, q["kind"], q["n"], q["D"])
        a, m = alone[i], mixed[i]
        assert a[0] == m[0] and np.array_equal(a[1], m[1]) and a[2] == m[2] and a[3] == m[3]
        pp, D = q["pp"], q["D"]
        ref = oracle.commit(q["kind"], q["x"].view(np.uint64), pp.num_rows, q["row_len"], D=D,
                            want_cw=True)
        assert a[0] == ref["root"]
        comb = oracle.combined_row(q["x"].view(np.uint64), q["al"], q["co"], pp.num_rows,
                                   q["row_len"], 1, D, 5)
        assert np.array_equal(a[1], comb)
        elem = (8 if q["kind"] == "iprs" else 4) * D
        cw = ref["cw"].reshape(pp.num_rows, q["N"], elem)
        for k, col in enumerate(q["cols"]):
            assert a[2][k] == cw[:, int(col), :].tobytes()
            assert oracle.merkle_verify(a[0], a[2][k], int(col), q["N"], a[3][k])


def test_server_admission_under_small_budget(zg, dev, oracle):
    reqs = requests(zg, 12, 6, 14, 16)
    # A budget below two requests' footprint: commits are admitted one or two
    # at a time, openings free the memory, and every request still finishes.
    srv = zg.CommitServer(hbm_budget=96 << 20, slots=4)
    mixed = run_mixed(srv, reqs)
    st = srv.stats()
    assert st["hbm_in_use"] == 0
    srv.close()
    for i, q in enumerate(reqs):
        ref = oracle.commit(q["kind"], q["x"].view(np.uint64), q["pp"].num_rows, q["row_len"],
                            D=q["D"])
        assert mixed[i][0] == ref["root"]


def test_server_device_inputs_and_errors(zg, dev, oracle):
    reqs = requests(zg, 13, 3, 12, 13)
    srv = zg.CommitServer(slots=2)
    xd = [torch.from_numpy(q["x"]).to(dev) for q in reqs]
    ids = [srv.commit(q["pp"], q["types"], x) for q, x in zip(reqs, xd)]
    with pytest.raises(zg.InvalidPcsParam):
        srv.open(12345, None, None, [])  # unknown request
    roots = {}
    for _ in ids:
        ev = srv.poll(60)
        assert ev["status"] == 0
        roots[ev["request"]] = ev["root"]
    for rid, q in zip(ids, reqs):
        ref = oracle.commit(q["kind"], q["x"].view(np.uint64), q["pp"].num_rows, q["row_len"],
                            D=q["D"])
        assert roots[rid] == ref["root"]
        with pytest.raises(zg.InvalidPcsParam):
            srv.open(rid, None, None, [q["N"]])  # column out of range
        srv.open(rid, None, None, [])  # release without opening anything
    for _ in ids:
        ev = srv.poll(60)
        assert ev["kind"] == "open" and ev["status"] == 0
    assert srv.poll(0.01) is None
    assert srv.stats()["hbm_in_use"] == 0
    srv.close()


def _point(seed, nv):
    w = synth.i128_words(seed, nv)
    return [int(np.uint64(w[i, 0].view(np.uint64))) | (int(w[i, 1]) << 64) for i in range(nv)]


def test_server_proofs_equal_direct_prove(zg, dev):
    # ZipPlus::prove run by the service (transcript from the root, field
    # drawn, proof written on a worker stream) equals the direct path
    # (zp_commit + PcsProverTranscript + zp_prove) byte for byte, for every
    # request of a concurrent mix.
    reqs = requests(zg, 23, 10, 12, 16)
    srv = zg.CommitServer(slots=4)
    ids = {srv.commit(q["pp"], q["types"], q["x"]): i for i, q in enumerate(reqs)}
    got, pending = {}, 2 * len(reqs)
    while pending:
        ev = srv.poll(120)
        assert ev is not None, "server stalled"
        assert ev["status"] == 0, ev
        i = ids[ev["request"]]
        if ev["kind"] == "commit":
            q = reqs[i]
            srv.prove(ev["request"], _point(100 + i, q["n"].bit_length() - 1))
        else:
            assert ev["kind"] == "prove"
            got[i] = ev["proof"]
        pending -= 1
    st = srv.stats()
    assert st["hbm_in_use"] == 0 and st["pending"] == 0
    srv.close()
    for i, q in enumerate(reqs):
        say("direct prove", i, q["kind"], q["n"], q["D"])
        zp = zg.ZipPlus(q["types"])
        x = torch.from_numpy(q["x"]).to(dev)
        hint, comm = zp.commit(q["pp"], [x])
        t = zg.PcsProverTranscript.new_from_commitment(comm)
        p = t.get_random_field_cfg()
        e = zp.prove(t, q["pp"], [x], _point(100 + i, q["n"].bit_length() - 1), hint, p)
        assert got[i] == (p, e, t.stream())
        hint.free()
