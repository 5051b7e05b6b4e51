"""Zip+ commit throughput on one MI355X.

Workload (the headline configuration): Zip+ commit of one polynomial with
2^20 evaluations of BinaryPoly<32> (i.i.d. Bernoulli(1/2) coefficient bits in
u64 lanes), IPRS code F65537 rate 1/4, row_len 2^14 x 64 rows (depth 2),
Cw = [i64; 32] (1 GiB of codewords), BLAKE3 column Merkle tree over 65536
leaves of 16 KiB.  One step = one ZipPlus::commit (encode + Merkle) of the
resident input; value = evaluations committed per second.

Also reported on the same line:
  roofline      the dominant kernel (iprs_mfma_kernel, the matrix-core IPRS
                encoder after its bit-plane pre-pass) timed with HIP events
                on the commit stream inside the timed commits; traffic from
                the committed rocprofv3 counter summary when present
  encode_rows   both launches of the encoder (pre-pass included), inside
                the commits and back to back on a stream of its own
  commit_roofline  the floor of the whole commit step: its HBM bytes and its
                VALU operations (BLAKE3 at its per-compression minimum, the
                radix-8 stages' multiply-adds) at the published peaks
  cpu_baseline  the reference's CPU commit of the same workload, restated on
                its parallel axes with its AVX-512 paths (all allotted host
                threads, and one thread)
  workloads     plain Zip 2^16 i32 commit; RAA 2^22 BinaryPoly<64> encode,
                commit and combined row; the sign-flipping RAA code's encode
                at the same shape; 64 mixed commits as one batch plan
                (HIP graph) against the plain per-poly path; the commit/open
                service under Poisson arrivals (p50/p99 latency) against one
                request at a time; both IPRS encoders at the headline shape.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "zinc-plus_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "Zip+ commit evals/s at 2^20 deg-32 IPRS; batched vs plain; p99 latency"
N_EVALS = 1 << 20
ROW_LEN = 1 << 14
D = 32
REP = 4
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of IPRS encode per evaluation: the u64 input lane plus
# REP * D * 8 bytes of [i64; D] codeword (both launches of one encode_rows).
ENCODE_BYTES_PER_EVAL = 8 + REP * D * 8
ENCODE_KERNEL = "bitplane_kernel + iprs_mfma_kernel (one encode_rows)"
# The dominant kernel alone, iprs_mfma_kernel: it reads the bit-plane words
# (D bits of each evaluation, D / 8 bytes; the base length 256 is a whole
# number of 64-wide k-steps, so there is no padding) and writes the codeword.
MAIN_BYTES_PER_EVAL = D // 8 + REP * D * 8
MAIN_KERNEL = "iprs_mfma_kernel (the encoder after its bit-plane pre-pass)"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "iprs_mfma_kernel_commit_2p20_d32_pmc.txt")
# VALU lane-operations per second at the published clock: 256 CUs x 64 lanes
# x 2.4 GHz of int32 or FP64 issue (FP64 vector 78.6 TFLOP/s = FMA at this
# rate).  A BLAKE3 compression that yields a chaining value takes at least
# 680 such operations on gfx950 (7 rounds x 8 G functions x 12: v_add3_u32,
# add, xor + v_alignbit_b32 rotations; then 8 output xors).
VALU_LANE_OPS_PEAK = 256 * 64 * 2.4e9
BLAKE3_OPS_PER_COMPRESSION = 680


def commit_roofline(ms_per_step, num_rows, N):
    """The floor of one whole commit step: HBM bytes (evaluations read,
    codewords written by the encoder and read back by the leaf hashing,
    chunk chaining values, Merkle layers) and VALU lane-operations (every
    BLAKE3 compression at its 680-op minimum, and the radix-8 stages at 7
    multiply-adds per output per stage); the step cannot beat the larger."""
    elem = D * 8
    cw = num_rows * N * elem
    leaf_bytes = num_rows * elem
    chunks = (leaf_bytes + 1023) // 1024
    hbm = N_EVALS * 8 + 2 * cw + 2 * N * chunks * 32 + (2 * N - 1) * 32
    blocks = cw // 64                      # 64-byte blocks of every leaf
    parents = N * (chunks - 1) + (N - 1)   # chunk-tree and Merkle parents
    depth = 2
    stage_fma = 7 * depth * num_rows * N * D
    valu = BLAKE3_OPS_PER_COMPRESSION * (blocks + parents) + stage_fma
    hbm_ms = hbm / (HBM_PEAK_GBS * 1e9) * 1e3
    valu_ms = valu / VALU_LANE_OPS_PEAK * 1e3
    floor = max(hbm_ms, valu_ms)
    return {"bound": "valu" if valu_ms >= hbm_ms else "hbm", "hbm_bytes": hbm,
            "hbm_floor_ms": hbm_ms, "valu_lane_ops": valu, "valu_floor_ms": valu_ms,
            "floor_ms": floor, "frac": floor / ms_per_step}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def traffic_per_launch(key="traffic_bytes_per_launch"):
    """HBM bytes per launch from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE, each scaled by the copy-kernel calibration): `key` names the
    encode_rows pair or iprs_mfma_kernel alone."""
    try:
        with open(TRAFFIC_FILE) as f:
            m = re.search(r"^" + re.escape(key) + r":\s*([0-9]+)", f.read(), re.M)
        return int(m.group(1)) if m else None
    except OSError:
        return None


def cpu_baseline(gpu_root, seconds=10.0):
    """The reference's CPU commit of the whole workload, restated
    (oracle/baseline.cpp): 2^20 BinaryPoly<32> under the same IPRS code, 64
    rows, on the reference's parallel axes (codeword indices, blocks, columns,
    pairs) with its AVX-512 widening and 16-lane BLAKE3 where the host has
    AVX-512; all allotted host threads (up to 16), then one thread.  Its root
    must equal the GPU's."""
    import numpy as np

    from oracle import oracle
    from zipgpu import synth

    threads = min(16, os.cpu_count() or 1)
    x = synth.binary_polys(2, N_EVALS, D).view(np.uint64)  # the GPU step's input
    rows = N_EVALS // ROW_LEN
    root = oracle.baseline_commit_iprs(x, rows, ROW_LEN, D=D, threads=threads)["root"]
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.baseline_commit_iprs(x, rows, ROW_LEN, D=D, threads=threads)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds / 2 or reps >= 10:
            break
    t1 = time.perf_counter()
    oracle.baseline_commit_iprs(x, rows, ROW_LEN, D=D, threads=1)
    one = time.perf_counter() - t1
    simd = "AVX-512" if oracle.baseline_has_avx512() else "portable (no AVX-512)"
    return {"value": N_EVALS * reps / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "one_thread_value": N_EVALS / one, "root_equals_gpu": root == gpu_root,
            "sample": f"the full workload: commit of 2^20 BinaryPoly<32>, IPRS rows 2^14 x 64, "
                      f"BLAKE3 Merkle over 65536 columns; {reps} commits in {el:.2f} s on "
                      f"{threads} threads ({simd}), one commit in {one:.2f} s on 1 thread"}


def wall_ms(torch, fn, steps, warmup=1):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def event_ms(torch, fn, stream, reps, warmup=2):
    """Average duration of fn's device work on `stream`, by HIP events."""
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            fn(stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn(stream)
        e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / reps


def leg_plain_zip(torch, zipgpu, synth, dev):
    """Plain Zip commit: 2^16 i32 evals, IPRS row 2^14 x 4 rows."""
    n = 1 << 16
    x = torch.from_numpy(synth.i32_evals(1, n)).to(dev)
    code = zipgpu.IprsCode.new_with_optimal_depth(1 << 14)
    zp = zipgpu.ZipPlus(zipgpu.IntTypes())
    pp = zp.setup(n, code)

    def step():
        h, _ = zp.commit(pp, [x])
        h.free()
    ms = wall_ms(torch, step, 20, 3)
    return {"workload": "Zip commit 2^16 i32, IPRS row 2^14 x 4, Cw i128", "commit_ms": ms,
            "evals_per_s": n / ms * 1e3}


def leg_raa(torch, zipgpu, synth, dev):
    """Zip+ RAA: 2^22 BinaryPoly<64>, row 2^11 x 2^11 rows, plus combined row."""
    n, row_len, d = 1 << 22, 1 << 11, 64
    x = torch.from_numpy(synth.binary_polys(3, n, d)).to(dev)
    code = zipgpu.RaaCode(row_len)
    types = zipgpu.BinPolyTypes(d)
    zp = zipgpu.ZipPlus(types)
    pp = zp.setup(n, code)
    out = torch.empty((pp.num_rows, code.codeword_len(), d), dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream()
    # Back-to-back launches from idle run fast for a few launches, slow down
    # for a few more and then settle (the board's power management), so the
    # average spans 40 launches after 20.
    enc = event_ms(torch, lambda s: code.encode_rows(x, pp.num_rows, types, out=out, stream=s),
                   stream, 40, warmup=20)
    del out

    def step():
        h, _ = zp.commit(pp, [x])
        h.free()
    ms = wall_ms(torch, step, 5, 1)
    al = torch.from_numpy(synth.i128_words(5, d)).to(dev)
    co = torch.from_numpy(synth.i128_words(6, pp.num_rows)).to(dev)
    comb = wall_ms(torch, lambda: zp.combined_row(pp, [x], al, co), 10, 2)
    enc_gbs = (8 + REP * d * 4) * n / (enc * 1e-3) / 1e9
    return {"workload": "Zip+ RAA commit 2^22 BinaryPoly<64>, row 2^11 x 2^11, Cw [i32;64]",
            "encode_ms": enc, "encode_gbs_algorithmic": enc_gbs,
            "encode_hbm_frac": enc_gbs / HBM_PEAK_GBS, "commit_ms": ms,
            "evals_per_s": n / ms * 1e3, "combined_row_ms": comb}


def leg_raa_signflip(torch, zipgpu, synth, dev):
    """RaaSignFlippingCode encode_rows: 2^22 BinaryPoly<64>, row 2^11 x 2^11 rows."""
    n, row_len, d = 1 << 22, 1 << 11, 64
    x = torch.from_numpy(synth.binary_polys(3, n, d)).to(dev)
    code = zipgpu.RaaSignFlippingCode(row_len)
    types = zipgpu.BinPolyTypes(d)
    pp = zipgpu.ZipPlus(types).setup(n, code)
    out = torch.empty((pp.num_rows, code.codeword_len(), d), dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream()
    enc = event_ms(torch, lambda s: code.encode_rows(x, pp.num_rows, types, out=out, stream=s),
                   stream, 40, warmup=20)
    gbs = (8 + REP * d * 4) * n / (enc * 1e-3) / 1e9
    return {"workload": "RaaSignFlippingCode encode_rows 2^22 BinaryPoly<64>, row 2^11 x 2^11",
            "encode_ms": enc, "encode_gbs_algorithmic": gbs, "encode_hbm_frac": gbs / HBM_PEAK_GBS}


def leg_encode_wide(torch, zipgpu, synth, dev):
    """LinearCode::encode_wide of CombR = Int<5> rows: the verifier's
    encoding of the combined row (one row), and 16 rows for throughput."""
    L, out = 5, {}
    for name, code in (("iprs_2p14", zipgpu.IprsCode.new_with_optimal_depth(1 << 14)),
                       ("raa_2p11", zipgpu.RaaCode(1 << 11)),
                       ("raa_flip_2p11", zipgpu.RaaSignFlippingCode(1 << 11))):
        rl, N = code.row_len(), code.codeword_len()
        res = {}
        for rows in (1, 16):
            x = torch.from_numpy(synth.comb_rows(8, rows * rl, L).reshape(rows, rl, L)).to(dev)
            stream = torch.cuda.Stream()
            ms = event_ms(torch, lambda s: code.encode_wide(x, stream=s), stream, 10)
            res[f"rows{rows}_ms"] = ms
            res[f"rows{rows}_gbs_algorithmic"] = rows * (rl + N) * L * 8 / (ms * 1e-3) / 1e9
        out[name] = res
    out["workload"] = "encode_wide Int<5>: IPRS row 2^14 (depth 2), RAA row 2^11, rate 1/4"
    return out


def leg_prove(torch, zipgpu, synth, dev, cpu=True):
    """ZipPlus::prove (alphas, b, combined row, 147 column openings, proof
    bytes) after a commit, at the reference bench's point (all ones)."""
    import numpy as np

    out = {}
    for name, kind, n, row_len, d in (("iprs_2p20_d32", "iprs", 1 << 20, 1 << 14, 32),
                                      ("raa_2p22_d64", "raa", 1 << 22, 1 << 11, 64)):
        types = zipgpu.BinPolyTypes(d)
        code = (zipgpu.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs"
                else zipgpu.RaaCode(row_len))
        zp = zipgpu.ZipPlus(types)
        pp = zp.setup(n, code)
        x = torch.from_numpy(synth.binary_polys(7, n, d)).to(dev)
        hint, comm = zp.commit(pp, [x])
        base = zipgpu.PcsProverTranscript.new_from_commitment(comm)
        p = base.get_random_field_cfg()
        point = [1] * (n.bit_length() - 1)
        last = {}

        def step():
            t = base.clone()
            last["ev"] = zp.prove(t, pp, [x], point, hint, p)
            last["t"] = t
        ms = wall_ms(torch, step, 10, 2)
        out[name] = {"prove_ms": ms, "proof_bytes": len(last["t"].stream())}
        hint.free()
    if cpu:
        from oracle import oracle
        threads = min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        n, row_len, rows = 1 << 20, 1 << 14, 64
        xs = synth.binary_polys(7, n, 32).view(np.uint64)
        c = oracle.commit("iprs", xs, rows, row_len, D=32, want_cw=True, want_layers=True)
        ts = oracle.Transcript()
        ts.absorb_slice(c["root"])
        p = ts.random_field_cfg()
        t0 = time.perf_counter()
        oracle.prove(ts, p, xs, rows, row_len, 4 * row_len, 256, c["cw"], c["layers"], [1] * 20,
                     D=32, L=5)
        out["iprs_2p20_d32"]["cpu_oracle_prove_ms"] = (time.perf_counter() - t0) * 1e3
        out["iprs_2p20_d32"]["cpu_cores"] = threads
    out["workload"] = ("prove after commit: 2^20 BinaryPoly<32> IPRS row 2^14 x 64, and 2^22 "
                       "BinaryPoly<64> RAA row 2^11 x 2^11; MontyField<4>, Int<5>, 147 openings")
    return out


def leg_step0(torch, zipgpu, synth, dev, cpu=True):
    """The protocol's step 0 (protocol/src/prover.rs:218-262) at its IPRS
    set-up (one row per poly, row_len = poly size 2^14): 4 binary
    (BinaryPoly<32>), 2 arbitrary (DensePolynomial<i64, 32>) and 4 int (i64)
    witness columns committed, one public column of each absorbed."""
    import numpy as np
    n, D = 1 << 14, 32
    groups = [("binpoly", zipgpu.BinPolyTypes(D, 8), 4), ("dense_i64", zipgpu.DensePolyTypes(D), 2),
              ("i64", zipgpu.I64Types(), 4)]
    code = zipgpu.IprsCode.new_with_optimal_depth(n)
    params, wit, pub, host = [], [], [], []
    for g, (et, types, cols) in enumerate(groups):
        pp = zipgpu.ZipPlus(types).setup(n, code)
        params.append((pp, types))
        xs = []
        for k in range(cols + 1):
            if et == "binpoly":
                xs.append(synth.binary_polys(300 + 10 * g + k, n, D))
            else:
                xs.append((synth.splitmix64(300 + 10 * g + k, n * types.eval_words)
                           .view(np.int64) >> np.int64(40)))
        host.append(xs)
        pub.append([torch.from_numpy(xs[0]).to(dev)])
        wit.append([torch.from_numpy(v).to(dev) for v in xs[1:]])

    def step():
        hints, _, _ = zipgpu.step0_commit(params, wit, pub)
        for h in hints:
            h.free()
    ms = wall_ms(torch, step, 10, 2)
    evals = n * sum(c for _, _, c in groups)
    out = {"workload": "step0_commit: 4 BinaryPoly<32> + 2 DensePolynomial<i64,32> + 4 i64 "
                       "witness columns of 2^14 evals, IPRS row 2^14, public columns absorbed",
           "step0_ms": ms, "witness_evals_per_s": evals / ms * 1e3}
    if cpu:
        from oracle import oracle
        oracle.set_threads(min(16, os.cpu_count() or 1))
        t0 = time.perf_counter()
        for g, (et, types, cols) in enumerate(groups):
            xs = np.concatenate(host[g][1:])
            if et == "binpoly":
                oracle.commit("iprs", xs.view(np.uint64), 1, n, batch=cols, D=D)
            else:
                oracle.commit_i64("iprs", xs, 1, n, cols, types.degree_plus_one, et)
        out["cpu_oracle_commits_ms"] = (time.perf_counter() - t0) * 1e3
    return out


def leg_batched(torch, zipgpu, synth, dev):
    """64 commits, log2 n uniform in 12..20, D 32/64, IPRS/RAA: one batch plan
    (HIP graph, autotuned variants) against one zp_commit per poly."""
    reqs = synth.request_mix(4, 64, 12, 20)
    total = sum(r["n"] for r in reqs)
    codes, items, xs = {}, [], []
    for r in reqs:
        key = (r["kind"], r["row_len"])
        if key not in codes:
            codes[key] = (zipgpu.IprsCode.new_with_optimal_depth(r["row_len"])
                          if r["kind"] == "iprs" else zipgpu.RaaCode(r["row_len"]))
        types = zipgpu.BinPolyTypes(r["D"])
        items.append((zipgpu.ZipPlus(types).setup(r["n"], codes[key]), types))
        xs.append(torch.from_numpy(synth.binary_polys(r["seed"], r["n"], r["D"])).to(dev))

    def plain():
        hs = [zipgpu.ZipPlus(t).commit(pp, [x])[0] for (pp, t), x in zip(items, xs)]
        for h in hs:
            h.free()
    plain_ms = wall_ms(torch, plain, 3, 1)
    plan = zipgpu.BatchCommit(items, graph=True, autotune=True, n_streams=8)
    roots = [c.root for c in plan.commit(xs)]
    plan_ms = wall_ms(torch, lambda: plan.commit(xs), 5, 1)
    grouped, streams = plan.schedule
    tuned_plain_ms = wall_ms(torch, plain, 3, 1)
    # Each schedule on its own plan (graph replays): grouped leaf hashing
    # after the encoders, or each item's hashing on its branch stream.
    schedules = {}
    for g in (True, False):
        for k in (1, 4, 8):
            p2 = zipgpu.BatchCommit(items, graph=True, autotune=False, n_streams=8)
            p2.set_schedule(g, k)
            p2.commit(xs)
            schedules[f"{'grouped' if g else 'per_item'}_{k}_streams_ms"] = wall_ms(
                torch, lambda: p2.commit(xs), 5, 1)
            del p2
    # One branch stream with per-item hashing: every item's device work back
    # to back with no host work between them (the serial device time).
    serial_ms = schedules["per_item_1_streams_ms"]
    # The plan's roots equal the plain path's.
    same = roots == [zipgpu.ZipPlus(t).commit(pp, [x])[1].root
                     for (pp, t), x in zip(items, xs)]
    info = plan.info()
    del plan
    return {"workload": "64 commits, log2 n ~ U{12..20}, D 32/64, IPRS (row min(n,2^14)) / "
                        "RAA (row isqrt(n) rounded up to 2^k), seeded mix 4",
            "total_evals": total, "plain_ms": plain_ms, "plain_evals_per_s": total / plain_ms * 1e3,
            "plain_autotuned_ms": tuned_plain_ms, "batched_graph_ms": plan_ms,
            "serial_graph_1_stream_ms": serial_ms, "batched_vs_serial": serial_ms / plan_ms,
            "batched_evals_per_s": total / plan_ms * 1e3, "batched_vs_plain": plain_ms / plan_ms,
            "roots_equal_plain": same, "hint_bytes": info["hint_bytes"],
            "scratch_bytes": info["scratch_bytes"],
            "tuned_schedule": {"grouped_hashing": grouped, "branch_streams": streams},
            "schedules": schedules}


def served_requests(zipgpu, synth, seed, count, lo, hi):
    import numpy as np
    codes, out = {}, []
    for r in synth.request_mix(seed, count, lo, hi):
        key = (r["kind"], r["row_len"])
        if key not in codes:
            codes[key] = (zipgpu.IprsCode.new_with_optimal_depth(r["row_len"])
                          if r["kind"] == "iprs" else zipgpu.RaaCode(r["row_len"]))
        types = zipgpu.BinPolyTypes(r["D"], 5)
        pp = zipgpu.ZipPlus(types).setup(r["n"], codes[key])
        N = codes[key].codeword_len()
        out.append({"pp": pp, "types": types, "n": r["n"],
                    "x": synth.binary_polys(r["seed"], r["n"], r["D"]),
                    "al": synth.i128_words(r["seed"] + 1, r["D"]),
                    "co": synth.i128_words(r["seed"] + 2, pp.num_rows),
                    "cols": (synth.splitmix64(r["seed"] + 3, 147) % np.uint64(N)),
                    "point": [int(v) for v in synth.splitmix64(r["seed"] + 4,
                                                               r["n"].bit_length() - 1)]})
    return out, codes


def serve(zipgpu, reqs, slots, arrivals, open_delays, prove=False):
    """Run the requests through a CommitServer: commit i arrives at
    arrivals[i] (seconds from the start), its opening (or, with prove, its
    full ZipPlus::prove at the point q["point"]) open_delays[i] after its
    root is back.  -> (wall seconds, commit latencies us, open latencies us,
    HBM peak, queue depths: jobs waiting for a slot or memory, sampled at
    each arrival before it is submitted)."""
    import heapq
    srv = zipgpu.CommitServer(slots=slots)
    t0 = time.perf_counter()
    due = [(a, 0, i) for i, a in enumerate(arrivals)]  # (time, kind 0 commit / 1 open, index)
    heapq.heapify(due)
    ids, lat_c, lat_o, depth, left = {}, [], [], [], 2 * len(reqs)
    while left:
        now = time.perf_counter() - t0
        while due and due[0][0] <= now:
            _, kind, i = heapq.heappop(due)
            q = reqs[i]
            if kind == 0:
                depth.append(srv.stats()["pending"])
                ids[srv.commit(q["pp"], q["types"], q["x"])] = i
            elif prove:
                srv.prove(q["rid"], q["point"])
            else:
                srv.open(q["rid"], q["al"], q["co"], q["cols"])
        wait = max(0.0, due[0][0] - (time.perf_counter() - t0)) if due else 1.0
        ev = srv.poll(min(wait, 1.0))
        if ev is None:
            continue
        if ev["status"] != 0:
            raise RuntimeError(f"served request failed: {ev}")
        i = ids[ev["request"]]
        if ev["kind"] == "commit":
            lat_c.append(ev["latency_us"])
            reqs[i]["rid"] = ev["request"]
            reqs[i]["root"] = ev["root"]
            heapq.heappush(due, (time.perf_counter() - t0 + open_delays[i], 1, i))
        else:
            lat_o.append(ev["latency_us"])
        left -= 1
    wall = time.perf_counter() - t0
    peak = srv.stats()["hbm_peak"]
    srv.close()
    return wall, lat_c, lat_o, peak, depth


def poisson_arrivals(synth, seed, rate, count=None, window=None):
    """Arrival times of a Poisson process of `rate` per second: the first
    `count`, or all within `window` seconds (times from 0)."""
    import numpy as np
    n = count if count is not None else int(rate * window * 1.5) + 16
    u = (synth.splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    gaps = -np.log1p(-u) / rate
    t = np.cumsum(gaps) - gaps[0]
    return t[:count] if count is not None else t[t < window]


def leg_served(torch, zipgpu, synth, dev):
    """Commit+open service: Poisson arrivals over a fixed window, log2 n
    uniform in 14..22, D 32/64, IPRS/RAA; openings (combined row + 147
    columns with Merkle paths) a seeded delay after each commit."""
    import numpy as np
    pool_n, rate, window = 48, 400.0, 1.5  # distinct requests, arrivals per second, seconds
    pool, _ = served_requests(zipgpu, synth, 5, pool_n, 14, 22)
    arrivals = poisson_arrivals(synth, 77, rate, window=window)
    count = len(arrivals)
    reqs = [dict(pool[i % pool_n]) for i in range(count)]
    total = sum(q["n"] for q in reqs)
    u = (synth.splitmix64(78, count) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    delays = -np.log1p(-u) * 1e-3  # openings: mean 1 ms after the root
    # Warm-up: every distinct request once, so the codes, the device memory
    # pool and the pinned staging reach their steady state (a long-running
    # server is warm); with only a few requests first, the pool grows during
    # the measured run and its allocations set the tail (p99 23-37 ms against
    # 8 ms warm).
    serve(zipgpu, [dict(q) for q in pool], 4, [0.0] * pool_n, [0.0] * pool_n)
    wall, lc, lo, peak, depth = serve(zipgpu, reqs, 4, arrivals, delays)
    # Roots of requests that share inputs agree.
    same = all(reqs[i]["root"] == reqs[i % pool_n]["root"] for i in range(count))
    # Saturated throughput with 4 slots and with 1 over the 48 distinct
    # requests: medians of 3 alternating runs.
    sat = [dict(q) for q in pool]
    s4, s1 = [], []
    for _ in range(3):
        s4.append(serve(zipgpu, sat, 4, [0.0] * pool_n, [0.0] * pool_n)[0])
        s1.append(serve(zipgpu, sat, 1, [0.0] * pool_n, [0.0] * pool_n)[0])
    sat_total = sum(q["n"] for q in pool)
    sat4, sat1 = float(np.median(s4)), float(np.median(s1))
    # Full proofs (Fiat-Shamir transcript from each root, b, combined row,
    # 147 openings, proof bytes) instead of openings: 32 requests, Poisson 300/s.
    pr = [dict(q) for q in pool[:32]]
    parr = poisson_arrivals(synth, 79, 300.0, count=32)
    pwall, _, lp, _, _ = serve(zipgpu, pr, 4, parr, delays[:32], prove=True)
    pct = lambda a, p: float(np.percentile(np.asarray(a), p))  # noqa: E731
    return {"workload": f"{count} commit+open requests (a pool of {pool_n} distinct inputs), "
                        f"Poisson {rate:.0f}/s over {window} s, log2 n ~ U{{14..22}}, D 32/64, "
                        "IPRS/RAA, host inputs; openings: combined row Int<5> + 147 columns with "
                        "Merkle paths, exp(1 ms) after each root; 4 slots, warmed by one pass "
                        "over the distinct inputs; queue depth = jobs "
                        "waiting for a slot or memory at each arrival; prove_*: 32 requests "
                        "at 300/s with full proofs (transcript from the root, 147 openings)",
            "requests": count, "total_evals": total, "wall_s": wall, "evals_per_s": total / wall,
            "commit_latency_us_p50": pct(lc, 50), "commit_latency_us_p99": pct(lc, 99),
            "open_latency_us_p50": pct(lo, 50), "open_latency_us_p99": pct(lo, 99),
            "queue_depth_mean": float(np.mean(depth)), "queue_depth_p99": pct(depth, 99),
            "queue_depth_max": int(max(depth)),
            "hbm_peak_bytes": peak,
            "prove_wall_s": pwall, "prove_latency_us_p50": pct(lp, 50),
            "prove_latency_us_p99": pct(lp, 99),
            "saturated_4_slots_evals_per_s": sat_total / sat4,
            "one_at_a_time_evals_per_s": sat_total / sat1,
            "served_vs_one_at_a_time": sat1 / sat4,
            "roots_equal_for_equal_inputs": same}


def copy_calibration():
    """Plain 1 GiB device copies, for scaling the FETCH_SIZE / WRITE_SIZE
    counters (run under rocprofv3 --pmc)."""
    import torch
    a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda:0").fill_(1)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("copy calibration: 3 copies of 1073741824 bytes", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # Commits of 0.7 ms: 40 warm-up steps (28 ms) bring the board's clocks
    # to their steady state before 100 timed steps (70 ms).
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-legs", action="store_true", help="skip the other workloads")
    ap.add_argument("--legs", default="", help="comma-separated workload names to run (default all)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run plain 1 GiB copies (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): this build runs on a "
              "single MI355X, and the row/column sharding over several GPUs is a design, not "
              "a measured path", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import numpy as np  # noqa: F401
    import torch

    import zipgpu
    from zipgpu import synth

    assert torch.cuda.is_available(), "bench needs a HIP device"
    assert zipgpu.lib().zp_init(0) == 0
    dev = torch.device("cuda:0")
    x = torch.from_numpy(synth.binary_polys(2, N_EVALS, D)).to(dev)
    code = zipgpu.IprsCode.new_with_optimal_depth(ROW_LEN)
    types = zipgpu.BinPolyTypes(D)
    zp = zipgpu.ZipPlus(types)
    pp = zp.setup(N_EVALS, code)
    log(f"workload: n=2^20 D={D} IPRS row_len={ROW_LEN} rows={pp.num_rows} "
        f"depth={code.depth} N={code.codeword_len()}")

    # Whole-commit throughput (one step = one commit of the resident input).
    for _ in range(args.warmup):
        h, c = zp.commit(pp, [x])
        h.free()
    torch.cuda.synchronize()
    # The encoder's launches inside the timed commits are bracketed by HIP
    # timing events on the commit stream: every 4th commit, from the first,
    # records two events (events made before the loop).
    every = 4
    with zipgpu.EncodeTiming(args.steps, every=every) as et:
        t0 = time.perf_counter()
        for _ in range(args.steps):
            h, c = zp.commit(pp, [x])
            h.free()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = N_EVALS * args.steps / el
    assert et.launches == (args.steps + every - 1) // every, (et.launches, args.steps)
    enc_ms, main_ms = et.mean_ms, et.mean_last_ms
    log(f"commit: {ms:.3f} ms/step, {value:.4g} evals/s, root {c.root.hex()[:16]}; "
        f"inside the commits iprs_mfma_kernel {main_ms:.3f} ms, encode_rows {enc_ms:.3f} ms")

    # A caller holding the evaluations in host memory, as the reference's
    # callers do: zp_commit_host (the upload of each row group overlaps the
    # encoding of the groups before it), and beside it a plain copy followed
    # by the device commit.
    pcie = {}
    for kind, xh in (("pinned", x.cpu().pin_memory()), ("pageable", x.cpu())):
        hh, hc = zp.commit(pp, [xh])
        hh.free()
        assert hc.root == c.root, "host-input commit root differs"

        def step_h():
            h, _ = zp.commit(pp, [xh])
            h.free()

        def step_c():
            x.copy_(xh, non_blocking=True)
            h, _ = zp.commit(pp, [x])
            h.free()
        ms_h = wall_ms(torch, step_h, 10, 2)
        ms_c = wall_ms(torch, step_c, 10, 2)
        pcie[kind] = {"ms_per_commit": ms_h, "evals_per_s": N_EVALS / ms_h * 1e3,
                      "copy_then_commit_ms": ms_c}
    log(f"host-input commit: {pcie}")

    # Dominant kernel: iprs_mfma_kernel, timed inside the commits (above)
    # from an event recorded after the bit-plane pre-pass to one after it.
    # The pair of launches (one encode_rows) is also timed back to back on a
    # stream of its own, where the sustained load lowers the clocks.
    stream = torch.cuda.Stream()
    out = torch.empty((pp.num_rows, code.codeword_len(), D), dtype=torch.int64, device=dev)

    def enc(s):
        code.encode_rows(x, pp.num_rows, types, out=out, stream=s)
    iso_ms = event_ms(torch, enc, stream, max(40, args.steps), warmup=20)
    zipgpu.set_option("iprs_encoder", 1)
    try:
        valu_ms = event_ms(torch, enc, stream, 5)
    finally:
        zipgpu.set_option("iprs_encoder", -1)
    del out
    enc_bytes = ENCODE_BYTES_PER_EVAL * N_EVALS
    main_bytes = MAIN_BYTES_PER_EVAL * N_EVALS
    achieved = main_bytes / (main_ms * 1e-3) / 1e9
    log(f"iprs_mfma_kernel: {main_ms:.3f} ms/launch inside the commits, {achieved:.1f} GB/s "
        f"algorithmic; encode_rows back to back {iso_ms:.3f} ms (VALU encoder {valu_ms:.3f} ms)")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": "Zip+ commit: 2^20 BinaryPoly<32> evals (Bernoulli(1/2) bits, "
                               "u64 lanes), IPRS F65537 rate 1/4, row_len 2^14 x 64 rows, "
                               "depth 2, Cw [i64;32], BLAKE3 column Merkle (65536 leaves)",
                   "n_evals": N_EVALS, "row_len": ROW_LEN, "num_rows": pp.num_rows,
                   "degree_plus_one": D, "code": "IPRS", "rep": REP},
        "roofline": {"bound": "hbm", "kernel": MAIN_KERNEL,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_per_launch("traffic_bytes_per_launch_iprs_mfma_kernel"),
                     "launch_ms": main_ms, "launches": et.launches,
                     "timed_in": "every 4th of the timed commits (HIP events on the commit "
                                 "stream, after the pre-pass to after the kernel)",
                     "bytes_per_launch": main_bytes,
                     "valu_encoder_launch_ms": valu_ms},
        "encode_rows": {"kernels": ENCODE_KERNEL, "bytes_per_launch": enc_bytes,
                        "in_commits_ms": enc_ms,
                        "in_commits_frac": enc_bytes / (enc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "in_commits_note": "from an event before the pre-pass: also holds the "
                                           "host's submission of the pre-pass to the idle stream",
                        "back_to_back_ms": iso_ms,
                        "back_to_back_frac": enc_bytes / (iso_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "traffic": traffic_per_launch()},
        "commit_roofline": commit_roofline(ms, pp.num_rows, code.codeword_len()),
        "host_input_commit": pcie,
    }
    if not args.no_legs:
        legs = {}
        for name, fn in (("plain_zip_2p16_i32", leg_plain_zip), ("raa_2p22_d64", leg_raa),
                         ("raa_signflip_2p22_d64", leg_raa_signflip),
                         ("batched_64_mixed", leg_batched), ("served_commit_open", leg_served),
                         ("encode_wide_int5", leg_encode_wide),
                         ("prove", lambda *a: leg_prove(*a, cpu=not args.no_cpu)),
                         ("step0_commit", lambda *a: leg_step0(*a, cpu=not args.no_cpu))):
            if args.legs and name not in args.legs.split(","):
                continue
            log(f"leg {name} ...")
            legs[name] = fn(torch, zipgpu, synth, dev)
            log(f"  {legs[name]}")
            torch.cuda.empty_cache()
        line["workloads"] = legs
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(c.root)
        log(f"cpu: {line['cpu_baseline']['value']:.4g} evals/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
