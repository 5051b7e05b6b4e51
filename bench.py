"""Zip+ commit throughput on one MI355X.

Workload (the headline configuration): Zip+ commit of one polynomial with
2^20 evaluations of BinaryPoly<32> (i.i.d. Bernoulli(1/2) coefficient bits in
u64 lanes), IPRS code F65537 rate 1/4, row_len 2^14 x 64 rows (depth 2),
Cw = [i64; 32] (1 GiB of codewords), BLAKE3 column Merkle tree over 65536
leaves of 16 KiB.  One step = one ZipPlus::commit (encode + Merkle) of the
resident input; value = evaluations committed per second.

Also reported: the dominant kernel's roofline (IPRS encode, timed with HIP
events on its own stream) and the CPU oracle on a bounded sample of the
same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "zinc-plus_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

N_EVALS = 1 << 20
ROW_LEN = 1 << 14
D = 32
REP = 4
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes of IPRS encode per evaluation: the u64 input lane plus
# REP * D * 8 bytes of [i64; D] codeword.
ENCODE_BYTES_PER_EVAL = 8 + REP * D * 8


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(seconds=12.0):
    """The oracle (CPU restatement of the reference) committing a bounded
    sample of the workload: 8 rows of the same code (2^17 evals)."""
    import numpy as np

    from oracle import oracle
    from zipgpu import synth

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    rows = 8
    n = rows * ROW_LEN
    x = synth.binary_polys(2, n, D).view(np.uint64)
    oracle.commit("iprs", x, rows, ROW_LEN, D=D)  # warm
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.commit("iprs", x, rows, ROW_LEN, D=D)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds or reps >= 50:
            break
    return {"value": n * reps / el, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"oracle commit of {rows} rows x 2^14 evals (2^17 BinaryPoly<32>, same "
                      f"IPRS code, BLAKE3 Merkle over 8-row columns), {reps} reps in {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU layout "
              "is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import numpy as np
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

    # Whole-commit throughput.
    for _ in range(args.warmup):
        h, c = zp.commit(pp, [x])
        h.free()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h, c = zp.commit(pp, [x])
        h.free()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    value = N_EVALS * args.steps / el
    log(f"commit: {ms:.3f} ms/step, {value:.4g} evals/s, root {c.root.hex()[:16]}")

    # Dominant kernel: IPRS encode (one launch over 64 rows) timed with HIP
    # events on the stream it is launched on.
    stream = torch.cuda.Stream()
    out = torch.empty((pp.num_rows, code.codeword_len(), D), dtype=torch.int64, device=dev)
    with torch.cuda.stream(stream):
        for _ in range(3):
            code.encode_rows(x, pp.num_rows, types, out=out, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k = max(5, args.steps)
        e0.record(stream)
        for _ in range(k):
            code.encode_rows(x, pp.num_rows, types, out=out, stream=stream)
        e1.record(stream)
    stream.synchronize()
    enc_ms = e0.elapsed_time(e1) / k
    enc_bytes = ENCODE_BYTES_PER_EVAL * N_EVALS
    achieved = enc_bytes / (enc_ms * 1e-3) / 1e9
    log(f"encode kernel: {enc_ms:.3f} ms/launch, {achieved:.1f} GB/s algorithmic")

    line = {
        "metric": "Zip+ commit evals/s at 2^20 deg-32 IPRS",
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
        "roofline": {"bound": "hbm", "kernel": "iprs_binary_kernel",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launch_ms": enc_ms, "bytes_per_launch": enc_bytes},
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline()
        log(f"cpu: {line['cpu_baseline']['value']:.4g} evals/s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
