"""Writes tests/golden/golden.json: frozen outputs of the commitment path.

The reference is Rust and cannot run here, so the vectors come from the CPU
oracle after it has passed the pins of tests/test_oracle.py (the reference's
own known answers, published BLAKE3 vectors, closed forms, brute force).
Freezing them guards both the oracle and the device path against drift: the
CPU tests check the oracle reproduces every entry, the GPU tests check the
device does.  Inputs are seeded (zipgpu.synth), so an entry stores its seed
and shape, the SHA-256 of the output bytes, and the leading output words.

Each entry names the pin that validated the oracle for that function.

    python tests/golden/make_golden.py            # rewrite golden.json
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "zinc-plus_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import oracle  # noqa: E402
from zipgpu import synth  # noqa: E402

PIN_IPRS = ("test_oracle: mod-p DFT closed form (radix8.rs:224-284), base layer / twiddle "
            "closed forms (params.rs:114-195), brute force, linearity, plane consistency")
PIN_RAA = ("test_oracle: raa.rs:396-413 known answer, repeat/accumulate brute force, "
           "linearity, in-place == precomputed perms")
PIN_B3 = "test_oracle: published BLAKE3 vectors (i mod 251 inputs), hazmat merge consistency"
PIN_MERKLE = "test_oracle: leaves == column hashes, prove/verify round trip (merkle.rs:443-469)"
PIN_COMB = "test_oracle: Python big-int restatement of phase_prove.rs:153-244"

IPRS_BIN = [  # (row_len, depth, D, rows)
    (8, 1, 32, 2), (64, 1, 32, 2), (256, 1, 32, 2), (512, 1, 64, 1), (1024, 1, 32, 1),
    (2048, 1, 64, 1), (4096, 2, 32, 1), (4096, 2, 64, 1), (2048, 3, 32, 1)]
IPRS_I32 = [(8, 1, 2), (256, 1, 2), (2048, 1, 1), (4096, 2, 1)]
RAA_BIN = [  # (row_len, D, rows, sign_flip)
    (4, 1, 1, False), (64, 32, 2, False), (256, 64, 2, False), (1024, 32, 1, False),
    (2048, 64, 1, False), (64, 32, 2, True), (512, 64, 1, True)]
MERKLE = [(1, 1, 8), (3, 4, 64), (4, 16, 256), (2, 32, 100)]  # (rows, width, elem bytes)
COMB = [(16, 256, 64, 1, 5), (64, 1024, 32, 2, 5), (64, 4096, 64, 1, 5), (1, 512, 64, 1, 5)]


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def head(a, n=8):
    return [int(v) for v in np.ascontiguousarray(a).reshape(-1)[:n]]


def entries():
    out = {"octet_reversal": [], "twiddles": [], "iprs_binary": [], "iprs_i32": [],
           "raa_perm": [], "raa_binary": [], "blake3": [], "merkle_root": [],
           "combined_row": []}
    for d in (1, 2, 3):
        out["octet_reversal"].append({"octets": d, "table": [oracle.octet_reversal(i, d)
                                                             for i in range(8 ** d)],
                                      "pin": "octet_reversal.rs:62-72 known results"})
    for row_len, depth in ((256, 1), (4096, 2), (16384, 2), (2048, 3)):
        bm = oracle.base_matrix(row_len, depth)
        tw = [oracle.twiddles(row_len, depth, k) for k in range(depth)]
        out["twiddles"].append({"row_len": row_len, "depth": depth,
                                "base_matrix_sha256": digest(bm), "base_matrix_head": head(bm),
                                "twiddles_sha256": [digest(t) for t in tw],
                                "twiddles_head": [head(t) for t in tw], "pin": PIN_IPRS})
    for row_len, depth, D, rows in IPRS_BIN:
        seed = 5000 + row_len + D
        x = synth.binary_polys(seed, rows * row_len, D)
        cw = oracle.iprs_encode_binary(x.view(np.uint64), rows, row_len, depth, D)
        out["iprs_binary"].append({"seed": seed, "row_len": row_len, "depth": depth, "D": D,
                                   "rows": rows, "sha256": digest(cw), "head": head(cw),
                                   "pin": PIN_IPRS})
    for row_len, depth, rows in IPRS_I32:
        seed = 6000 + row_len
        x = synth.i32_evals(seed, rows * row_len)
        cw = oracle.iprs_encode_i32(x, rows, row_len, depth)
        out["iprs_i32"].append({"seed": seed, "row_len": row_len, "depth": depth, "rows": rows,
                                "sha256": digest(cw), "head": head(cw), "pin": PIN_IPRS})
    for n in (16, 256, 4096, 8192):
        for s in (1, 2):
            p = oracle.raa_perm(n, s)
            e = {"n": n, "seed": s, "sha256": digest(p.astype(np.uint64)), "head": head(p, 16),
                 "pin": PIN_RAA}
            if n == 16:
                e["perm"] = [int(v) for v in p]
            out["raa_perm"].append(e)
    for row_len, D, rows, flip in RAA_BIN:
        seed = 7000 + row_len + D + flip
        x = synth.binary_polys(seed, rows * row_len, D)
        cw = oracle.raa_encode_binary(x.view(np.uint64), rows, row_len, D, sign_flip=flip)
        out["raa_binary"].append({"seed": seed, "row_len": row_len, "D": D, "rows": rows,
                                  "sign_flip": flip, "sha256": digest(cw), "head": head(cw),
                                  "pin": PIN_RAA})
    for n in (0, 1, 63, 64, 65, 1023, 1024, 1025, 4096, 5000):
        data = bytes((31 * i + 7) % 256 for i in range(n))
        out["blake3"].append({"len": n, "input": "byte i = (31*i + 7) mod 256",
                              "hash": oracle.blake3(data).hex(), "pin": PIN_B3})
    for rows, width, elem in MERKLE:
        seed = 8000 + rows * width + elem
        raw = synth.splitmix64(seed, (rows * width * elem + 7) // 8).view(np.uint8)
        raw = raw[:rows * width * elem]
        out["merkle_root"].append({"seed": seed, "rows": rows, "width": width, "elem": elem,
                                   "input": "first rows*width*elem bytes of splitmix64(seed)",
                                   "root": oracle.merkle_root(raw.tobytes(), rows, width,
                                                              elem).hex(),
                                   "pin": PIN_MERKLE})
    for rows, row_len, D, batch, L in COMB:
        seed = 9000 + rows + row_len + D
        x = synth.binary_polys(seed, batch * rows * row_len, D)
        al = synth.i128_words(seed + 1, batch * D)
        co = synth.i128_words(seed + 2, rows)
        r = oracle.combined_row(x.view(np.uint64), al, co, rows, row_len, batch, D, L)
        out["combined_row"].append({"seed": seed, "rows": rows, "row_len": row_len, "D": D,
                                    "batch": batch, "L": L,
                                    "inputs": "evals binary_polys(seed), alphas "
                                              "i128_words(seed+1, batch*D), coeffs "
                                              "i128_words(seed+2, rows)",
                                    "sha256": digest(r), "head": head(r), "pin": PIN_COMB})
    return out


def main():
    oracle.set_threads(min(8, os.cpu_count() or 1))
    data = entries()
    body = json.dumps(data, indent=1, sort_keys=True)
    doc = {"manifest_sha256": hashlib.sha256(body.encode()).hexdigest(), "vectors": data}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", os.path.join(HERE, "golden.json"), doc["manifest_sha256"])


if __name__ == "__main__":
    main()
