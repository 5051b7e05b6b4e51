"""Frozen vectors (tests/golden/golden.json, written by make_golden.py).

CPU: the oracle reproduces every entry and the manifest digest holds.
GPU: the device path reproduces every entry through the C ABI.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from zipgpu import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def head(a, n=8):
    return [int(v) for v in np.ascontiguousarray(a).reshape(-1)[:n]]


def merkle_input(e):
    raw = synth.splitmix64(e["seed"], (e["rows"] * e["width"] * e["elem"] + 7) // 8)
    return raw.view(np.uint8)[:e["rows"] * e["width"] * e["elem"]]


def comb_inputs(e):
    x = synth.binary_polys(e["seed"], e["batch"] * e["rows"] * e["row_len"], e["D"])
    return (x, synth.i128_words(e["seed"] + 1, e["batch"] * e["D"]),
            synth.i128_words(e["seed"] + 2, e["rows"]))


# ---------------------------------------------------------------------------
# CPU: the oracle
# ---------------------------------------------------------------------------

def test_manifest(golden):
    body = json.dumps(golden["vectors"], indent=1, sort_keys=True)
    assert hashlib.sha256(body.encode()).hexdigest() == golden["manifest_sha256"]


def test_oracle_octets_and_tables(oracle, golden):
    v = golden["vectors"]
    for e in v["octet_reversal"]:
        assert [oracle.octet_reversal(i, e["octets"]) for i in range(8 ** e["octets"])] == \
            e["table"]
    for e in v["twiddles"]:
        bm = oracle.base_matrix(e["row_len"], e["depth"])
        assert digest(bm) == e["base_matrix_sha256"] and head(bm) == e["base_matrix_head"]
        for k in range(e["depth"]):
            t = oracle.twiddles(e["row_len"], e["depth"], k)
            assert digest(t) == e["twiddles_sha256"][k]


def test_oracle_encoders(oracle, golden):
    v = golden["vectors"]
    for e in v["iprs_binary"]:
        x = synth.binary_polys(e["seed"], e["rows"] * e["row_len"], e["D"])
        cw = oracle.iprs_encode_binary(x.view(np.uint64), e["rows"], e["row_len"], e["depth"],
                                       e["D"])
        assert digest(cw) == e["sha256"], e
    for e in v["iprs_i32"]:
        x = synth.i32_evals(e["seed"], e["rows"] * e["row_len"])
        assert digest(oracle.iprs_encode_i32(x, e["rows"], e["row_len"], e["depth"])) == \
            e["sha256"], e
    for e in v["raa_perm"]:
        p = oracle.raa_perm(e["n"], e["seed"])
        assert digest(p.astype(np.uint64)) == e["sha256"]
        if "perm" in e:
            assert [int(x) for x in p] == e["perm"]
    for e in v["raa_binary"]:
        x = synth.binary_polys(e["seed"], e["rows"] * e["row_len"], e["D"])
        cw = oracle.raa_encode_binary(x.view(np.uint64), e["rows"], e["row_len"], e["D"],
                                      sign_flip=e["sign_flip"])
        assert digest(cw) == e["sha256"], e


def test_oracle_hashes(oracle, golden):
    v = golden["vectors"]
    for e in v["blake3"]:
        data = bytes((31 * i + 7) % 256 for i in range(e["len"]))
        assert oracle.blake3(data).hex() == e["hash"]
    for e in v["merkle_root"]:
        raw = merkle_input(e)
        assert oracle.merkle_root(raw.tobytes(), e["rows"], e["width"], e["elem"]).hex() == \
            e["root"]
    for e in v["combined_row"]:
        x, al, co = comb_inputs(e)
        r = oracle.combined_row(x.view(np.uint64), al, co, e["rows"], e["row_len"], e["batch"],
                                e["D"], e["L"])
        assert digest(r) == e["sha256"]


# ---------------------------------------------------------------------------
# GPU: the device path
# ---------------------------------------------------------------------------

def to_dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.gpu
def test_device_encoders(zg, dev, golden):
    v = golden["vectors"]
    for e in v["iprs_binary"]:
        print("[case] iprs_binary", e["row_len"], e["depth"], e["D"], flush=True)
        x = synth.binary_polys(e["seed"], e["rows"] * e["row_len"], e["D"])
        code = zg.IprsCode(e["row_len"], e["depth"])
        cw = code.encode_rows(to_dev(x, dev), e["rows"], zg.BinPolyTypes(e["D"])).cpu().numpy()
        assert digest(cw) == e["sha256"], e
    for e in v["iprs_i32"]:
        print("[case] iprs_i32", e["row_len"], flush=True)
        x = synth.i32_evals(e["seed"], e["rows"] * e["row_len"])
        code = zg.IprsCode(e["row_len"], e["depth"])
        cw = code.encode_rows(to_dev(x, dev), e["rows"], zg.IntTypes()).cpu().numpy()
        assert digest(cw) == e["sha256"], e
    for e in v["raa_binary"]:
        print("[case] raa_binary", e["row_len"], e["D"], e["sign_flip"], flush=True)
        x = synth.binary_polys(e["seed"], e["rows"] * e["row_len"], e["D"])
        code = (zg.RaaSignFlippingCode if e["sign_flip"] else zg.RaaCode)(e["row_len"])
        cw = code.encode_rows(to_dev(x, dev), e["rows"], zg.BinPolyTypes(e["D"])).cpu().numpy()
        assert digest(cw) == e["sha256"], e


@pytest.mark.gpu
def test_device_hashes(zg, dev, golden):
    v = golden["vectors"]
    for e in v["blake3"]:
        if e["len"] == 0 or e["len"] % 4:
            continue  # device elements are whole 32-bit words (every codeword type is)
        raw = bytes((31 * i + 7) % 256 for i in range(e["len"]))
        data = np.frombuffer(raw, dtype=np.uint8).copy()
        # One row of one element: the root of a single-leaf tree is the leaf.
        root = zg.merkle_root(to_dev(data, dev).reshape(1, 1, -1), e["len"])
        assert root.hex() == e["hash"], e["len"]
    for e in v["merkle_root"]:
        raw = merkle_input(e)
        t = to_dev(raw, dev).reshape(e["rows"], e["width"], e["elem"])
        assert zg.merkle_root(t, e["elem"]).hex() == e["root"], e
    for e in v["combined_row"]:
        print("[case] combined_row", e["rows"], e["row_len"], e["D"], flush=True)
        x, al, co = comb_inputs(e)
        zp = zg.ZipPlus(zg.BinPolyTypes(e["D"], e["L"]))
        pp = zp.setup(e["rows"] * e["row_len"], zg.RaaCode(e["row_len"]))
        xd = to_dev(x, dev).reshape(e["batch"], -1)
        r = zp.combined_row(pp, [xd[b] for b in range(e["batch"])], to_dev(al, dev),
                            to_dev(co, dev)).cpu().numpy()
        assert digest(r.astype(np.uint64)) == e["sha256"], e
