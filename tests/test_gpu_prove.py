"""ZipPlus::prove on the device (zp_prove through the C ABI) against the
oracle's restatement: the same evaluation and the same proof bytes, and the
oracle's verifier accepts the device proof (phase_prove.rs:85-304,
phase_verify.rs:100-305)."""
import numpy as np
import pytest
import torch

from test_opening_oracle import expected_proof_bytes
from zipgpu import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def make_code(zg, kind, row_len):
    return {"iprs": lambda: zg.IprsCode.new_with_optimal_depth(row_len),
            "raa": lambda: zg.RaaCode(row_len),
            "raa_signflip": lambda: zg.RaaSignFlippingCode(row_len)}[kind]()


def encoder(oracle, kind, row_len, L):
    if kind == "iprs":
        return lambda comb: oracle.iprs_encode_wide(comb.reshape(1, row_len, L), row_len, -1, L)[0]
    return lambda comb: oracle.raa_encode_wide(comb.reshape(1, row_len, L), row_len, L,
                                               sign_flip=kind == "raa_signflip")[0]


def run_case(zg, dev, oracle, kind, row_len, num_rows, D, batch, L, eval_type, point_kind,
             num_openings=147, seed=1):
    n = num_rows * row_len
    if eval_type == "binpoly":
        x = synth.binary_polys(seed, batch * n, D)
        types = zg.BinPolyTypes(D, comb_limbs=L)
        elem, cb = (8 * D, 8) if kind == "iprs" else (4 * D, 4)
        ox = x.view(np.uint64)
    else:
        x = synth.i32_evals(seed, batch * n)
        types = zg.IntTypes(comb_limbs=L)
        elem, cb = 16, 16
        ox = x
    types.num_column_openings = num_openings
    code = make_code(zg, kind, row_len)
    zp = zg.ZipPlus(types)
    pp = zp.setup(n, code)
    polys = [to_dev(x[b * n:(b + 1) * n], dev) for b in range(batch)]
    hint, comm = zp.commit(pp, polys)
    nv = n.bit_length() - 1
    if point_kind == "ones":
        point = [1] * nv
    else:
        w = synth.i128_words(seed + 5, nv)
        point = [int(np.uint64(w[i, 0].view(np.uint64))) | (int(w[i, 1]) << 64) for i in range(nv)]
    t = zg.PcsProverTranscript.new_from_commitment(comm)
    p = t.get_random_field_cfg()
    ev = zp.prove(t, pp, polys, point, hint, p)
    got = t.stream()
    c = oracle.commit(kind, ox, num_rows, row_len, batch=batch, D=D, eval_type=eval_type,
                      want_cw=True, want_layers=True)
    assert c["root"] == comm.root
    ts = oracle.Transcript()
    ts.absorb_slice(c["root"])
    assert ts.random_field_cfg() == p
    want_ev = oracle.prove(ts, p, ox, num_rows, row_len, 4 * row_len, elem, c["cw"], c["layers"],
                           point, batch=batch, D=D, L=L, eval_type=eval_type,
                           num_openings=num_openings)
    assert ev == want_ev
    assert len(got) == len(ts.stream)
    assert got == bytes(ts.stream)
    assert len(got) == expected_proof_bytes(num_rows, row_len, 4 * row_len, elem, L, batch,
                                            num_openings)
    tv = oracle.Transcript(got)
    tv.absorb_slice(comm.root)
    tv.random_field_cfg()
    assert oracle.verify(tv, p, comm.root, point, ev, encoder(oracle, kind, row_len, L), num_rows,
                         row_len, 4 * row_len, elem, batch=batch, D=D, L=L, eval_type=eval_type,
                         coeff_bytes=cb, num_openings=num_openings)
    return zp, pp, polys, hint, comm, point, p


@pytest.mark.parametrize("kind,row_len,num_rows,D,batch,L,eval_type,point_kind", [
    ("iprs", 1024, 16, 32, 1, 5, "binpoly", "random"),
    ("iprs", 1 << 14, 4, 64, 1, 5, "binpoly", "ones"),
    ("raa", 64, 64, 64, 2, 5, "binpoly", "random"),
    ("raa_signflip", 256, 16, 8, 1, 6, "binpoly", "random"),
    ("iprs", 512, 1, 1, 1, 5, "binpoly", "random"),
    ("iprs", 256, 8, 7, 3, 8, "binpoly", "random"),
    ("iprs", 1024, 4, 1, 1, 3, "i32", "random"),
    ("iprs", 1 << 14, 4, 1, 1, 3, "i32", "ones"),
    # 147 openings of 128 KiB: the proof copies back in two pieces.
    ("raa", 1 << 11, 512, 64, 1, 5, "binpoly", "ones"),
])
def test_prove_matches_oracle(zg, dev, oracle, kind, row_len, num_rows, D, batch, L, eval_type,
                              point_kind):
    say("prove", kind, row_len, num_rows, D, batch, L, eval_type, point_kind)
    run_case(zg, dev, oracle, kind, row_len, num_rows, D, batch, L, eval_type, point_kind)


def test_prove_at_commit_size(zg, dev, oracle):
    # The benchmarked commitment: 2^20 BinaryPoly<32>, IPRS row 2^14 x 64.
    say("prove 2^20")
    run_case(zg, dev, oracle, "iprs", 1 << 14, 64, 32, 1, 5, "binpoly", "random", seed=3)


def test_prove_twice_from_clones_is_deterministic(zg, dev, oracle):
    say("prove twice")
    zp, pp, polys, hint, comm, point, p = run_case(zg, dev, oracle, "raa", 128, 32, 16, 1, 5,
                                                   "binpoly", "random", num_openings=9)
    t0 = zg.PcsProverTranscript.new_from_commitment(comm)
    t0.get_random_field_cfg()
    t1 = t0.clone()
    e0 = zp.prove(t0, pp, polys, point, hint, p)
    e1 = zp.prove(t1, pp, polys, point, hint, p)
    assert e0 == e1 and t0.stream() == t1.stream()


def test_prove_errors(zg, dev):
    say("prove errors")
    types = zg.BinPolyTypes(16)
    code = zg.IprsCode.new_with_optimal_depth(256)
    zp = zg.ZipPlus(types)
    pp = zp.setup(1 << 12, code)
    x = to_dev(synth.binary_polys(2, 1 << 12, 16), dev)
    hint, comm = zp.commit(pp, [x])
    t = zg.PcsProverTranscript.new_from_commitment(comm)
    p = t.get_random_field_cfg()
    with pytest.raises(zg.InvalidPcsParam):
        zp.prove(t, pp, [x], [1] * 11, hint, p)  # 12 variables
    with pytest.raises(zg.InvalidPcsParam):
        zp.prove(t, pp, [x], [1] * 12, hint, p + 1)  # even modulus
    other = zg.ZipPlus(zg.BinPolyTypes(8))
    with pytest.raises(zg.ShapeError):
        other.prove(t, pp, [x & 0xFF], [1] * 12, hint, p)  # hint of another eval shape
