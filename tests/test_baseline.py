"""The timed CPU baseline (oracle/baseline.cpp) equals the oracle's
ZipPlus::commit bit for bit: codewords and root, with and without the
AVX-512 paths, on one and several threads (CPU only)."""
import numpy as np
import pytest

from zipgpu import synth


@pytest.mark.parametrize("row_len,rows,D", [(1 << 10, 4, 32), (1 << 12, 4, 64), (1 << 11, 8, 32),
                                            (1 << 9, 16, 24)])
@pytest.mark.parametrize("simd,threads", [(True, 4), (False, 3)])
def test_baseline_commit_equals_oracle(oracle, row_len, rows, D, simd, threads):
    if simd and not oracle.baseline_has_avx512():
        pytest.skip("no AVX-512 on this host")
    x = synth.binary_polys(row_len + D, rows * row_len, D).view(np.uint64)
    ref = oracle.commit("iprs", x, rows, row_len, D=D, want_cw=True)
    got = oracle.baseline_commit_iprs(x, rows, row_len, D=D, threads=threads, simd=simd,
                                      want_cw=True)
    assert got["cw"].view(np.uint8).tobytes() == ref["cw"].tobytes()
    assert got["root"] == ref["root"]


def test_baseline_single_row_single_chunk(oracle):
    # One row of BinaryPoly<8>: every leaf is a single 64-byte chunk (ROOT in
    # the chunk), the other BLAKE3 case.
    x = synth.binary_polys(5, 1 << 10, 8).view(np.uint64)
    ref = oracle.commit("iprs", x, 1, 1 << 10, D=8)
    assert oracle.baseline_commit_iprs(x, 1, 1 << 10, D=8, threads=2)["root"] == ref["root"]
