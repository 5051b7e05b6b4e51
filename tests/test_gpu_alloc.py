"""The device cache and hints read on other streams.

A hint read by work on a second stream (through zp_hint_codewords) and then
freed must not be overwritten by the next commit on the commit stream
before that read has run: zp_hint_record_stream makes the next user of the
memory wait for it."""
import ctypes

import numpy as np
import pytest
import torch

from zipgpu import synth

pytestmark = pytest.mark.gpu

HIP_D2D = 3  # hipMemcpyDeviceToDevice


def hip():
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMemcpyAsync.restype = ctypes.c_int
    lib.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_int, ctypes.c_void_p]
    return lib


def test_hint_read_on_second_stream_survives_free_and_recommit(zg, dev):
    n, row_len, D = 1 << 18, 1 << 14, 32
    code = zg.IprsCode.new_with_optimal_depth(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    x1 = torch.from_numpy(synth.binary_polys(61, n, D)).to(dev)
    x2 = torch.from_numpy(synth.binary_polys(62, n, D)).to(dev)
    s1 = torch.cuda.Stream()  # a fresh commit stream: its cache starts empty
    s2 = torch.cuda.Stream()
    torch.cuda.synchronize()
    hint, _ = zp.commit(pp, [x1], stream=s1)
    want = hint.codewords(stream=s1)
    p1, nbytes = hint.device_codewords_ptr(), hint.cw_bytes_per_poly
    dst = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    with torch.cuda.stream(s2):
        torch.cuda._sleep(20_000_000)  # hold s2 so that its read runs late
    assert hip().hipMemcpyAsync(dst.data_ptr(), p1, nbytes, HIP_D2D, s2.cuda_stream) == 0
    hint.record_stream(s2)
    hint.free()
    # Same shape on the same stream: the cache hands the freed block out
    # again, and this commit must wait for s2's read before it writes.
    hint2, _ = zp.commit(pp, [x2], stream=s1)
    assert hint2.device_codewords_ptr() == p1
    s2.synchronize()
    s1.synchronize()
    assert dst.cpu().numpy().tobytes() == want.tobytes()
    # The second commit's codewords are its own (it did write the block).
    assert hint2.codewords(stream=s1).tobytes() != want.tobytes()
    hint2.free()


def test_commits_on_two_streams_share_the_cache(zg, dev, oracle):
    # Blocks freed on one stream stay with that stream; commits that
    # alternate between two streams keep their results.
    n, row_len, D = 1 << 14, 1 << 11, 32
    code = zg.RaaCode(row_len)
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    pp = zp.setup(n, code)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for k in range(4):
        x = synth.binary_polys(70 + k, n, D)
        h, c = zp.commit(pp, [torch.from_numpy(x).to(dev)], stream=sa if k % 2 else sb)
        assert c.root == oracle.commit("raa", x.view(np.uint64), pp.num_rows, row_len,
                                       D=D)["root"]
        h.free()
