"""zp_commit_noalloc (the commit with its scratch in a static device array)
against zp_commit, through the C ABI, bit for bit: codewords, every Merkle
layer and the root, on two calls in a row (the second smaller, with another
code and a batch), then an oversized commit that is refused before any
work."""
import ctypes

import numpy as np
import pytest
import torch

from zipgpu import _lib, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def hip():
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMemcpy.restype = ctypes.c_int
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


def noalloc_commit(zg, code, D, x, num_rows, batch, stream):
    lib = _lib.load()
    N = code.codeword_len()
    elem = code.cw_elem_bytes(zg.BinPolyTypes(D))
    cw = torch.empty(batch * num_rows * N * elem, dtype=torch.uint8, device=x.device)
    layers = torch.empty((2 * N - 1) * 32, dtype=torch.uint8, device=x.device)
    torch.cuda.synchronize()
    rc = lib.zp_commit_noalloc(code._h, _lib.ZP_EVAL_BINPOLY, D, ctypes.c_void_p(x.data_ptr()),
                               num_rows, batch, ctypes.c_void_p(cw.data_ptr()),
                               ctypes.c_void_p(layers.data_ptr()),
                               ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, _lib.last_error()
    stream.synchronize()
    return cw, layers


def plain_commit(zg, pp, D, polys):
    """zp_commit, with the hint's codewords and layers copied to device tensors."""
    zp = zg.ZipPlus(zg.BinPolyTypes(D))
    hint, comm = zp.commit(pp, polys)
    N = hint.num_leaves
    cw = torch.empty(hint.cw_bytes_per_poly * hint.batch, dtype=torch.uint8, device=polys[0].device)
    assert hip().hipMemcpy(cw.data_ptr(), hint.device_codewords_ptr(), cw.numel(), 3) == 0
    layers = np.concatenate([hint.layer(level) for level in range(hint.height)])
    hint.free()
    assert layers.shape == (2 * N - 1, 32)
    return comm.root, cw, layers


def check(zg, kind, n, D, batch, row_len, seed, dev, stream):
    code = (zg.IprsCode.new_with_optimal_depth(row_len) if kind == "iprs"
            else zg.RaaCode(row_len))
    pp = zg.ZipPlus(zg.BinPolyTypes(D)).setup(n, code)
    x = torch.from_numpy(synth.binary_polys(seed, batch * n, D)).to(dev)
    polys = list(x.reshape(batch, -1))
    say("noalloc", kind, n, D, batch)
    cw, layers = noalloc_commit(zg, code, D, x, pp.num_rows, batch, stream)
    root, cw_ref, layers_ref = plain_commit(zg, pp, D, polys)
    assert torch.equal(cw, cw_ref), "codewords differ from zp_commit"
    got_layers = layers.cpu().numpy().reshape(-1, 32)
    assert np.array_equal(got_layers, layers_ref), "Merkle layers differ from zp_commit"
    assert got_layers[-1].tobytes() == root


def test_noalloc_commit_equals_commit_twice_then_refuses_oversize(zg, dev):
    stream = torch.cuda.Stream()
    # The benchmarked commit: 2^20 BinaryPoly<32>, IPRS rows of 2^14 (1 GiB
    # of codewords, 8 MiB of bit-plane words, 48 MiB of chunk values).
    check(zg, "iprs", 1 << 20, 32, 1, 1 << 14, 501, dev, stream)
    # Smaller, another code, two polys under one root.
    check(zg, "raa", 1 << 14, 64, 2, 1 << 7, 502, dev, stream)
    # 2^22 BinaryPoly<64> under RAA rows of 2^11: 512 chunks per column need
    # 192 MiB of chunk values, more than the static buffer.
    code = zg.RaaCode(1 << 11)
    N, rows = code.codeword_len(), 1 << 11
    x = torch.zeros(1, dtype=torch.int64, device=dev)
    out = torch.zeros(1, dtype=torch.uint8, device=dev)
    rc = _lib.load().zp_commit_noalloc(code._h, _lib.ZP_EVAL_BINPOLY, 64,
                                       ctypes.c_void_p(x.data_ptr()), rows, 1,
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(stream.cuda_stream))
    assert rc == _lib.ZP_ERR_OOM and b"static buffer" in _lib.load().zp_last_error()
    assert N == 8192
    # Unsupported eval types are refused too.
    rc = _lib.load().zp_commit_noalloc(code._h, _lib.ZP_EVAL_I32, 1,
                                       ctypes.c_void_p(x.data_ptr()), rows, 1,
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(stream.cuda_stream))
    assert rc == _lib.ZP_ERR_INVALID_PARAM
