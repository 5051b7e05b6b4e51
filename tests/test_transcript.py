"""The library's Fiat-Shamir transcript (host code behind the C ABI) against
the oracle's: the same absorbs give the same challenges, indices, prime and
stream bytes (transcript/src/lib.rs:10-72, transcript/src/traits.rs:358-412,
zip-plus/src/pcs_transcript.rs:27-221).  CPU only: no device calls."""
import random

import pytest


@pytest.fixture(scope="module")
def zgh():
    import zipgpu
    return zipgpu


def test_challenges_and_indices_match(zgh, oracle):
    rng = random.Random(1)
    t, o = zgh.PcsProverTranscript(), oracle.Transcript()
    for step in range(400):
        op = rng.randrange(4)
        if op == 0:
            data = bytes(rng.randrange(256) for _ in range(rng.choice([0, 1, 31, 32, 64, 65, 700])))
            t.absorb_slice(data)
            o.absorb_slice(data)
        elif op == 1:
            n = rng.choice([1, 4, 16, 32, 33, 64, 65, 200])
            assert t.get_challenge(n) == o.challenge(n), step
        elif op == 2:
            cap = rng.choice([1, 7, 1 << 16, 65536 * 3 + 1])
            assert t.squeeze_challenge_idx(cap) == o.squeeze_idx(cap), step
        else:
            assert t.get_challenge_i128() == o.challenge_i128(), step


@pytest.mark.parametrize("sizes", [[1024], [1023, 1, 1], [5000], [1024 * 7 + 3, 64, 2049],
                                   [3] * 800, [64] * 40])
def test_incremental_hash_across_chunks(zgh, oracle, sizes):
    # Absorbs that fill and cross 1 KiB chunks exercise the chaining-value
    # stack; the challenge must equal the XOF of the whole absorbed string.
    rng = random.Random(len(sizes))
    t, o = zgh.PcsProverTranscript(), oracle.Transcript()
    for n in sizes:
        data = bytes(rng.randrange(256) for _ in range(n))
        t.absorb_slice(data)
        o.absorb_slice(data)
    assert t.get_challenge(131) == o.challenge(131)


@pytest.mark.parametrize("sizes", [[8190], [8191], [8192], [8193], [9000], [8192 * 3 + 5],
                                   [1, 8192 * 2 + 1, 20000], [1021, 65536 + 7], [300000],
                                   [1024 * 8 - 2, 1024 * 8 - 2]])
def test_multi_chunk_absorbs(zgh, oracle, sizes):
    # Long absorbs go eight chunks at a time (SIMD lanes) once the hasher
    # sits at a chunk boundary with more than eight chunks left; the oracle
    # hashes one block at a time.  Sizes straddle that threshold from every
    # starting offset (each absorb_slice adds two tag bytes).
    import numpy as np
    rng = np.random.default_rng(sum(sizes))
    t, o = zgh.PcsProverTranscript(), oracle.Transcript()
    for n in sizes:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        t.absorb_slice(data)
        o.absorb_slice(data)
        assert t.get_challenge(32) == o.challenge(32), n
    entries = rng.integers(0, 256, 256 * 4096, dtype=np.uint8).tobytes()
    t.absorb_entries(entries, 256)
    for i in range(4096):
        o.absorb_slice(entries[256 * i:256 * (i + 1)])
    assert t.get_challenge(64) == o.challenge(64)


@pytest.mark.parametrize("entry,count", [(300000, 2), (262142, 3), (262143, 2), (1, 70000)])
def test_absorb_entries_bulk_and_large(zgh, oracle, entry, count):
    # Entries are framed into 256 KiB pieces; a framed entry larger than a
    # piece (entry + 2 tag bytes > 262144) is absorbed entry by entry.
    import numpy as np
    data = np.random.default_rng(entry + count).integers(0, 256, entry * count,
                                                          dtype=np.uint8).tobytes()
    t, o = zgh.PcsProverTranscript(), oracle.Transcript()
    t.absorb_slice(b"x")
    o.absorb_slice(b"x")
    t.absorb_entries(data, entry)
    for i in range(count):
        o.absorb_slice(data[entry * i:entry * (i + 1)])
    assert t.get_challenge(32) == o.challenge(32)


def test_random_field_cfg_matches(zgh, oracle):
    for seed in range(4):
        root = bytes((seed * 37 + i) % 256 for i in range(32))
        t, o = zgh.PcsProverTranscript(), oracle.Transcript()
        t.absorb_slice(root)
        o.absorb_slice(root)
        assert t.get_random_field_cfg() == o.random_field_cfg()
        assert t.get_challenge(32) == o.challenge(32)


def test_clone_and_stream(zgh, oracle):
    t = zgh.PcsProverTranscript()
    t.absorb_slice(b"root")
    t.write(b"abc")
    u = t.clone()
    assert u.get_challenge(16) == t.get_challenge(16)
    u.write(b"de")
    assert t.stream() == b"abc" and u.stream() == b"abcde"


def test_squeeze_zero_cap_is_an_error(zgh):
    t = zgh.PcsProverTranscript()
    with pytest.raises(zgh.ShapeError):
        t.squeeze_challenge_idx(0)


def test_absorb_entries_is_absorb_slice_per_entry(zgh, oracle):
    import numpy as np
    data = bytes(np.random.default_rng(3).integers(0, 256, 40 * 24, dtype=np.uint8))
    t, o = zgh.PcsProverTranscript(), oracle.Transcript()
    t.absorb_entries(data, 40)
    for i in range(24):
        o.absorb_slice(data[40 * i:40 * (i + 1)])
    assert t.get_challenge(32) == o.challenge(32)


_SCALAR_TRACE = r'''
import sys
import numpy as np
sys.path[:0] = [sys.argv[1], sys.argv[2]]
import zipgpu
rng = np.random.default_rng(7)
t = zipgpu.PcsProverTranscript()
out = []
for n in [0, 1, 63, 64, 65, 1023, 1024, 1025, 8193, 16 * 1024 + 1, 40000, 300000]:
    t.absorb_slice(rng.integers(0, 256, n, dtype=np.uint8).tobytes())
    out.append(t.get_challenge(16).hex())
    out.append(t.get_challenge(100).hex())
    out.append(str(t.squeeze_challenge_idx(1000003)))
print(" ".join(out))
'''


def test_scalar_and_simd_hashers_agree(zgh, oracle):
    # ZIPGPU_HOST_SIMD=0 keeps the hasher on its scalar compression and
    # one-chunk-at-a-time hashing (a host without AVX2); this process uses
    # the SIMD paths the host has.  Both must give the oracle's challenges.
    import os
    import subprocess
    import sys
    import numpy as np
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    runs = []
    for simd in ("0", "1"):
        env = dict(os.environ, ZIPGPU_HOST_SIMD=simd)
        r = subprocess.run([sys.executable, "-c", _SCALAR_TRACE, root,
                            os.path.join(root, "zinc-plus_amd")],
                           env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        runs.append(r.stdout.split())
    rng = np.random.default_rng(7)
    o = oracle.Transcript()
    want = []
    for n in [0, 1, 63, 64, 65, 1023, 1024, 1025, 8193, 16 * 1024 + 1, 40000, 300000]:
        o.absorb_slice(rng.integers(0, 256, n, dtype=np.uint8).tobytes())
        want.append(o.challenge(16).hex())
        want.append(o.challenge(100).hex())
        want.append(str(o.squeeze_idx(1000003)))
    assert runs[0] == want
    assert runs[1] == want
