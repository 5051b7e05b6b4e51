"""Transcript absorbs of device-resident entries
(zp_transcript_absorb_entries_device) against the host absorb of the same
bytes: the full BLAKE3 chunks of the framed stream (0x6 || entry || 0x7 per
entry, transcript/src/traits.rs:388-392, protocol/src/lib.rs:305-316) are
hashed on the GPU, so the chunk counters, the chaining-value stack and the
host-side head and tail must line up for every starting offset."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _pair(zg, prefix):
    a, b = zg.PcsProverTranscript(), zg.PcsProverTranscript()
    if prefix is not None:
        a.absorb_slice(prefix)
        b.absorb_slice(prefix)
    return a, b


# Stream positions before the absorb: none, 2, 1024 (a full chunk, no tail),
# 1025, 1026, 2048 and 3000 bytes (absorb_slice adds two tag bytes).
PREFIXES = [None, b"", bytes(1022), bytes(1023), bytes(1024), bytes(2046), bytes(2998)]


@pytest.mark.parametrize("entry,count", [(256, 1 << 14), (8, 5000), (40, 24), (0, 20000),
                                         (1, 20000), (3000, 40), (1022, 64)])
@pytest.mark.parametrize("pi", range(len(PREFIXES)))
def test_device_absorb_matches_host(zg, dev, entry, count, pi):
    import torch
    say("device absorb", entry, count, pi)
    rng = np.random.default_rng(entry * 7 + count + pi)
    data = rng.integers(0, 256, entry * count, dtype=np.uint8)
    host, gpu = _pair(zg, PREFIXES[pi])
    host.absorb_entries(data.tobytes(), entry)
    gpu.absorb_entries(torch.from_numpy(data).to(dev), entry)
    assert gpu.get_challenge(64) == host.get_challenge(64)
    # The chaining-value stack left behind must keep agreeing.
    tail = bytes(rng.integers(0, 256, 3000, dtype=np.uint8))
    host.absorb_slice(tail)
    gpu.absorb_slice(tail)
    assert gpu.squeeze_challenge_idx(1 << 20) == host.squeeze_challenge_idx(1 << 20)
    assert gpu.get_challenge(32) == host.get_challenge(32)


def test_device_absorb_matches_oracle(zg, dev, oracle):
    # A DensePolynomial<i64, 32> public column of 2^14 evals (256-byte
    # entries) after the three roots of step 0.
    import torch
    say("device absorb vs oracle")
    rng = np.random.default_rng(5)
    roots = [bytes(rng.integers(0, 256, 32, dtype=np.uint8)) for _ in range(3)]
    data = rng.integers(-2**40, 2**40, (1 << 14) * 32, dtype=np.int64)
    t, o = zg.PcsProverTranscript(), oracle.Transcript()
    for r in roots:
        t.absorb_slice(r)
        o.absorb_slice(r)
    t.absorb_entries(torch.from_numpy(data).to(dev), 256)
    raw = data.tobytes()
    for i in range(1 << 14):
        o.absorb_slice(raw[256 * i:256 * (i + 1)])
    assert t.get_challenge(32) == o.challenge(32)


def test_device_absorb_rejects_partial_entry(zg, dev):
    import torch
    t = zg.PcsProverTranscript()
    with pytest.raises(zg.ShapeError):
        t.absorb_entries(torch.zeros(10, dtype=torch.uint8, device=dev), 4)
