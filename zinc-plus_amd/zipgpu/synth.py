"""Seeded synthetic inputs for the commitment workloads.

The reference benches draw inputs from an unseeded ThreadRng
(zip-plus/benches/zip_common.rs:116,197), so reproducible workloads need a
generator of their own.  This one is SplitMix64 in counter form: word i of
stream `seed` is mix(seed + (i + 1) * 0x9E3779B97F4A7C15), which numpy
evaluates for a whole array at once.  The same bytes feed the device and the
CPU checker.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n, offset=0):
    """n u64 words of stream `seed`, starting at word `offset`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def binary_polys(seed, n, degree_plus_one):
    """n BinaryPoly<D> evaluations: the low D bits are i.i.d. Bernoulli(1/2)
    (the StandardUniform sample of binary_u64.rs:282-294), returned as the
    int64 bit pattern of the u64 lane."""
    w = splitmix64(seed, n)
    if degree_plus_one < 64:
        w &= np.uint64((1 << degree_plus_one) - 1)
    return w.view(np.int64)


def i32_evals(seed, n):
    """n i32 evaluations uniform over all 2^32 bit patterns."""
    return (splitmix64(seed, n) >> np.uint64(32)).astype(np.uint32).view(np.int32)


def i128_words(seed, n):
    """n i128 values uniform over all 2^128 bit patterns, as int64 [n, 2]
    little-endian words."""
    return splitmix64(seed, 2 * n).view(np.int64).reshape(n, 2)


def comb_rows(seed, n, limbs, small_words=0):
    """n CombR = Int<limbs> values as int64 [n, limbs] little-endian limbs.

    small_words = 0: uniform over all 2^(64 limbs) bit patterns.  Otherwise
    uniform signed values of small_words 64-bit words, sign-extended (the
    range of a combined row built from i128 sums)."""
    if small_words == 0:
        return splitmix64(seed, n * limbs).view(np.int64).reshape(n, limbs)
    w = splitmix64(seed, n * small_words).view(np.int64).reshape(n, small_words)
    ext = np.where(w[:, -1:] < 0, np.int64(-1), np.int64(0))
    return np.concatenate([w, np.repeat(ext, limbs - small_words, axis=1)], axis=1)


def row_split(kind, n, max_iprs_row=1 << 14):
    """Row length of a commitment of n evals.

    RAA: isqrt(n).next_power_of_two() (zip-plus/benches/zip_plus_benches.rs:84).
    IPRS: min(n, 2^14), the largest row the F65537 pNTT takes at rate 1/4
    (rows of 2^15 and above give codewords of 2^17 > 65537 points,
    code/iprs/pntt/radix8/params.rs:59-103).
    """
    if kind == "iprs":
        return min(n, max_iprs_row)
    r = int(np.sqrt(n))
    while r * r > n:
        r -= 1
    while (r + 1) * (r + 1) <= n:
        r += 1
    p = 1
    while p < r:
        p <<= 1
    return p


def request_mix(seed, count, log2_lo, log2_hi):
    """`count` commit requests: log2 n uniform over log2_lo..log2_hi, D in
    {32, 64} and code in {IPRS, RAA} each with probability 1/2, from one
    SplitMix64 stream.  -> list of dict(kind, n, D, row_len, seed)."""
    w = splitmix64(seed, 3 * count).reshape(count, 3)
    span = log2_hi - log2_lo + 1
    out = []
    for i in range(count):
        log2n = log2_lo + int(w[i, 0] % np.uint64(span))
        D = 32 if int(w[i, 1] >> np.uint64(63)) == 0 else 64
        kind = "iprs" if int(w[i, 2] >> np.uint64(63)) == 0 else "raa"
        n = 1 << log2n
        out.append({"kind": kind, "n": n, "D": D, "row_len": row_split(kind, n),
                    "seed": (seed * 1000003 + i) & 0xFFFFFFFF})
    return out
