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
