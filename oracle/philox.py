"""CPU restatement of the engine's device random numbers (test infrastructure:
only tests/ use it, as the checker).

The reference draws from numpy's default_rng; the engine's device mode uses
Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
as 1, 2, 3", SC11): 10 rounds of two 32x32->64 multiplications (0xD2511F53,
0xCD9E8D57) with the Weyl key schedule (0x9E3779B9, 0xBB67AE85).  Pinned by
the published Random123 known-answer vectors (tests/test_philox_kat.py).
The counter layout and the conversions restate the engine's specification:
counter = (element pair or 0, chain, proposal low word,
stream << 24 | proposal high 24 bits), key = the 64-bit seed (low, high);
53-bit uniforms from two words; Box-Muller with u1 = 1 - U in (0, 1].
"""
import math

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF
STREAM_MOMENTUM, STREAM_SCALARS, STREAM_EXCHANGE = 0, 1, 2


def philox4x32_10(counter, key):
    c = [int(v) & MASK for v in counter]
    k0, k1 = int(key[0]) & MASK, int(key[1]) & MASK
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k0, p1 & MASK, (p0 >> 32) ^ c[3] ^ k1, p0 & MASK]
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return c


def u53(a, b):
    """53-bit uniform in [0, 1) from two 32-bit words."""
    return ((a >> 5) * 67108864.0 + (b >> 6)) * (1.0 / 9007199254740992.0)


def _counter(first, chain, prop, stream):
    return [first, chain, prop & MASK, (stream << 24) | ((prop >> 32) & 0xFFFFFF)]


def _key(seed):
    return [seed & MASK, (seed >> 32) & MASK]


def normal_pair(seed, chain, prop, pair):
    """Momentum normals of elements 2*pair and 2*pair+1 of a chain."""
    r = philox4x32_10(_counter(pair, chain, prop, STREAM_MOMENTUM), _key(seed))
    u1 = 1.0 - u53(r[0], r[1])
    u2 = u53(r[2], r[3])
    rad = math.sqrt(-2.0 * math.log(u1))
    ang = 6.283185307179586476925286766559 * u2
    return rad * math.cos(ang), rad * math.sin(ang)


def scalars(seed, chain, prop):
    """(step jitter in [0.5, 1.5), accept uniform in [0, 1))."""
    r = philox4x32_10(_counter(0, chain, prop, STREAM_SCALARS), _key(seed))
    return 0.5 + u53(r[0], r[1]), u53(r[2], r[3])
