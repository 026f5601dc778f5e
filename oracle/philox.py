"""Philox4x64-10 and the draw keying shared by the oracle and the device engine.

The reference draws a random measurement outcome with ``random.choice``
(``sdim/tableau/tableau_prime.py:332``) and noise with ``np.random``
(``sdim/program.py:489-503``).  Neither stream can be reproduced on a GPU, so
both implementations take every draw from a counter-based Philox4x64-10
block keyed by (seed, stream) with counter (op_index, shot, 0, 0):

    out = philox4x64_10(counter=(op_index, shot, 0, 0), key=(seed, stream))

* measurement outcome      m = mulhi64(out[0], d)                in [0, d)
* noise coin               apply iff (out[0] >> 11) >= keep53,
                           keep53 = ceil((1 - prob) * 2**53)
  (the reference's ``u < 1 - prob`` -> identity test, ``program.py:500-502``,
  with u = (out[0] >> 11) / 2**53)
* noise exponents          channel 'd': r = 1 + mulhi64(out[1], d*d - 1),
                           a = r % d, b = r // d   (``program.py:487-491``)
                           channel 'f': a = 1 + mulhi64(out[1], d - 1), b = 0
                           channel 'p': a = 0, b = 1 + mulhi64(out[1], d - 1)

The block function is the published Random123 Philox4x64 with 10 rounds
(Salmon et al., SC'11).  It is pinned against ``numpy.random.Philox``, which
implements the same generator (numpy increments its counter before each block).
"""

MASK64 = (1 << 64) - 1
PHILOX_M0 = 0xD2E7470EE14C6C93
PHILOX_M1 = 0xCA5A826395121157
PHILOX_W0 = 0x9E3779B97F4A7C15
PHILOX_W1 = 0xBB67AE8584CAA73B


def philox4x64_10(counter, key):
    """One Philox4x64-10 block: 4 x uint64 counter, 2 x uint64 key -> 4 x uint64."""
    c0, c1, c2, c3 = (int(v) & MASK64 for v in counter)
    k0, k1 = (int(v) & MASK64 for v in key)
    for r in range(10):
        if r:
            k0 = (k0 + PHILOX_W0) & MASK64
            k1 = (k1 + PHILOX_W1) & MASK64
        p0 = PHILOX_M0 * c0
        p1 = PHILOX_M1 * c2
        hi0, lo0 = p0 >> 64, p0 & MASK64
        hi1, lo1 = p1 >> 64, p1 & MASK64
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return (c0, c1, c2, c3)


def mulhi64(a, b):
    return ((int(a) & MASK64) * int(b)) >> 64


def block(seed, stream, op_index, shot):
    return philox4x64_10((op_index, shot, 0, 0), (seed, stream))


def measurement_outcome(seed, stream, op_index, shot, d):
    """Uniform outcome in [0, d) replacing ``random.choice(range(d))`` (tableau_prime.py:332)."""
    return mulhi64(block(seed, stream, op_index, shot)[0], d)


CHANNEL_CODES = {"d": 0, "f": 1, "p": 2}


def keep_threshold53(prob):
    """Integer threshold: the noise op is the identity iff (out0 >> 11) < keep53.

    Equivalent to the reference's ``u < 1.0 - prob`` (program.py:500-502) for
    u = k / 2**53; (1 - prob) * 2**53 is exact in binary floating point.
    """
    import math
    x = (1.0 - float(prob)) * 9007199254740992.0
    if x <= 0.0:
        return 0
    if x >= 9007199254740992.0:
        return 1 << 53
    return int(math.ceil(x))


def noise_exponents(seed, stream, op_index, shot, d, channel, keep53):
    """(a, b) exponents of the X^a Z^b error of an N1 op, (0, 0) when the coin says identity."""
    out = block(seed, stream, op_index, shot)
    if (out[0] >> 11) < keep53:
        return 0, 0
    if channel == 0:  # 'd'
        r = 1 + mulhi64(out[1], d * d - 1)
        return r % d, r // d
    if channel == 1:  # 'f'
        return 1 + mulhi64(out[1], d - 1), 0
    if channel == 2:  # 'p'
        return 0, 1 + mulhi64(out[1], d - 1)
    raise ValueError("Must specify a valid noise channel.")
