"""The x sweep's division by a launch-invariant divisor (FastDiv in
sonicsolver_amd/csrc/sf_device.h): floor(x / d) = (umulhi(x, m) + x) >> l
with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1 (Granlund and
Montgomery, PLDI 1994, section 4).  The same integer steps are taken here
and compared with exact division: every x below 2^16 for the small
divisors, and for divisors up to the largest face count per line, the
x at and next to each multiple of d near 2^32 and random x."""
import numpy as np


def make(d):
    l = 0
    while l < 32 and (1 << l) < d:
        l += 1
    m = ((((1 << l) - d) << 32) // d + 1) & 0xFFFFFFFF
    return m, l


def fast_div(x, m, l):
    x = np.asarray(x, dtype=np.uint64)
    t = (x * np.uint64(m)) >> np.uint64(32)   # umulhi of two 32-bit values
    return (t + x) >> np.uint64(l)


def test_small_divisors_exhaustive():
    x = np.arange(1 << 16, dtype=np.uint64)
    for d in range(1, 600):
        m, l = make(d)
        assert np.array_equal(fast_div(x, m, l), x // np.uint64(d)), d


def test_large_range():
    rng = np.random.default_rng(5)
    top = (1 << 32) - 1
    divisors = list(range(1, 70)) + [127, 128, 129, 255, 256, 257, 513, 1024, 2049, 4097,
                                      8193, 65535, 65536, 65537, 1 << 20, (1 << 31) - 1,
                                      1 << 31, top]
    divisors += [int(v) for v in rng.integers(2, 1 << 31, 200)]
    for d in divisors:
        m, l = make(d)
        assert 0 <= m < (1 << 32) and 0 <= l <= 32
        q = top // d
        xs = [0, 1, d - 1, d, d + 1, top, top - 1]
        for k in (q, q - 1, max(q // 2, 1)):
            xs += [k * d - 1, k * d, k * d + d - 1]
        xs = np.array([x for x in xs if 0 <= x <= top], dtype=np.uint64)
        xs = np.concatenate([xs, rng.integers(0, top, 2000, dtype=np.uint64, endpoint=True)])
        assert np.array_equal(fast_div(xs, m, l), xs // np.uint64(d)), d
