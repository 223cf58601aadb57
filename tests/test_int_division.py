"""The exact integer division by an invariant divisor that G2P uses to turn cell keys into rows and
planes (floor_div in csrc/mpm_kernels.hip, div_magic in csrc/mpm_runtime.hip; Granlund and
Montgomery, PLDI 1994): with l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1 and
t = floor(n m / 2^32), floor(n / d) = (t + ((n - t) >> 1)) >> (l - 1) for every 32-bit n. Checked
here against Python's floor division for every divisor up to 1024 (grid resolutions), large and
power-of-two divisors, and numerators at and around every multiple boundary plus the largest keys."""
import numpy as np


def magic(d):
    l = 0
    while (1 << l) < d:
        l += 1
    m = 0 if l == 0 else ((1 << 32) * ((1 << l) - d)) // d + 1
    assert 0 <= m < (1 << 32)
    return m, l


def div(n, m, l):
    n = np.asarray(n, dtype=np.uint64)
    if l == 0:
        return n
    t = (n * np.uint64(m)) >> np.uint64(32)
    return (t + ((n - t) >> np.uint64(1))) >> np.uint64(l - 1)


def test_magic_division_is_exact():
    rng = np.random.default_rng(0)
    divisors = list(range(1, 1025)) + [2 ** k for k in range(11, 31)] + [int(x) for x in rng.integers(1025, 2 ** 31, 200)]
    for d in divisors:
        m, l = magic(d)
        base = np.array([0, 1, d - 1, d, d + 1, 2 ** 31 - 1, 2 ** 32 - 1], dtype=np.uint64)
        q = np.arange(0, 2 ** 32 // d, max(1, (2 ** 32 // d) // 500), dtype=np.uint64) * np.uint64(d)
        near = np.concatenate([q, q + np.uint64(d - 1), q[1:] - np.uint64(1)])
        n = np.concatenate([base, near, rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)])
        n = n[n < 2 ** 32]
        assert np.array_equal(div(n, m, l), n // np.uint64(d)), d
