"""The exact division by the Halton bases behind the sampler's digit loop (device_sampling.h).

floor(n / b) is computed as mulhi(m, n) >> s when a 32-bit m = ceil(2^(32+s) / b) satisfies
m * b - 2^(32+s) <= 2^s (Granlund & Montgomery 1994, Theorem 4.2 with N = 32), otherwise through
their 33-bit multiplier (figure 4.1).  The library's table builder applies the same criterion;
these checks confirm, by brute force over boundary and random 32-bit numerators, that both forms
equal integer division for each of the first 128 primes.
"""
import numpy as np


def _primes(n):
    ps, x = [], 2
    while len(ps) < n:
        if all(x % p for p in ps if p * p <= x):
            ps.append(x)
        x += 1
    return ps


def _simple(b):
    for s in range(32):
        p2 = 1 << (32 + s)
        m = -(-p2 // b)
        if m < (1 << 32) and m * b - p2 <= (1 << s):
            return m, s
    return None


def _numerators(b, rng):
    k = np.arange(0, (1 << 32) // b, max(1, ((1 << 32) // b) // 4096), dtype=np.uint64)
    edge = np.concatenate([k * b, k * b + b - 1, k * b + b // 2]).astype(np.uint64)
    rnd = rng.integers(0, 1 << 32, 20000, dtype=np.uint64)
    top = np.array([(1 << 32) - 1, (1 << 32) - 2, (((1 << 32) - 1) // b) * b], np.uint64)
    n = np.concatenate([edge, rnd, top])
    return n[n < (1 << 32)]


def test_single_multiplier_division_is_exact():
    rng = np.random.default_rng(11)
    count = 0
    for b in _primes(128):
        ms = _simple(b)
        if ms is None:
            continue
        count += 1
        m, s = ms
        n = _numerators(b, rng)
        q = (n * np.uint64(m)) >> np.uint64(32 + s)     # n < 2^32, m < 2^32: fits in 64 bits
        assert np.array_equal(q, n // np.uint64(b)), b
    assert count == 94


def test_33_bit_multiplier_division_is_exact():
    rng = np.random.default_rng(12)
    for b in _primes(128):
        l = (b - 1).bit_length()                          # ceil(log2 b)
        magic = (((1 << l) - b) << 32) // b + 1
        sh1, sh2 = min(l, 1), max(l - 1, 0)
        n = _numerators(b, rng)
        t = (n * np.uint64(magic)) >> np.uint64(32)
        q = (t + ((n - t) >> np.uint64(sh1))) >> np.uint64(sh2)
        assert np.array_equal(q, n // np.uint64(b)), b
