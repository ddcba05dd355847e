"""GPU parity of the prime / pi-table / LPF / Moebius / Phi drop-ins against
the oracle (CudaSieve countPrimes + getDevicePrimes, get_d_piTable,
class PiTable, gen_d_lpf / gen_d_mu, class Phi, divXbyY)."""
import random

import pytest

pytestmark = pytest.mark.gpu


def test_wave_register_scan(ctx):
    # the DPP scan/sum used by every segment stage, lane by lane against a serial scan
    for seed in (1, 2, 3):
        print("case selftest_wave seed %d" % seed, flush=True)
        assert ctx.selftest_wave(seed, 4096) == 0


def test_count_primes_ranges(ctx, oracle):
    rnd = random.Random(3)
    cases = [(0, 10), (0, 2), (0, 3), (2, 3), (3, 100), (0, 10 ** 6), (10 ** 6, 10 ** 6 + 1000),
             (10 ** 9, 10 ** 9 + 10 ** 7), (2 ** 32 - 1000, 2 ** 32 + 1000)]
    for _ in range(10):
        lo = rnd.randrange(0, 10 ** 10)
        cases.append((lo, lo + rnd.randrange(1, 10 ** 6)))
    for lo, hi in cases:
        print("case countPrimes[%d,%d)" % (lo, hi), flush=True)
        assert ctx.countPrimes(lo, hi) == oracle.count_primes(lo, hi), (lo, hi)


def test_count_primes_brute_1e11(ctx, oracle):
    # the plain segmented-Eratosthenes count (beyond the resident table)
    assert ctx.countPrimes(0, 10 ** 11 + 1) == 4118054813


def test_get_device_primes(ctx, oracle):
    assert ctx.getDevicePrimes(0, 100) == oracle.primes(99)
    got = ctx.getDevicePrimes(10 ** 6, 10 ** 6 + 5000)
    ref = [p for p in oracle.primes(10 ** 6 + 4999) if p >= 10 ** 6]
    assert got == ref


def _sieve_range(lo, hi, oracle):
    """Primes of [lo, hi) by numpy Eratosthenes with the oracle's primes to sqrt(hi)."""
    import math
    import numpy as np
    a = np.ones(hi - lo, dtype=bool)
    for p in oracle.primes(math.isqrt(hi) + 1):
        s0 = max(p * p, (lo + p - 1) // p * p)
        a[s0 - lo::p] = False
    return [lo + int(i) for i in np.flatnonzero(a) if lo + int(i) >= 2]


def _is_prime_u64(n):
    """Deterministic Miller-Rabin for n < 2^64 (bases 2, 325, 9375, 28178,
    450775, 9780504, 1795265022)."""
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, r = n - 1, 0
    while d % 2 == 0:
        d //= 2
        r += 1
    for a in (2, 325, 9375, 28178, 450775, 9780504, 1795265022):
        a %= n
        if a == 0:
            continue
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_range_primes_segmented(oracle):
    # ranges above 2^24 the pi table does not reach go through the segmented
    # bit sieve (2^20 integers per block, sieving primes to sqrt(hi) only):
    # across segment and word boundaries, unaligned ends, one prime, none
    # (a fresh context, whose pi table reaches none of them)
    import cudapix_amd as C
    with C.Context(device=0) as ctx:
        _range_primes(ctx, oracle)


def _range_primes(ctx, oracle):
    cases = [(2 ** 24 - 1000, 2 ** 24 + 5000), (10 ** 12, 10 ** 12 + 3 * 2 ** 20 + 7),
             (10 ** 12 + 37, 10 ** 12 + 37 + 2 ** 20 + 13), (10 ** 13 + 1, 10 ** 13 + 2),
             (10 ** 13 + 37, 10 ** 13 + 38), (3 * 10 ** 13 + 1, 3 * 10 ** 13 + 64)]
    for lo, hi in cases:
        print("case getDevicePrimes[%d,%d)" % (lo, hi), flush=True)
        want = _sieve_range(lo, hi, oracle)
        assert ctx.getDevicePrimes(lo, hi) == want, (lo, hi)
        assert ctx.countPrimes(lo, hi) == len(want)
    assert ctx.getHostPrimes(10 ** 12, 10 ** 12 + 100000) == _sieve_range(10 ** 12, 10 ** 12 + 100000, oracle)
    assert ctx.getHostPrimes(0, 100) == oracle.primes(99)
    # near 2^63: every odd n of the range classified by Miller-Rabin
    lo, hi = 2 ** 63 - 30001, 2 ** 63 + 10000
    print("case getDevicePrimes near 2^63", flush=True)
    got = ctx.getDevicePrimes(lo, hi)
    assert got == [n for n in range(lo, hi) if _is_prime_u64(n)]


def test_device_primes32_and_bit_sieve(ctx, oracle):
    import cudapix_amd as C
    assert ctx.getDevicePrimes32(0, 10 ** 6) == oracle.primes(10 ** 6 - 1)
    assert ctx.getDevicePrimes32(2 ** 32 - 100000, 2 ** 32) == _sieve_range(2 ** 32 - 100000, 2 ** 32, oracle)
    with pytest.raises(C.CpxError) as e:
        ctx.getDevicePrimes32(0, 2 ** 32 + 1)
    assert e.value.code == -4
    # genDeviceBitSieve(0, y) as the hard-leaf code reads it: bit (m % 64) / 2
    # of word m / 64 is clear iff odd m < y is prime
    for lo, hi in ((0, 10 ** 6 + 7), (2 ** 40, 2 ** 40 + 3 * 2 ** 20 + 100)):
        print("case genDeviceBitSieve[%d,%d)" % (lo, hi), flush=True)
        words = ctx.genDeviceBitSieve(lo, hi)
        assert len(words) == (hi - lo + 63) // 64
        pr = set(_sieve_range(lo, hi, oracle))
        for w, word in enumerate(words):
            for i in range(32):
                n = lo + 64 * w + 2 * i + 1
                want = 0 if (n < hi and n in pr) else 1
                assert (word >> i) & 1 == want, n
    with pytest.raises(C.CpxError) as e:
        ctx.genDeviceBitSieve(10, 1000)
    assert e.value.code == -1


def test_get_d_pitable_layout(ctx, oracle):
    hi = 200000
    tab = ctx.get_d_piTable(hi)
    pr = oracle.primes(hi)
    import bisect
    for k in list(range(0, 200)) + list(range(hi // 2 - 200, hi // 2)):
        n = 2 * k - 1
        want = bisect.bisect_right(pr, n) if n >= 1 else 0
        assert tab[k] == want, k


def test_pitable_windows(ctx, oracle):
    import bisect
    import cudapix_amd as C
    pr = oracle.primes(3 * 10 ** 6)
    t = C.PiTable(ctx, 10 ** 6, 0, 10 ** 6)
    for get in (t.getCurrent, t.getNextUp, t.getNextDown, t.getNextDown):
        base, pb, vals = get()
        assert pb == bisect.bisect_right(pr, base - 1)
        for i in (0, 1, 2, 7, len(vals) // 2, len(vals) - 1):
            n = base + 2 * i - 1
            assert vals[i] == bisect.bisect_right(pr, n) - pb
    t.close()


def test_lpf_mu_windows(ctx, oracle):
    mu, lpf = oracle.mu_lpf_odd(400001)
    got_l = ctx.gen_lpf(0, 400002)
    got_m = ctx.gen_mu(0, 400002)
    assert got_m == mu[:len(got_m)]
    assert [v if v != 2 ** 32 - 1 else 2 ** 64 - 1 for v in got_l] == lpf[:len(got_l)]
    # an offset window
    lo = 300000
    gl = ctx.gen_lpf(lo, lo + 100000)
    gm = ctx.gen_mu(lo, lo + 100000)
    assert gl == lpf[lo // 2: lo // 2 + len(gl)]
    assert gm == mu[lo // 2: lo // 2 + len(gm)]


def test_phi_class(ctx, oracle):
    import cudapix_amd as C
    ph = C.Phi(ctx, 100000)
    rnd = random.Random(9)
    for _ in range(30):
        x = rnd.randrange(0, 99999)
        a = rnd.randrange(0, 12)
        assert ph.phi(x, a) == oracle.phi_legendre(x, a), (x, a)
    rng = ph.generateRange(20000, 4)
    for i in (0, 1, 2, 50, 9999):
        assert rng[i] == oracle.phi_legendre(2 * i + 1, 4)
    ph.close()


def test_div_x_by_y(ctx):
    # the quotient lands in a uint64 array (include/P2.cuh:21): exact when it
    # fits 64 bits, reduced mod 2^64 otherwise (x / 3 here)
    x = 10 ** 20 + 12345
    ys = [3, 17, 2 ** 32 + 15, 10 ** 12, 2 ** 63 + 5]
    assert ctx.divXbyY(x, ys) == [(x // v) % 2 ** 64 for v in ys]
    assert all(x // v < 2 ** 64 for v in ys[1:])


def test_device_bytes_keep_the_resident_wheels():
    """A context with fresh tables rebuilds its size-dependent tables for every
    query; the bootstrap tables and wheels stay allocated and stay counted, so
    the reported device bytes are the same after each solve and include the
    c = 8 wheel (4849845 eight-byte pattern words)."""
    import cudapix_amd as C
    with C.Context(device=0, c=8, fresh_tables=True) as fc:
        assert fc.pi(10 ** 12) == 37607912018
        b1 = fc.device_bytes()
        assert fc.pi(10 ** 12) == 37607912018
        b2 = fc.device_bytes()
    assert b1 == b2
    assert b1 >= 8 * 4849845


def _hip_free_bytes():
    import ctypes
    h = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_allocation_failure_is_clean():
    """A device allocation that fails (forced from the host on the 1st, 2nd,
    3rd, 5th and 8th allocation of a solve that builds tables and a batch
    arena) gives CPX_ENOMEM, the context then solves correctly, and after
    the context is destroyed the device holds no more memory than before."""
    import cudapix_amd as C
    want = 346065536839
    free0 = _hip_free_bytes()
    failed = 0
    with C.Context(device=0, fresh_tables=True) as fc:
        for n in (1, 2, 3, 5, 8):
            C.debug_fail_alloc(n)
            try:
                v = fc.pi(10 ** 13)
                assert v == want                   # the solve needed fewer allocations
            except C.CpxError as e:
                assert e.code == -2, e             # CPX_ENOMEM
                failed += 1
            finally:
                C.debug_fail_alloc(0)
            assert fc.pi(10 ** 13) == want
        b_after = fc.device_bytes()
    with C.Context(device=0, fresh_tables=True) as ref:
        assert ref.pi(10 ** 13) == want
        b_ref = ref.device_bytes()
    assert failed >= 3
    assert b_after == b_ref
    assert _hip_free_bytes() >= free0 - (64 << 20)
