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
