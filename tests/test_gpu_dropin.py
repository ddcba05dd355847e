"""GPU parity of the drop-in pieces added around the terms (tolerance 0):
64-bit LPF / Moebius windows (gen_d_lpf<uint64_t>, gen_d_mu with 64-bit
bottoms, src/sieve/lpf_mu.cu:57-87,153-197) against the oracle's trial
division and sympy; the Phi getters and makeModCache (include/phi.cuh:23-32);
PiTable set_pi_base / reallocate / getMaxRange (include/pitable.cuh:34-43)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lo,hi", [(0, 40002), (2 ** 32 - 4000, 2 ** 32 + 4000), (2 ** 40, 2 ** 40 + 20000),
                                   (2 ** 50 + 10 ** 6, 2 ** 50 + 10 ** 6 + 4000)])
def test_lpf64_and_mu_windows(ctx, oracle, lo, hi):
    print("case lpf64 window [%d, %d)" % (lo, hi), flush=True)
    mu, lpf = oracle.mu_lpf_window(lo, hi)
    assert ctx.gen_lpf64(lo, hi) == lpf
    assert ctx.gen_mu(lo, hi) == mu
    if hi < 2 ** 32:                       # the 32-bit entries agree where they apply
        assert [v if v != 2 ** 32 - 1 else 2 ** 64 - 1 for v in ctx.gen_lpf(lo, hi)] == lpf


def test_lpf64_window_near_2_62(ctx):
    sympy = pytest.importorskip("sympy")
    lo = 2 ** 62
    print("case lpf64 window at 2^62", flush=True)
    lpf = ctx.gen_lpf64(lo, lo + 2000)
    mu = ctx.gen_mu(lo, lo + 2000)
    for i in range(0, 1000, 37):
        f = sympy.factorint(lo + 2 * i + 1)
        assert lpf[i] == min(f)
        assert mu[i] == (0 if max(f.values()) > 1 else (-1) ** len(f))


def test_phi_getters_and_mod_cache(ctx, oracle):
    import cudapix_amd as C
    mu, lpf = oracle.mu_lpf_odd(99999)
    primes = [p for p in range(2, 1001) if all(p % d for d in range(2, int(p ** 0.5) + 1))]
    for pmax, bound in ((0, 316), (1000, 1000)):
        ph = C.Phi(ctx, 100000, pmax)
        try:
            got = ph.lpf()
            assert [v if v != 2 ** 32 - 1 else 2 ** 64 - 1 for v in got] == lpf[:len(got)]
            h, d = ph.prime_list()
            assert h == d == [p for p in primes if 41 <= p <= bound]
            # phi(p1...pk, k) = prod (p_i - 1): counted on the device up to 30030
            assert ph.makeModCache(10) == [1, 2, 8, 48, 480, 5760, 92160, 1658880, 36495360]
        finally:
            ph.close()


def test_pitable_pi_base_and_reallocate(ctx, oracle):
    import cudapix_amd as C
    base, rng = 10 ** 6, 100000
    t = C.PiTable(ctx, base, 0, rng)
    u = C.PiTable(ctx, base, 0, rng)
    try:
        b0, pb, vals = t.getCurrent()
        assert pb == oracle.count_primes(0, base)
        u.set_pi_base(pb + 5)                   # reported as given; the table stays relative
        _, pb2, vals2 = u.getCurrent()
        assert pb2 == pb + 5 and vals2 == vals
        _, pb3, _ = u.getNextUp()               # ... and moved by the window's primes
        assert pb3 == pb + 5 + oracle.count_primes(base, base + rng)
        t.reallocate(50000)
        _, _, half = t.getCurrent()
        assert half == vals[:len(half)]
        r = t.getMaxRange()
        assert r > 0 and r % 2 == 0
    finally:
        t.close()
        u.close()
