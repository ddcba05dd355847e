"""GPU parity of every DR term against the CPU oracle (bit-exact integers).

Each term is computed by the HIP library through the C ABI and compared with
the oracle's restatement of the reference term on the same (x, y, c).
"""
import pytest

pytestmark = pytest.mark.gpu

KNOWN = {10 ** 4: 1229, 10 ** 5: 9592, 10 ** 6: 78498, 10 ** 7: 664579, 10 ** 8: 5761455,
         10 ** 9: 50847534, 10 ** 10: 455052511, 10 ** 11: 4118054813,
         10 ** 12: 37607912018, 10 ** 13: 346065536839}

XS = [16384, 16385, 20000, 65537, 10 ** 5, 123457, 10 ** 6, 999999, 10 ** 7,
      3 ** 15, 10 ** 8, 2 ** 30, 10 ** 9, 10 ** 10, 10 ** 11, 10 ** 12]


@pytest.mark.parametrize("x", XS)
@pytest.mark.parametrize("alpha", [1.0, 2.0, 4.0])
def test_terms_match_oracle(ctx, oracle, x, alpha):
    y = oracle.choose_y(x, alpha)
    print("case x=%d alpha=%g y=%d" % (x, alpha, y), flush=True)
    t = ctx.pi(x, y=y, c=6, terms=True)
    ref = oracle.pi_dr(x, y, 6)
    for k in ("ordinary", "trivial", "easy", "hard", "p2", "pi"):
        assert t[k] == ref[k], (k, t[k], ref[k])
    if x in KNOWN:
        assert t["pi"] == KNOWN[x]


@pytest.mark.parametrize("x", [10 ** 5, 10 ** 7, 10 ** 10])
@pytest.mark.parametrize("c", [1, 2, 3, 4, 5, 6, 7, 8])
def test_c_invariance(ctx, oracle, x, c):
    y = oracle.choose_y(x, 2.0)
    print("case x=%d c=%d" % (x, c), flush=True)
    t = ctx.pi(x, y=y, c=c, terms=True)
    cc = oracle.clamp_c(x, c)
    ref = oracle.pi_dr(x, y, cc)
    for k in ("ordinary", "trivial", "easy", "hard", "p2", "pi"):
        assert t[k] == ref[k], (k, t[k], ref[k])
    assert t["pi"] == KNOWN[x]


def test_single_term_entry_points(ctx, oracle):
    x = 10 ** 11
    y = oracle.choose_y(x, 2.0)
    assert ctx.P2(x, y) == oracle.P2(x, y)
    assert ctx.Ordinary(x, y, 6) == oracle.Ordinary(x, y, 6)
    assert ctx.S2_trivial(x, y) == oracle.S2_trivial(x, y)
    assert ctx.V(x, y) == oracle.V(x, y)
    assert ctx.S3(x, y, 6) == oracle.S3(x, y, 6)
    va, vb = ctx.V_split(x, y)
    assert va + vb == oracle.V(x, y)


def test_small_x_table_path(ctx, oracle):
    for x in [0, 1, 2, 3, 4, 10, 100, 1000, 1023, 4096, 10 ** 4, 16383]:
        assert ctx.pi(x) == oracle.count_primes(0, x + 1)


@pytest.mark.parametrize("x,alpha", [(10 ** 9, 1.0), (10 ** 10, 3.0), (10 ** 12, 2.0), (10 ** 12, 4.0),
                                     (2 ** 40 + 15, 6.0), (10 ** 13, 1.5), (10 ** 14, 8.0)])
def test_s2_sigma_form_matches_oracle(ctx, oracle, x, alpha):
    """The sigma form of the easy leaves (A and sigma1..sigma6 of the reference's
    S2.cu / sigma.cu) term by term against the oracle; where it is exact (no
    prime in (x^1/4, x/y^2]) its sum equals the direct V of the same y."""
    y = oracle.choose_y(x, alpha)
    print("case sigma x=%d y=%d" % (x, y), flush=True)
    g = ctx.s2_sigma(x, y)
    r = oracle.sigma(x, y)
    assert g["y"] == y
    for k in ("A", "s1", "s2", "s3", "s4", "s5", "s6", "S2"):
        assert g[k] == r[k], k
    v = ctx.V(x, y)
    assert (g["S2"] == v) == g["exact"], (g["S2"], v, g["exact"])


def _iroot(n, k):
    r = int(round(n ** (1.0 / k)))
    while r ** k > n:
        r -= 1
    while (r + 1) ** k <= n:
        r += 1
    return r


@pytest.mark.parametrize("x", [10 ** 18, 2 ** 64 + 1, 10 ** 20])
def test_s2_sigma_form_equals_v_beyond_the_oracle(ctx, x):
    """Past the oracle's reach: with y > x^(3/8), x/y^2 < x^(1/4), so no prime
    lies in (x^1/4, x/y^2] and the sigma form's sum must equal the direct V of
    the same y exactly -- two partitions of the easy leaves, each on the GPU
    (the 128-bit path at 2^64 + 1)."""
    y = _iroot(x ** 3, 8) + 1
    print("case sigma x=%d y=%d" % (x, y), flush=True)
    g = ctx.s2_sigma(x, y)
    assert g["y"] == y and g["exact"]
    assert g["S2"] == ctx.V(x, y)
