"""Large x at full size: published pi(10^n) and pi(2^64), and agreement of
the total across alpha tunings (the u128 path for x > 2^64)."""
import json
import os

import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "known.json")) as f:
    KNOWN = json.load(f)


@pytest.mark.parametrize("n", [13, 14, 15, 16])
def test_pi_10n(ctx, n):
    print("case pi(1e%d)" % n, flush=True)
    assert ctx.pi(10 ** n) == KNOWN["pi10"][str(n)]


def test_pi_1e16_two_alphas(ctx):
    x = 10 ** 16
    vals = set()
    for alpha in (2.0, 5.0):
        y = int(alpha * 215443)
        t = ctx.pi(x, y=y, terms=True)
        vals.add(t["pi"])
    assert vals == {KNOWN["pi10"]["16"]}


@pytest.mark.slow
def test_pi_1e17(ctx):
    assert ctx.pi(10 ** 17) == KNOWN["pi10"]["17"]


@pytest.mark.slow
def test_pi_2_64_u128(ctx):
    # 2^64 needs the 128-bit path (x does not fit 64 bits)
    assert ctx.pi(2 ** 64) == KNOWN["pi2"]["64"]


@pytest.mark.slow
def test_pi_1e20_two_alphas(ctx):
    # 128-bit x with 64-bit primes and u128 leaf quotients; the same total for
    # alpha = 2 and 4 (and the built-in alpha), each the published pi(10^20)
    x = 10 ** 20
    cb = 4641588                                   # icbrt(10^20)
    for alpha in (2.0, 4.0, 0.0):
        y = int(alpha * cb)
        print("case pi(1e20) alpha=%g y=%d" % (alpha, y), flush=True)
        t = ctx.pi(x, y=y, terms=True)
        assert t["pi"] == KNOWN["pi10"]["20"], (alpha, t)


def test_batch_known_values(ctx):
    # one batch (query groups, P2 by table lookups) of published values
    xs = [10 ** n for n in range(9, 16)] + [2 ** k for k in (30, 32, 40)]
    want = [KNOWN["pi10"][str(n)] for n in range(9, 16)] + [KNOWN["pi2"][str(k)] for k in (30, 32, 40)]
    assert ctx.pi_batch(xs) == want


@pytest.mark.slow
def test_batch_with_u128_member(ctx):
    # a 128-bit x (P2 from its sieve segments, a group of its own) in a batch
    # whose other members take P2 from table lookups
    xs = [2 ** 64] + [10 ** n for n in range(9, 17)]
    want = [KNOWN["pi2"]["64"]] + [KNOWN["pi10"][str(n)] for n in range(9, 17)]
    assert ctx.pi_batch(xs) == want


def test_pi_1e21_wide_stage_quotients(ctx):
    """x = 1e21: x / p exceeds 2^64 for the first hard stages (p < 55), so the
    segment tiles carry its high bits; the published pi(10^21)."""
    print("case pi(1e21)", flush=True)
    t = ctx.pi(10 ** 21, terms=True)
    print("pi(1e21) y=%d in %.0f ms" % (t["y"], t["ms"]), flush=True)
    assert t["pi"] == KNOWN["pi10"]["21"]
