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
