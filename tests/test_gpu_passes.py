"""Hard segments in passes over segment ranges (cpx_opts.pass_bytes): a
single query whose hard rows (T per active stage of every segment, sum(mu)
per replica and active stage) exceed the pass size is solved range by range,
each stage's survivor prefix carried from the earlier ranges.  Forced tiny
pass sizes must give every term of the one-pass solve; x = 1e22 (about
1.5 TB of rows in one pass) needs passes and gives the published pi(10^22)."""
import json
import os

import pytest

import cudapix_amd as C

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "known.json")) as f:
    KNOWN = json.load(f)

TERMS = ("pi", "ordinary", "trivial", "easy", "hard", "p2", "y", "z", "c", "pi_y")


@pytest.mark.parametrize("x,pass_bytes", [(10 ** 10, 4), (10 ** 12, 2048), (10 ** 14, 65536),
                                          (10 ** 16, 1 << 20), (2 ** 64 + 1, 1 << 28)])
def test_passes_equal_one_pass(ctx, x, pass_bytes):
    print("case x=%d pass_bytes=%d" % (x, pass_bytes), flush=True)
    (t1,), s1 = ctx.pi_batch([x], terms=True, stats=True)
    with C.Context(device=0, pass_bytes=pass_bytes) as cp:
        (tp,), sp = cp.pi_batch([x], terms=True, stats=True)
        # and a second solve on the same context (the pass state is reused)
        (tq,), _ = cp.pi_batch([x], terms=True, stats=True)
    assert s1["passes"] == 1 and sp["passes"] > 1, (s1["passes"], sp["passes"])
    for k in TERMS:
        assert tp[k] == t1[k] == tq[k], (k, tp[k], t1[k], tq[k])
    n = {10 ** 10: "10", 10 ** 12: "12", 10 ** 14: "14", 10 ** 16: "16"}.get(x)
    if n:
        assert tp["pi"] == KNOWN["pi10"][n]
