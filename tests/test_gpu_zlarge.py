"""x = 1e22, the reference README's range: its hard rows (T per active stage
of every segment, sum(mu) per replica and active stage; about 1.5 TB in one
pass) run in passes over segment ranges of at most cpx_opts.pass_bytes
(48 GiB by default), each stage's survivor prefix carried between passes.
On a context of its own, closed afterwards (its tables reach 1e11 and its
arena holds one pass of rows).  The library prints a line per pass to stderr
when CPX_PROGRESS is set."""
import json
import os

import pytest

import cudapix_amd as C

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLD, "known.json")) as f:
    KNOWN = json.load(f)


def test_pi_1e22_in_passes():
    os.environ["CPX_PROGRESS"] = "1"
    try:
        print("case pi(1e22)", flush=True)
        with C.Context(device=0) as c:
            (t,), st = c.pi_batch([10 ** 22], terms=True, stats=True)
    finally:
        del os.environ["CPX_PROGRESS"]
    print("pi(1e22) y=%d passes=%d in %.0f ms" % (t["y"], st["passes"], st["ms_total"]), flush=True)
    assert st["passes"] > 1
    assert t["pi"] == KNOWN["pi10"]["22"]
