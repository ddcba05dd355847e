"""The command line front-end (python -m cudapix_amd, the counterpart of the
reference's `pix x [y [c]]`, src/main.cpp:48-72): its integer-expression
parser on the CPU, and one solve and one batch on the GPU against published
values and the oracle's terms."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_parse_int_expressions():
    from cudapix_amd.__main__ import parse_int
    assert parse_int("1e16") == 10 ** 16
    assert parse_int("10^16") == 10 ** 16
    assert parse_int("2**64 + 1") == 2 ** 64 + 1
    assert parse_int("2^64-1") == 2 ** 64 - 1
    assert parse_int("2.5e15") == 25 * 10 ** 14
    assert parse_int("(3 * 7) << 4") == 336
    assert parse_int(" 37607912018 ") == 37607912018
    for bad in ("1.5", "1e-3", "__import__('os')", "x + 1", "2 ** 999", "1/2"):
        with pytest.raises((ValueError, SyntaxError)):
            parse_int(bad)


def run(args, stdin=None):
    p = subprocess.run([sys.executable, "-m", "cudapix_amd"] + args, cwd=ROOT, input=stdin,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return [json.loads(ln) for ln in p.stdout.splitlines() if ln.strip()]


@pytest.mark.gpu
def test_cli_single_and_batch(oracle):
    print("case cli 1e12 y=20000 c=6", flush=True)
    (r,) = run(["1e12", "20000", "6", "--terms"])
    assert r["pi"] == "37607912018" and r["y"] == 20000 and r["c"] == 6
    x, y = 10 ** 12, 20000
    assert int(r["p2"]) == oracle.P2(x, y)
    assert int(r["ordinary"]) == oracle.Ordinary(x, y, 6)
    assert int(r["trivial"]) == oracle.S2_trivial(x, y, 6)
    print("case cli batch", flush=True)
    out = run(["--batch", "-"], stdin="1e9\n10^10\n# comment\n2^40\n")
    assert [o["pi"] for o in out] == ["50847534", "455052511", str(oracle.count_primes(0, 2 ** 40 + 1))]
