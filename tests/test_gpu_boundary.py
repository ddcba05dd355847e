"""GPU parity at the drop-in boundary: every 64-bit entry point (the exact
signatures of include/P2.cuh:16, ordinary.cuh:6, trivial.cuh:8, S3.cuh:3,
V.cuh:6) against the CPU oracle, the 128-bit ordinary and trivial terms at
x >= 2^64 term by term, V_a and V_b each against the oracle's split at
p = sqrt(x/y) (src/V.cu:58-124), and the reference-style C++ caller's device
branch run on the GPU.  Tolerance 0 everywhere (integer terms)."""
import json
import os
import shutil
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("x,alpha", [(10 ** 12, 2.0), (10 ** 12, 4.0), (987654321987, 1.0),
                                     (2 ** 40 + 1, 3.0)])
def test_u64_entry_points_match_oracle(ctx, oracle, x, alpha):
    y = oracle.choose_y(x, alpha)
    print("case u64 x=%d y=%d" % (x, y), flush=True)
    got = ctx.terms_u64(x, y, 6)
    assert got["p2"] == oracle.P2(x, y)
    assert got["ordinary"] == oracle.Ordinary(x, y, 6)
    assert got["trivial"] == oracle.S2_trivial(x, y, 6)
    assert got["easy"] == oracle.V(x, y, 6)
    assert got["hard"] == oracle.S3(x, y, 6)
    ref = oracle.pi_dr(x, y, 6)
    assert got["pi"] == ref["pi"]


@pytest.mark.parametrize("x,alpha", [(10 ** 20, 2.0), (10 ** 20, 4.0), (2 ** 64, 2.0),
                                     (2 ** 64 + 1, 2.0), (2 ** 64 - 1, 3.0), (3 * 10 ** 20, 2.0)])
def test_u128_ordinary_and_trivial_match_oracle(ctx, oracle, x, alpha):
    # S0 by the two-digit 128/64 divide per ordinary leaf, S1+U with x/p^2
    # from a 128-bit x, each against the oracle's unsigned __int128 restatement
    y = oracle.choose_y(x, alpha)
    print("case u128 x=%d y=%d" % (x, y), flush=True)
    assert ctx.Ordinary(x, y, 6) == oracle.Ordinary(x, y, 6)
    assert ctx.S2_trivial(x, y) == oracle.S2_trivial(x, y, 6)


@pytest.mark.parametrize("x", [10 ** 8, 123456789, 10 ** 10, 10 ** 11, 10 ** 12, 3 * 10 ** 12])
@pytest.mark.parametrize("alpha", [1.0, 2.0, 4.0])
def test_v_a_and_v_b_match_oracle_split(ctx, oracle, x, alpha):
    y = oracle.choose_y(x, alpha)
    print("case V split x=%d y=%d" % (x, y), flush=True)
    va, vb = ctx.V_split(x, y)
    assert va == oracle.V_a(x, y)
    assert vb == oracle.V_b(x, y)
    assert ctx.V_a(x, y) + ctx.V_b(x, y) == ctx.V(x, y)


def test_v_split_golden(ctx):
    # the split against the brute-force leaf classification itself
    with open(os.path.join(GOLD, "terms.json")) as f:
        cases = [t for t in json.load(f) if t["c"] == 6 and t["x"] >= 16384]
    for t in cases:
        assert ctx.V_split(t["x"], t["y"]) == (t["easy_a"], t["easy_b"]), t


def test_reference_style_caller_device_branch(tmp_path, oracle):
    """The C++ caller of include/cudapix.hpp (reference call sites unchanged)
    with its device branch: CUDASieve calls, the 64-bit DR assembly at 1e12,
    Phi and PiTable."""
    from test_dropin_header import build_caller
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    exe = build_caller(tmp_path)
    print("case drop-in caller device branch", flush=True)
    out = subprocess.run([str(exe), "device"], check=True, capture_output=True, text=True,
                         timeout=300).stdout.split()
    vals = [int(v) for v in out]
    pi_y = oracle.count_primes(0, 20001)
    assert vals[0] == 5 * 9592                          # pi(1e5) five ways
    assert vals[1] + pi_y == 37607912018                # dr64(1e12, 20000, 6) + pi(y)
    assert vals[2] == oracle.phi_legendre(1000, 3)
    # V through the reference's V() body (PrimeArray V_a + V_b) and V itself
    assert vals[3] == vals[4] == oracle.V(10 ** 12, 20000, 6)
    # Phi(1e5, 1000): primes 41..1000, the LPF array, phi of the primorials
    assert vals[5:9] == [168 - 12, 41, 1, 5760]
    # gen_h_lpf / gen_h_mu <uint64_t> at n = 2^40 + 1 = 257 * 4278255361
    assert vals[9:11] == [257, 1]
    # PiTable::set_pi_base is carried, getMaxRange is positive
    assert vals[11:13] == [78498, 1]
    assert vals[-1] == 0                                # host self check
