"""CPU tests: the oracle pinned by fixtures, published values and mathematics.

The oracle (oracle/oracle.c) restates the reference's terms; here it is checked
against (a) golden vectors from an independent numpy brute force that
enumerates every special leaf from the definition (tests/golden/terms.json),
(b) published pi(10^n), pi(2^k) and Mertens values (tests/golden/known.json),
(c) brute semiprime counts for P2, (d) sympy factorisations for mu / lpf, and
(e) invariance of pi(x) under different alpha and c.
"""
import json
import os
import random

import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


@pytest.mark.parametrize("case", load("terms.json"), ids=lambda t: "x%d-a%g-c%d" % (t["x"], t["alpha"], t["c"]))
def test_oracle_terms_match_brute_force_leaves(oracle, case):
    r = oracle.pi_dr(case["x"], case["y"], case["c"])
    for k in ("ordinary", "trivial", "easy", "hard", "p2", "pi_y", "pi"):
        assert r[k] == case[k], k


@pytest.mark.parametrize("case", load("terms.json"), ids=lambda t: "x%d-a%g-c%d" % (t["x"], t["alpha"], t["c"]))
def test_oracle_v_split_matches_brute_force_leaves(oracle, case):
    """V_a (x^(1/4) < p <= sqrt(x/y)) and V_b (sqrt(x/y) < p <= cbrt x), each
    against the brute-force leaf classification on its own."""
    assert oracle.V_a(case["x"], case["y"], case["c"]) == case["easy_a"]
    assert oracle.V_b(case["x"], case["y"], case["c"]) == case["easy_b"]


@pytest.mark.parametrize("case", load("p2_brute.json"), ids=lambda t: "x%d-y%d" % (t["x"], t["y"]))
def test_oracle_p2_matches_semiprime_count(oracle, case):
    assert oracle.P2(case["x"], case["y"]) == case["p2"]


def test_known_pi10_small_by_brute_sieve(oracle):
    k = load("known.json")["pi10"]
    for n in range(1, 10):
        assert oracle.count_primes(0, 10 ** n + 1) == k[str(n)]


def test_known_pi2_small_by_brute_sieve(oracle):
    k = load("known.json")["pi2"]
    for e in ("10", "20", "30", "32"):
        assert oracle.count_primes(0, 2 ** int(e) + 1) == k[e]


@pytest.mark.parametrize("n", [9, 10, 11, 12, 13])
def test_oracle_dr_matches_published_pi(oracle, n):
    x = 10 ** n
    k = load("known.json")["pi10"]
    for alpha in (1.0, 3.0):
        y = oracle.choose_y(x, alpha)
        assert oracle.pi_dr(x, y, 6)["pi"] == k[str(n)]


def test_mertens(oracle):
    k = load("known.json")["mertens10"]
    for n in range(1, 8):
        assert oracle.mertens(10 ** n) == k[str(n)]


def test_mu_lpf_against_sympy(oracle):
    sympy = pytest.importorskip("sympy")
    mu, lpf = oracle.mu_lpf_odd(20001)
    for i, n in enumerate(range(1, 20002, 2)):
        if n == 1:
            assert mu[i] == 1 and lpf[i] == 2 ** 64 - 1
            continue
        f = sympy.factorint(n)
        assert lpf[i] == min(f)
        assert mu[i] == (0 if max(f.values()) > 1 else (-1) ** len(f))


def test_mu_lpf_window_oracle(oracle):
    # the trial-division windows against the full sieve below 4e5, and against
    # sympy's factorisations for 64-bit windows (2^33, 2^44, 2^50)
    sympy = pytest.importorskip("sympy")
    mu, lpf = oracle.mu_lpf_odd(400001)
    for lo, hi in ((0, 20000), (300000, 320002)):
        wm, wl = oracle.mu_lpf_window(lo, hi)
        assert wm == mu[lo // 2: hi // 2] and wl == lpf[lo // 2: hi // 2]
    for lo in (2 ** 33, 2 ** 44 + 1000, 2 ** 50):
        wm, wl = oracle.mu_lpf_window(lo, lo + 400)
        for i in range(0, 200, 7):
            f = sympy.factorint(lo + 2 * i + 1)
            assert wl[i] == min(f)
            assert wm[i] == (0 if max(f.values()) > 1 else (-1) ** len(f))


def test_phi_legendre_equals_definition(oracle):
    rnd = random.Random(5)
    for _ in range(40):
        n = rnd.randrange(0, 3000)
        b = rnd.randrange(0, 9)
        assert oracle.phi_legendre(n, b) == oracle.phi_brute(n, b)


def test_identity_phi_equals_leaves(oracle):
    # phi(x, pi(y)) = S0 + S1 + U + V + S3, checked against Legendre's phi
    for x in (20000, 50000, 10 ** 5):
        y = oracle.choose_y(x, 2.0)
        r = oracle.pi_dr(x, y, 6)
        phi = r["ordinary"] + r["trivial"] + r["easy"] + r["hard"]
        assert phi == oracle.phi_legendre(x, r["pi_y"])


@pytest.mark.parametrize("x", [10 ** 7 + 19, 3 ** 17, 2 ** 34 + 1, 10 ** 11 + 7])
def test_alpha_and_c_invariance(oracle, x):
    vals = set()
    for alpha in (1.0, 1.7, 3.0, 5.5):
        for c in (1, 4, 6, 8):
            y = oracle.choose_y(x, alpha)
            vals.add(oracle.pi_dr(x, y, c)["pi"])
    assert len(vals) == 1
    if x < 10 ** 9:
        assert vals.pop() == oracle.count_primes(0, x + 1)


def test_integer_roots_exact(oracle):
    rnd = random.Random(11)
    for _ in range(2000):
        x = rnd.getrandbits(rnd.randrange(1, 101))
        s, c, q = oracle.isqrt(x), oracle.icbrt(x), oracle.iroot4(x)
        assert s * s <= x < (s + 1) ** 2
        assert c ** 3 <= x < (c + 1) ** 3
        assert q ** 4 <= x < (q + 1) ** 4


def test_query_fixtures_regenerate():
    # the committed workloads are what the generator produces
    import importlib.util
    spec = importlib.util.spec_from_file_location("mg", os.path.join(GOLD, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = mg.splitmix64(3)
    xs = [int(10 ** (9 + 6 * mg.u53(g))) for _ in range(512)]
    assert xs == load("queries_c3.json")["x"]
    assert all(10 ** 9 <= x <= 10 ** 15 for x in xs)
    c4 = load("queries_c4.json")["streams"]
    assert sorted(int(k) for k in c4) == [50, 200, 500, 1000, 2000]
    for s in c4.values():
        assert len(s["x"]) == 2000 and all(10 ** 8 <= x <= 10 ** 17 for x in s["x"])
        assert all(a < b for a, b in zip(s["t_ms"], s["t_ms"][1:]))


def _sigma_brute(x, y):
    """A and sigma2..sigma6 from their definitions with a numpy sieve (for the
    oracle's orc_sigma), and the direct V."""
    import numpy as np
    sq = int(x ** 0.5) + 2
    s = np.ones(sq + 1, bool)
    s[:2] = False
    for i in range(2, int(sq ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    pc = np.cumsum(s)
    pr = [int(v) for v in np.nonzero(s)[0]]

    def root(v, k):
        r = int(round(v ** (1.0 / k)))
        while r ** k > v:
            r -= 1
        while (r + 1) ** k <= v:
            r += 1
        return r

    def pi(n):
        return int(pc[n])

    q4, c3, sz = root(x, 4), root(x, 3), root(x // y, 2)
    P1 = [p for p in pr if q4 < p <= c3]
    A = sum((2 if x // (p * q) < y else 1) * pi(x // (p * q))
            for p in P1 for q in pr if p < q <= root(x // p, 2))
    V = sum(pi(x // (p * q)) - pi(p) + 2 for p in P1 for q in pr if p < q <= min(y, x // (p * p)))
    a, b, c, d = pi(y), pi(c3), pi(q4), pi(sz)
    s2 = a * (c * (c - 3) // 2 - d * (d - 3) // 2)
    s3 = b * (b - 1) * (2 * b - 1) // 6 - b - c * (c - 1) * (2 * c - 1) // 6 + c
    s4 = a * sum(pi(x // (p * y)) for p in P1 if p <= sz)
    s5 = sum(pi(min(y, x // (p * p))) for p in P1 if p > sz)
    s6 = -sum(pi(root(x // p, 2)) ** 2 for p in P1)
    exact = not any(p * y * y <= x for p in P1)
    return {"A": A, "s2": s2, "s3": s3, "s4": s4, "s5": s5, "s6": s6}, V, exact


def test_sigma_partition_against_definitions(oracle):
    """orc_sigma against the definitions (numpy sieve), and the identity
    A + s2 + s3 + s4 + 2 s5 + s6 = V exactly when no prime lies in
    (x^1/4, x/y^2] -- and not otherwise (A then also counts q > y)."""
    rnd = random.Random(21)
    seen = {True: 0, False: 0}
    for _ in range(24):
        x = rnd.randrange(10 ** 6, 3 * 10 ** 8)
        cb = oracle.icbrt(x)
        y = rnd.randrange(cb, min(oracle.isqrt(x), 5 * cb) + 1)
        want, V, exact = _sigma_brute(x, y)
        got = oracle.sigma(x, y)
        for k, v in want.items():
            assert got[k] == v, (x, y, k)
        assert (got["S2"] == V) == exact, (x, y)
        seen[exact] += 1
    assert seen[True] and seen[False]


def test_sigma_partition_equals_golden_easy_leaves(oracle):
    # where the partition is exact, it gives the brute-force easy leaves
    n = 0
    for case in load("terms.json"):
        x, y = case["x"], case["y"]
        if x < 16384:
            continue
        _, _, exact = _sigma_brute(x, y)
        if exact:
            assert oracle.sigma(x, y)["S2"] == case["easy"], (x, y)
            n += 1
    assert n >= 3
