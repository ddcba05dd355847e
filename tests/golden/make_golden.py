"""Generate the committed golden fixtures under tests/golden/.

Independent of the C oracle: every value here comes from a numpy brute force
written from the definitions (no reference code is imported or run):

* terms.json -- for small x and several alpha, every DR special leaf
    (p, m): p_c < p <= y, lpf(m) > p, m <= y < m p, mu(m) != 0
  enumerated over ALL m (not only primes), with phi(n, b) read from cumulative
  sums of "no prime factor among the first b primes" indicator arrays, then
  classified by p into S1 (p > cbrt x), U (x/(p m) < p), V (x^(1/4) < p) and
  S3 (p <= x^(1/4)); V is also split at p = isqrt(x // y) into V_a (p at or
  below) and V_b (above); S0 from its definition with the same phi; P2 by
  counting n <= x whose two prime factors (with multiplicity) both exceed y.
* known.json -- published values: pi(10^n) (OEIS A006880), pi(2^k)
  (OEIS A007053), Mertens M(10^n) (OEIS A084237).
* queries_c3.json / queries_c4.json -- the batched and served workloads
  (splitmix64 seeds 3 and 4), written once so every consumer sees the same x.

Run:  python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SMALLP = [2, 3, 5, 7, 11, 13, 17, 19]


def iroot(x, k):
    r = int(round(x ** (1.0 / k)))
    while r ** k > x:
        r -= 1
    while (r + 1) ** k <= x:
        r += 1
    return r


def spf_sieve(n):
    """Smallest prime factor of 0..n (spf[1] = 0)."""
    spf = np.zeros(n + 1, dtype=np.int64)
    for p in range(2, n + 1):
        if spf[p] == 0:
            spf[p::p][spf[p::p] == 0] = p
    return spf


def factor_info(n):
    """mu, lpf (lpf(1) = inf), Omega for 1..n from a smallest-prime-factor sieve."""
    spf = spf_sieve(n)
    mu = np.zeros(n + 1, dtype=np.int64)
    omega = np.zeros(n + 1, dtype=np.int64)
    mu[1] = 1
    for k in range(2, n + 1):
        p = spf[k]
        r = k // p
        omega[k] = omega[r] + 1
        mu[k] = 0 if r % p == 0 else -mu[r]
    lpf = spf.copy()
    lpf[1] = 1 << 62
    return mu, lpf, omega, spf


def leaves_terms(x, alpha, c):
    cb, sq, q4 = iroot(x, 3), iroot(x, 2), iroot(x, 4)
    y = int(alpha * cb)
    y = max(cb, min(y, sq))
    kc = sum(1 for p in SMALLP if p <= q4)
    c = min(c, kc)
    z = x // y
    N = max(z, y) + 1
    mu, lpf, omega, spf = factor_info(N)
    primes = [int(p) for p in range(2, N + 1) if spf[p] == p]
    a = sum(1 for p in primes if p <= y)
    # phi tables: cum[b][n] = #{1 <= k <= n : no prime factor among first b primes}
    bmax = a
    ks = np.arange(N + 1)
    alive = ks >= 1
    cums = {}
    need_b = set(range(0, min(bmax, len(primes)) + 1))
    for b in range(0, bmax + 1):
        if b > 0:
            p = primes[b - 1]
            alive = alive & (ks % p != 0)
        if b in need_b:
            cums[b] = np.cumsum(alive.astype(np.int64))

    def phi(n, b):
        assert n <= N, (n, N)
        return int(cums[b][n])

    # S0, with phi(u, c) = sum over squarefree d | W_c of mu(d) floor(u / d)
    # (Legendre's inclusion-exclusion; the argument x/n is far above the tables)
    divs = [(1, 1)]
    for p in SMALLP[:c]:
        divs += [(d * p, -s) for d, s in divs]

    def phi_c(u):
        return sum(s * (u // d) for d, s in divs)

    s0 = 0
    pc = SMALLP[c - 1]
    for n in range(1, y + 1):
        if mu[n] != 0 and lpf[n] > pc:
            s0 += int(mu[n]) * phi_c(x // n)
    # special leaves
    S1 = U = V = S3 = Va = 0
    sqz = iroot(z, 2)
    for bi, p in enumerate(primes):
        if p <= pc or p > y:
            continue
        b = bi + 1                      # p = p_b
        for m in range(y // p + 1, y + 1):
            if mu[m] == 0 or lpf[m] <= p:
                continue
            n = x // (m * p)
            val = -int(mu[m]) * phi(n, b - 1)
            if p > cb:
                assert n < p and val == 1
                S1 += val
            elif p <= q4:
                S3 += val
            elif n < p:
                assert val == 1
                U += val
            else:
                assert mu[m] == -1          # m prime for every easy leaf
                V += val
                if p <= sqz:
                    Va += val
    # P2 by brute count
    P2 = 0
    parr = np.array(primes, dtype=np.int64)
    for p in primes:
        if p <= y:
            continue
        if p * p > x:
            break
        # primes q with p <= q <= x/p   (x/p < z <= N)
        P2 += int(np.searchsorted(parr, x // p, side="right") - np.searchsorted(parr, p))
    phi_xa = s0 + S1 + U + V + S3
    pi = phi_xa + a - 1 - P2
    return {"x": x, "alpha": alpha, "y": y, "c": c, "ordinary": s0, "trivial": S1 + U,
            "easy": V, "easy_a": Va, "easy_b": V - Va, "hard": S3, "p2": P2, "pi_y": a, "pi": pi}


def p2_brute(x, y):
    """#{n <= x : n = p q, y < p <= q primes} by Omega / lpf counting."""
    mu, lpf, omega, spf = factor_info(x)
    n = np.arange(x + 1)
    sel = (omega == 2) & (spf > y)
    return int(np.count_nonzero(sel[2:]))


def splitmix64(seed):
    s = seed & ((1 << 64) - 1)
    while True:
        s = (s + 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        yield z ^ (z >> 31)


def u53(g):
    return (next(g) >> 11) / float(1 << 53)


def main():
    known = {
        "pi10": {str(n): v for n, v in enumerate([
            0, 4, 25, 168, 1229, 9592, 78498, 664579, 5761455, 50847534, 455052511,
            4118054813, 37607912018, 346065536839, 3204941750802, 29844570422669,
            279238341033925, 2623557157654233, 24739954287740860, 234057667276344607,
            2220819602560918840, 21127269486018731928, 201467286689315906290])},
        "pi2": {"10": 172, "20": 82025, "30": 54400028, "32": 203280221, "40": 41203088796,
                "64": 425656284035217743},
        "mertens10": {"1": -1, "2": 1, "3": 2, "4": -23, "5": -48, "6": 212, "7": 1037,
                      "8": 1928, "9": -222},
        "source": "pi(10^n): OEIS A006880; pi(2^k): OEIS A007053; M(10^n): OEIS A084237",
    }
    del known["pi10"]["0"]
    with open(os.path.join(HERE, "known.json"), "w") as f:
        json.dump(known, f, indent=1)

    terms = []
    for x in [20000, 65537, 10 ** 5, 123457, 10 ** 6]:
        for alpha in (1.0, 2.0, 4.0):
            for c in (6, 3):
                t = leaves_terms(x, alpha, c)
                terms.append(t)
                print(t, flush=True)
    with open(os.path.join(HERE, "terms.json"), "w") as f:
        json.dump(terms, f, indent=1)

    p2 = []
    for x, y in [(10 ** 5, 46), (10 ** 6, 200), (2 * 10 ** 6, 300), (4 * 10 ** 6, 500)]:
        p2.append({"x": x, "y": y, "p2": p2_brute(x, y)})
    with open(os.path.join(HERE, "p2_brute.json"), "w") as f:
        json.dump(p2, f, indent=1)

    g = splitmix64(3)
    c3 = [int(10 ** (9 + 6 * u53(g))) for _ in range(512)]
    with open(os.path.join(HERE, "queries_c3.json"), "w") as f:
        json.dump({"seed": 3, "rule": "x = floor(10^(9 + 6u)), u = top 53 bits of splitmix64 / 2^53",
                   "x": c3}, f)

    g = splitmix64(4)
    c4 = {"seed": 4, "rule": "gap = -ln(1-u)/lambda s; x = floor(10^(8 + 9u')), u, u' from "
                             "splitmix64 in that order", "streams": {}}
    for lam in (50, 200, 500, 1000, 2000):
        t, ts, xs = 0.0, [], []
        for _ in range(2000):
            t += -math.log(1.0 - u53(g)) / lam
            ts.append(round(t * 1000.0, 6))
            xs.append(int(10 ** (8 + 9 * u53(g))))
        c4["streams"][str(lam)] = {"t_ms": ts, "x": xs}
    with open(os.path.join(HERE, "queries_c4.json"), "w") as f:
        json.dump(c4, f)


if __name__ == "__main__":
    main()
