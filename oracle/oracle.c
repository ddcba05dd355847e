/*
 * oracle.c -- CPU restatement of the Deleglise-Rivat combinatorial pi(x) terms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links or calls this
 * file.  It is used by tests/ and by __graft_entry__.smoke() (as the values the
 * GPU results are compared with) and by bench.py's cpu_baseline leg (timed
 * beside the GPU).
 *
 * Every term is restated from the reference's algorithm (file:line below,
 * relative to the reference tree) with its known defects fixed, and with every
 * bound computed by exact integer roots instead of floating point:
 *
 *   P2         src/P2.cu:76-133           sum_{y<p<=sqrt x} [pi(x/p) - pi(p) + 1]
 *   Ordinary   src/ordinary.cu:23-57,101-115  sum_{odd n<=y, lpf(n)>p_c} mu(n) phi(x/n, c)
 *   S2_trivial src/trivial.cu:31-62       C(pi(y)-pi(cbrt x), 2) + sum_{sqrt z<p<=cbrt x} [pi(y)-pi(x/p^2)]
 *   V (easy)   src/V.cu:22-124            sum_{x^1/4<p<=cbrt x} sum_{p<q<=min(y,x/p^2)} [pi(x/pq)-pi(p)+2]
 *   V_a / V_b  src/V.cu:58-87 / :89-124   V split at p = sqrt(x/y) (orc_easy_part)
 *   S3 (hard)  src/sieve/S2_hard_device.cu:264-355 + S2_hard_host.cu:115-191 (DR split p <= x^1/4,
 *              src/Deleglise-Rivat/S3.cu:8): -sum_{p_c<p<=x^1/4} sum_{y/p<m<=y, lpf(m)>p} mu(m) phi(x/(mp), pi(p)-1)
 *   assembly   src/Deleglise-Rivat/deleglise-rivat.cpp:77
 *              pi(x) = S0 + trivial + V + S3 + pi(y) - 1 - P2
 *   mu / lpf   src/sieve/lpf_mu.cu:25-197 (odd-only arrays, lpf(1) = "infinity")
 *   pi table   src/pitable.cu:172-218     (pi(n) for every n below the table top)
 *
 * The CUDA reference cannot be built here (nvcc, CUDASieve and CUDA-uint128 are
 * absent), so these restatements are pinned by mathematics in tests/: published
 * pi(10^n), Mertens values, brute-force leaf enumeration with a naive phi on
 * tiny x, and invariance of the total under different alpha/c tunings.
 *
 * Integers: x is an unsigned 128-bit value; every signed sum is an __int128.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

static inline u128 mk128(uint64_t hi, uint64_t lo) { return ((u128)hi << 64) | lo; }
static inline void put128(u128 v, uint64_t out[2]) { out[0] = (uint64_t)v; out[1] = (uint64_t)(v >> 64); }

/* ------------------------------------------------------------------ roots */

uint64_t orc_isqrt(uint64_t xhi, uint64_t xlo)
{
  u128 x = mk128(xhi, xlo);
  if (x == 0) return 0;
  long double d = sqrtl((long double)x);
  uint64_t r = d >= 18446744073709551615.0L ? UINT64_MAX : (uint64_t)d;
  while ((u128)r * r > x) r--;
  while (r < UINT64_MAX && (u128)(r + 1) * (r + 1) <= x) r++;
  return r;
}

uint64_t orc_icbrt(uint64_t xhi, uint64_t xlo)
{
  u128 x = mk128(xhi, xlo);
  if (x == 0) return 0;
  uint64_t r = (uint64_t)cbrtl((long double)x);
  while (r > 0 && (u128)r * r * r > x) r--;
  while ((u128)(r + 1) * (r + 1) * (r + 1) <= x) r++;
  return r;
}

uint64_t orc_iroot4(uint64_t xhi, uint64_t xlo)
{
  /* floor(sqrt(floor(sqrt(x)))) == floor(x^(1/4)) */
  return orc_isqrt(0, orc_isqrt(xhi, xlo));
}

/* ------------------------------------------------------- small sieves */

/* primes <= limit (limit < 2^32), returned as a malloc'd array; *count set. */
static uint32_t *primes_upto(uint64_t limit, size_t *count)
{
  if (limit < 2) { *count = 0; return (uint32_t *)malloc(sizeof(uint32_t)); }
  uint8_t *comp = (uint8_t *)calloc(limit + 1, 1);
  for (uint64_t i = 2; i * i <= limit; i++)
    if (!comp[i])
      for (uint64_t j = i * i; j <= limit; j += i) comp[j] = 1;
  size_t n = 0;
  for (uint64_t i = 2; i <= limit; i++) n += !comp[i];
  uint32_t *p = (uint32_t *)malloc((n + 1) * sizeof(uint32_t));
  n = 0;
  for (uint64_t i = 2; i <= limit; i++) if (!comp[i]) p[n++] = (uint32_t)i;
  free(comp);
  *count = n;
  return p;
}

/* number of primes <= v in a sorted list */
static uint64_t count_le(const uint32_t *pr, size_t n, uint64_t v)
{
  size_t lo = 0, hi = n;
  while (lo < hi) { size_t mid = (lo + hi) / 2; if (pr[mid] <= v) lo = mid + 1; else hi = mid; }
  return lo;
}

size_t orc_primes(uint64_t limit, uint32_t *out, size_t cap)
{
  size_t n;
  uint32_t *p = primes_upto(limit, &n);
  size_t m = n < cap ? n : cap;
  if (out) memcpy(out, p, m * sizeof(uint32_t));
  free(p);
  return n;
}

/* Count primes in [lo, hi) with a segmented odd-only sieve of Eratosthenes
 * (the "plain segmented-Eratosthenes count" oracle).  OpenMP over segments. */
uint64_t orc_count_primes(uint64_t lo, uint64_t hi, int threads)
{
  if (hi <= lo) return 0;
  uint64_t total = 0;
  if (lo <= 2 && hi > 2) total = 1;                  /* the prime 2 */
  uint64_t r = orc_isqrt(0, hi - 1);
  size_t np; uint32_t *pr = primes_upto(r, &np);
  const uint64_t SEG = 1ull << 21;                   /* numbers per segment (even) */
  uint64_t base = lo & ~1ull;
  uint64_t nseg = (hi - base + SEG - 1) / SEG;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  uint64_t acc = 0;
#pragma omp parallel reduction(+:acc)
  {
    uint8_t *s = (uint8_t *)malloc(SEG / 2);
#pragma omp for schedule(dynamic, 4)
    for (uint64_t k = 0; k < nseg; k++) {
      uint64_t s0 = base + k * SEG;                 /* even; bit i <-> s0 + 2i + 1 */
      uint64_t s1 = s0 + SEG; if (s1 > hi) s1 = hi;
      uint64_t nb = (s1 - s0) / 2;                  /* odd numbers s0+1 .. <= s1-1 */
      memset(s, 1, nb);
      for (size_t i = 1; i < np; i++) {
        uint64_t p = pr[i];
        if (p * p >= s1) break;
        uint64_t st = p * p;
        if (st < s0 + 1) { st = ((s0 + 1 + p - 1) / p) * p; if (!(st & 1)) st += p; }
        for (uint64_t m = st; m < s1; m += 2 * p) s[(m - s0 - 1) / 2] = 0;
      }
      uint64_t c = 0;
      for (uint64_t i = 0; i < nb; i++) {
        uint64_t n = s0 + 2 * i + 1;
        if (s[i] && n >= lo && n > 1) c++;
      }
      acc += c;
    }
    free(s);
  }
  free(pr);
  return total + acc;
}

/* Brute mu / lpf for 1..n (index by n).  lpf(1) = UINT64_MAX as in
 * src/sieve/lpf_mu.cu:54 (d_lpf[0] = -1). */
static void mu_lpf_upto(uint64_t n, int8_t **mu_out, uint64_t **lpf_out)
{
  int8_t *mu = (int8_t *)malloc(n + 1);
  uint64_t *lpf = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i <= n; i++) mu[i] = 1;
  for (uint64_t p = 2; p <= n; p++) {
    if (lpf[p]) continue;                            /* composite */
    for (uint64_t m = p; m <= n; m += p) { if (!lpf[m]) lpf[m] = p; mu[m] = (int8_t)-mu[m]; }
    if (p <= n / p)
      for (uint64_t m = p * p; m <= n; m += p * p) mu[m] = 0;
  }
  lpf[1] = UINT64_MAX;
  mu[0] = 0;
  *mu_out = mu; *lpf_out = lpf;
}

void orc_mu_lpf_odd(uint64_t y, int8_t *mu, uint64_t *lpf)
{
  /* odd-only layout of src/sieve/lpf_mu.cu: entry i <-> n = 2i + 1, n <= y */
  int8_t *m; uint64_t *l;
  mu_lpf_upto(y, &m, &l);
  for (uint64_t n = 1; n <= y; n += 2) { mu[n / 2] = m[n]; lpf[n / 2] = l[n]; }
  free(m); free(l);
}

/* mu / lpf of the odd n = bottom + 2i + 1, i < count (bottom even, the window
 * below 2^63), each by trial division of its own cofactor by the primes up to
 * its square root -- no sieve, so it checks the sieved windows of any size
 * independently (small windows only: ~sqrt(n)/ln sqrt(n) divisions per prime
 * n).  lpf(1) = UINT64_MAX.  (The 64-bit windows of src/sieve/lpf_mu.cu:57-87,
 * 153-197.) */
void orc_mu_lpf_window(uint64_t bottom, uint64_t count, int8_t *mu, uint64_t *lpf)
{
  const uint64_t top = bottom + 2 * count;
  uint64_t r = 1;
  while ((r + 1) <= top / (r + 1)) r++;          /* isqrt(top) */
  size_t np;
  uint32_t *pr = primes_upto(r + 1, &np);
  for (uint64_t i = 0; i < count; i++) {
    uint64_t n = bottom + 2 * i + 1, m = n, lp = 0;
    int par = 0, sq = 0;
    for (size_t k = 1; k < np && (uint64_t)pr[k] <= m / pr[k]; k++) {   /* odd primes, p^2 <= m */
      const uint64_t p = pr[k];
      if (m % p) continue;
      if (!lp) lp = p;
      m /= p;
      par ^= 1;
      if (m % p == 0) { sq = 1; while (m % p == 0) m /= p; }
    }
    if (m > 1) { par ^= 1; if (!lp) lp = m; }
    lpf[i] = n == 1 ? UINT64_MAX : lp;
    mu[i] = (int8_t)(sq ? 0 : (par ? -1 : 1));
  }
  free(pr);
}

int64_t orc_mertens(uint64_t n)
{
  int8_t *m; uint64_t *l;
  mu_lpf_upto(n, &m, &l);
  int64_t s = 0;
  for (uint64_t i = 1; i <= n; i++) s += m[i];
  free(m); free(l);
  return s;
}

/* phi(n, b): count of 1 <= k <= n with no prime factor among the first b
 * primes -- by direct trial of each k (tiny n only; the definition itself). */
uint64_t orc_phi_brute(uint64_t n, uint32_t b)
{
  size_t np; uint32_t *pr = primes_upto(b ? 64 + (uint64_t)b * 32 : 2, &np);
  while (np < b) { free(pr); pr = primes_upto(4 * (uint64_t)b * 32 + 64, &np); }
  uint64_t c = 0;
  for (uint64_t k = 1; k <= n; k++) {
    int ok = 1;
    for (uint32_t i = 0; i < b; i++) if (k % pr[i] == 0) { ok = 0; break; }
    c += ok;
  }
  free(pr);
  return c;
}

/* Legendre recursion phi(x, a) = phi(x, a-1) - phi(x/p_a, a-1), phi(x,0)=x. */
static uint64_t legendre(uint64_t x, uint32_t a, const uint32_t *pr)
{
  if (a == 0) return x;
  if (x == 0) return 0;
  if (x <= pr[a - 1]) return 1;           /* only k = 1 survives */
  return legendre(x, a - 1, pr) - legendre(x / pr[a - 1], a - 1, pr);
}

uint64_t orc_phi_legendre(uint64_t x, uint32_t a)
{
  size_t np; uint32_t *pr = primes_upto(a ? 64 + (uint64_t)a * 32 : 2, &np);
  uint64_t r = legendre(x, a, pr);
  free(pr);
  return r;
}

/* ------------------------------------------------------------- params */

/* Parameters of one DR evaluation; every bound is an exact integer root. */
typedef struct {
  u128 x;
  uint64_t y, z_lo, z_hi, sqrtx, cbrtx, qrtx, sqrtz;
  uint32_t c;
} params_t;

/* c is clamped to pi(x^(1/4)) so that p_c <= x^(1/4) <= sqrt(z) <= cbrt(x)
 * and every leaf class below starts above p_c. */
static uint32_t clamp_c(uint32_t c, uint64_t qrtx)
{
  static const uint32_t sp[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  uint32_t k = 0;
  while (k < 8 && sp[k] <= qrtx) k++;
  return c < k ? c : k;
}

static void make_params(params_t *P, u128 x, uint64_t y, uint32_t c)
{
  P->x = x; P->y = y;
  u128 z = x / y;
  P->z_lo = (uint64_t)z; P->z_hi = (uint64_t)(z >> 64);
  P->sqrtx = orc_isqrt((uint64_t)(x >> 64), (uint64_t)x);
  P->cbrtx = orc_icbrt((uint64_t)(x >> 64), (uint64_t)x);
  P->qrtx = orc_isqrt(0, P->sqrtx);
  P->sqrtz = orc_isqrt(P->z_hi, P->z_lo);
  P->c = clamp_c(c, P->qrtx);
}

uint32_t orc_clamp_c(uint64_t xhi, uint64_t xlo, uint32_t c)
{
  return clamp_c(c, orc_iroot4(xhi, xlo));
}

/* ----------------------------------------------------------------- P2 */

/* P2(x, y) = sum over primes y < p <= sqrt(x) of [pi(x/p) - pi(p) + 1]
 * (src/P2.cu:76-133).  The reference walks hi-windows of primes above sqrt x
 * and counts pi(x/p) by upper_bound; here [sqrt x, z] is cut into chunks, each
 * chunk counts its primes and the pi(x/p) that land in it, then a prefix over
 * chunks makes the counts global. */
void orc_p2(uint64_t xhi, uint64_t xlo, uint64_t y, int threads, uint64_t out[2])
{
  u128 x = mk128(xhi, xlo);
  uint64_t sq = orc_isqrt(xhi, xlo);
  if (y >= sq) { put128(0, out); return; }
  u128 z128 = x / (y + 1);                         /* largest x/p, p > y */
  uint64_t zt = (uint64_t)z128;
  size_t np; uint32_t *pr = primes_upto(sq, &np);  /* all primes <= sqrt x */
  uint64_t a = count_le(pr, np, y);                /* pi(y) */
  /* P2 primes: pr[a .. np) ; x/p for these lies in [sqrt x, zt] */
  uint64_t lo = sq, hi = zt + 1;                   /* chunked range [lo, hi) */
  uint64_t rr = orc_isqrt(0, hi);
  size_t nb; uint32_t *bp = primes_upto(rr, &nb);
  const uint64_t CH = 1ull << 24;
  uint64_t base = lo & ~1ull;
  uint64_t nch = (hi - base + CH - 1) / CH;
  uint64_t *cnt = (uint64_t *)calloc(nch, sizeof(uint64_t));   /* primes in chunk */
  uint64_t *nq = (uint64_t *)calloc(nch, sizeof(uint64_t));    /* # p with x/p in chunk */
  i128 *loc = (i128 *)calloc(nch, sizeof(i128));                /* sum of local counts */
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
  {
    uint8_t *s = (uint8_t *)malloc(CH / 2);
#pragma omp for schedule(dynamic, 1)
    for (uint64_t k = 0; k < nch; k++) {
      uint64_t s0 = base + k * CH, s1 = s0 + CH; if (s1 > hi) s1 = hi;
      uint64_t nbits = (s1 - s0) / 2;
      memset(s, 1, nbits);
      for (size_t i = 1; i < nb; i++) {
        uint64_t p = bp[i];
        if (p * p >= s1) break;
        uint64_t st = p * p;
        if (st < s0 + 1) { st = ((s0 + 1 + p - 1) / p) * p; if (!(st & 1)) st += p; }
        for (uint64_t m = st; m < s1; m += 2 * p) s[(m - s0 - 1) / 2] = 0;
      }
      /* number of primes in [s0, s1) that are >= lo; s0 >= lo-1 so only n=1,2 edge */
      uint64_t c = 0;
      for (uint64_t i = 0; i < nbits; i++) { uint64_t n = s0 + 2 * i + 1; if (s[i] && n >= lo && n > 1) c++; }
      if (s0 <= 2 && s1 > 2 && lo <= 2) c++;
      cnt[k] = c;
      /* P2 primes p with s0 <= x/p < s1, i.e. x/s1 < p <= x/s0 */
      u128 pmax = s0 ? x / s0 : x; u128 pmin = x / s1;   /* p in (pmin, pmax] */
      uint64_t j0 = count_le(pr, np, pmin > sq ? sq : (uint64_t)pmin);
      uint64_t j1 = count_le(pr, np, pmax > sq ? sq : (uint64_t)pmax);
      if (j0 < a) j0 = a;
      i128 sl = 0; uint64_t q = 0;
      uint64_t i = 0, lc = (s0 <= 2 && lo <= 2) ? 1 : 0;   /* running count of primes <= v */
      for (uint64_t j = j1; j > j0; j--) {               /* p descending -> v ascending */
        uint64_t v = (uint64_t)(x / pr[j - 1]);           /* in [s0, s1) */
        for (; i < nbits && s0 + 2 * i + 1 <= v; i++) {
          uint64_t n = s0 + 2 * i + 1; if (s[i] && n >= lo && n > 1) lc++;
        }
        sl += lc; q++;
      }
      loc[k] = sl; nq[k] = q;
    }
    free(s);
  }
  /* pi(lo - 1) = pi(sqrt x - 1) */
  uint64_t pi_below = count_le(pr, np, lo - 1);
  i128 total = 0, pref = pi_below;
  for (uint64_t k = 0; k < nch; k++) { total += loc[k] + (i128)nq[k] * pref; pref += cnt[k]; }
  /* minus sum (pi(p) - 1) over P2 primes: pi(pr[j]) = j + 1 */
  for (uint64_t j = a; j < np; j++) total -= (i128)j;
  put128((u128)total, out);
  free(cnt); free(nq); free(loc); free(bp); free(pr);
}

/* ----------------------------------------------------------- Ordinary */

static const uint32_t SMALLP[8] = {2, 3, 5, 7, 11, 13, 17, 19};

/* S0 = sum over odd n <= y with lpf(n) > p_c of mu(n) * phi(x/n, c), with
 * phi(u, c) = floor(u/W_c)*totient(W_c) + phi(u mod W_c, c)   (src/ordinary.cu:23-57,101-115;
 * the reference's c-- / d_small[c] indexing makes p_c the c-th prime). */
void orc_ordinary(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, int threads, uint64_t out[2])
{
  u128 x = mk128(xhi, xlo);
  c = clamp_c(c, orc_iroot4(xhi, xlo));
  uint64_t W = 1, tot = 1;
  for (uint32_t i = 0; i < c; i++) { W *= SMALLP[i]; tot *= SMALLP[i] - 1; }
  uint32_t pc = c ? SMALLP[c - 1] : 1;
  /* phi(r, c) for 0 <= r < W by direct count */
  uint32_t *ph = (uint32_t *)malloc((W + 1) * sizeof(uint32_t));
  uint32_t run = 0; ph[0] = 0;
  for (uint64_t r = 1; r <= W; r++) {
    int ok = 1;
    for (uint32_t i = 0; i < c; i++) if (r % SMALLP[i] == 0) { ok = 0; break; }
    run += ok; ph[r] = run;
  }
  int8_t *mu; uint64_t *lpf;
  mu_lpf_upto(y, &mu, &lpf);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  i128 s = 0;
  int64_t nodd = (int64_t)((y + 1) / 2);
#pragma omp parallel
  {
    i128 ls = 0;
#pragma omp for schedule(static)
    for (int64_t i = 0; i < nodd; i++) {
      uint64_t n = 2 * (uint64_t)i + 1;
      if (lpf[n] > pc && mu[n]) {
        u128 m = x / n;
        u128 ph_ = (m / W) * tot + ph[(uint64_t)(m % W)];
        ls += mu[n] > 0 ? (i128)ph_ : -(i128)ph_;
      }
    }
#pragma omp critical
    s += ls;
  }
  put128((u128)s, out);
  free(ph); free(mu); free(lpf);
}

/* --------------------------------------------------------- S2_trivial */

/* S1 + U = C(pi(y) - pi(cbrt x), 2) + sum_{sqrt(z) < p <= cbrt(x)} [pi(y) - pi(x/p^2)]
 * (src/trivial.cu:31-62).  For p > sqrt(z) we have x/p^2 < y, for p <= cbrt x
 * x/p^2 >= p, so every counted q satisfies max(p, x/p^2) < q <= y. */
void orc_trivial(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, uint64_t out[2])
{
  u128 x = mk128(xhi, xlo);
  params_t P; make_params(&P, x, y, c);
  size_t np; uint32_t *pr = primes_upto(y, &np);
  uint64_t piy = np;
  uint64_t picb = count_le(pr, np, P.cbrtx < y ? P.cbrtx : y);
  uint64_t k = piy - picb;                         /* primes in (cbrt x, y] */
  u128 s = k ? (u128)k * (k - 1) / 2 : 0;
  uint64_t pclim = P.c ? SMALLP[P.c - 1] : 1;
  for (uint64_t j = count_le(pr, np, P.sqrtz); j < np && pr[j] <= P.cbrtx; j++) {
    uint64_t p = pr[j];
    if (p <= pclim) continue;
    u128 q = x / ((u128)p * p);
    uint64_t pq = count_le(pr, np, q > y ? y : (uint64_t)q);
    s += piy - pq;
  }
  put128(s, out);
  free(pr);
}

/* ------------------------------------------------------------ V (easy) */

/* V = sum_{x^(1/4) < p <= cbrt(x)} sum_{p < q <= min(y, x/p^2)} [pi(x/(pq)) - pi(p) + 2]
 * (src/V.cu:22-124, whose V_a/V_b split at p = sqrt(x/y) is the same set).
 * Every such leaf has x/(pq) < p^2, so phi(x/(pq), pi(p)-1) = pi(x/(pq)) - pi(p) + 2. */
/* part: 0 = all of V; 1 = V_a, x^(1/4) < p <= sqrt(x/y) (q runs to y: src/V.cu:58-87);
 * 2 = V_b, sqrt(x/y) < p <= cbrt(x) (q runs to x/p^2: src/V.cu:89-124).  The
 * split point is isqrt(floor(x/y)) = floor(sqrt(x/y)). */
void orc_easy_part(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, int threads, int part,
                   uint64_t out[2])
{
  u128 x = mk128(xhi, xlo);
  params_t P; make_params(&P, x, y, c);
  size_t np; uint32_t *pr = primes_upto(P.sqrtx, &np);   /* pi lookups up to sqrt x */
  uint64_t piy = count_le(pr, np, y);
  uint64_t j0 = count_le(pr, np, P.qrtx), j1 = count_le(pr, np, P.cbrtx);
  uint64_t js = count_le(pr, np, P.sqrtz);
  if (part == 1 && j1 > js) j1 = js;
  if (part == 2 && j0 < js) j0 = js;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  i128 s = 0;
#pragma omp parallel
  {
    i128 ls = 0;
#pragma omp for schedule(dynamic, 1)
    for (int64_t jj = (int64_t)j0; jj < (int64_t)j1; jj++) {
      uint64_t p = pr[jj];
      u128 xp = x / p;
      u128 qm = xp / p; uint64_t qmax = qm > y ? y : (uint64_t)qm;
      uint64_t k1 = count_le(pr, np, qmax);
      uint64_t pip = (uint64_t)jj + 1;
      for (uint64_t k = (uint64_t)jj + 1; k < k1 && k < piy; k++) {
        uint64_t u = (uint64_t)(xp / pr[k]);
        ls += (i128)count_le(pr, np, u) - (i128)pip + 2;
      }
    }
#pragma omp critical
    s += ls;
  }
  put128((u128)s, out);
  free(pr);
}

void orc_easy(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, int threads, uint64_t out[2])
{
  orc_easy_part(xhi, xlo, y, c, threads, 0, out);
}

/* ------------------------------------------------------------ S3 (hard) */

/* Hard leaves for p_c < p <= x^(1/4):
 *   S3 = - sum_b sum_{m} mu(m) * phi(x/(m p_b), b - 1),  y/p_b < m <= y, lpf(m) > p_b.
 * Restates the staged sieve of src/sieve/S2_hard_device.cu:264-355: the odd
 * numbers of [1, z] start crossed by p_1..p_c; at stage b the leaves of p_b
 * read phi(n, b-1) = survivors <= n, then p_b is crossed off.  Leaves of one
 * stage are walked with m descending, so n = x/(m p) ascends and
 * one running count per stage suffices (the reference's countUpSieve walk).
 * OpenMP: each thread owns a contiguous run of bitmap words; per stage the
 * threads count their survivors, take the counts of the runs below theirs as
 * the carry into their own run, evaluate the leaves whose n falls in their run
 * and cross p off in it (the per-block carry of src/sieve/S2_hard_host.cu:115-191). */
void orc_hard(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, int threads, uint64_t out[2])
{
  u128 x = mk128(xhi, xlo);
  params_t P; make_params(&P, x, y, c);
  c = P.c;
  size_t np; uint32_t *pr = primes_upto(y, &np);
  uint64_t B = count_le(pr, np, P.qrtx);            /* hard stages b = c+1 .. B */
  if (B <= c) { put128(0, out); free(pr); return; }
  int8_t *mu; uint64_t *lpf;
  mu_lpf_upto(y, &mu, &lpf);
  u128 z128 = x / y;
  uint64_t z = (uint64_t)z128;                       /* leaves have n <= z */
  uint64_t nbits = z / 2 + 1;                        /* bit i <-> 2i+1, covers 1..z(+1) */
  uint64_t nw = (nbits + 63) / 64;
  uint64_t *bm = (uint64_t *)malloc(nw * sizeof(uint64_t));
  int T = 1;
#ifdef _OPENMP
  T = threads > 0 ? threads : omp_get_max_threads();
#endif
  uint64_t *cnt = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
  i128 *part = (i128 *)calloc((size_t)T + 1, sizeof(i128));
#pragma omp parallel num_threads(T)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num(); nt = omp_get_num_threads();
#endif
    uint64_t per = (nw + (uint64_t)nt - 1) / (uint64_t)nt;
    uint64_t w0 = (uint64_t)t * per; if (w0 > nw) w0 = nw;
    uint64_t w1 = w0 + per; if (w1 > nw) w1 = nw;
    uint64_t bs = 64 * w0, be = 64 * w1;             /* own bits [bs, be) */
    for (uint64_t w = w0; w < w1; w++) {
      uint64_t v = 0;
      for (int k = 0; k < 64; k++) {
        uint64_t n = 2 * (w * 64 + k) + 1;
        int ok = 1;
        for (uint32_t i = 1; i < c; i++) if (n % SMALLP[i] == 0) { ok = 0; break; }
        if (ok) v |= 1ull << k;
      }
      bm[w] = v;
    }
    i128 s3 = 0;
    for (uint64_t b = c + 1; b <= B; b++) {
      uint64_t p = pr[b - 1];
      u128 xp = x / p;
      uint64_t own = 0;
      for (uint64_t w = w0; w < w1; w++) own += (uint64_t)__builtin_popcountll(bm[w]);
      cnt[t] = own;
#pragma omp barrier
      uint64_t run = 0;                              /* survivors in words [0, w) */
      for (int i = 0; i < t; i++) run += cnt[i];
      /* leaves with bit((x/p)/m) in [bs, be): x/(p m) >= 2 bs + 1 and < 2 be + 1 */
      uint64_t mlo = y / p;                          /* m > y/p */
      u128 top = xp / (2 * (u128)bs + 1), bot = xp / (2 * (u128)be + 1);
      uint64_t mhi = top < (u128)y ? (uint64_t)top : y;
      if (bot > (u128)mlo) mlo = (uint64_t)bot;
      uint64_t w = w0;
      for (uint64_t m = mhi; m > mlo; m--) {
        if (!(m & 1) || !mu[m] || lpf[m] <= p) continue;
        uint64_t n = (uint64_t)(xp / m);
        uint64_t bit = (n - 1) / 2;                  /* largest odd <= n */
        uint64_t tw = bit / 64;
        while (w < tw) run += (uint64_t)__builtin_popcountll(bm[w++]);
        uint64_t msk = (bit % 64 == 63) ? ~0ull : ((2ull << (bit % 64)) - 1);
        uint64_t phi = run + (uint64_t)__builtin_popcountll(bm[tw] & msk);
        s3 -= mu[m] * (i128)phi;
      }
      /* cross off odd multiples of p (p itself included) among own bits below nbits */
      uint64_t i0 = (p - 1) / 2, lim = be < nbits ? be : nbits;
      if (i0 < bs) i0 += (bs - i0 + p - 1) / p * p;
      for (uint64_t i = i0; i < lim; i += p) bm[i / 64] &= ~(1ull << (i % 64));
#pragma omp barrier
    }
    part[t] = s3;
  }
  i128 s3 = 0;
  for (int i = 0; i <= T; i++) s3 += part[i];
  put128((u128)s3, out);
  free(part); free(cnt); free(bm); free(mu); free(lpf); free(pr);
}

/* ------------------------------------------------ sigma partition of S2 */

/* The reference's second form of the easy leaves (src/Deleglise-Rivat/S2.cu:30-70
 * with the sigma terms of src/Deleglise-Rivat/sigma.cu:25-112, taking the
 * offsets of their 128-bit twins, sigma.cu:118-214):
 *   A  = sum_{x^1/4 < p <= x^1/3} sum_{p < q <= sqrt(x/p)} chi(x/pq) pi(x/pq),
 *        chi(n) = 2 if n < y else 1                              (S2.cu:72-110)
 *   s1 = C(pi(y) - pi(x^1/3), 2)
 *   s2 = pi(y) [pi(x^1/4)(pi(x^1/4) - 3)/2 - pi(sqrt z)(pi(sqrt z) - 3)/2]
 *   s3 = pi3(pi3 - 1)(2 pi3 - 1)/6 - pi3 - pi4(pi4 - 1)(2 pi4 - 1)/6 + pi4,
 *        pi3 = pi(x^1/3), pi4 = pi(x^1/4)
 *   s4 = pi(y) sum_{x^1/4 < p <= sqrt z} pi(x/(p y))
 *   s5 = sum_{sqrt z < p <= x^1/3} pi(min(y, x/p^2))
 *   s6 = - sum_{x^1/4 < p <= x^1/3} pi(floor(sqrt(x/p)))^2
 *   S2 = A + s2 + s3 + s4 + 2 s5 + s6  (s1 belongs to the trivial leaves).
 * S2 equals the direct V exactly when no prime lies in (x^1/4, x/y^2]; with
 * such primes A also counts pairs y < q <= sqrt(x/p).  out (two 64-bit words
 * each, two's complement): A, s1 .. s6, S2.  Needs x/x^(1/4) < 2^64. */
void orc_sigma(uint64_t xhi, uint64_t xlo, uint64_t y, int threads, uint64_t out[16])
{
  u128 x = mk128(xhi, xlo);
  uint64_t q4 = orc_iroot4(xhi, xlo), c3 = orc_icbrt(xhi, xlo), sq = orc_isqrt(xhi, xlo);
  u128 z128 = x / y;
  uint64_t sz = orc_isqrt((uint64_t)(z128 >> 64), (uint64_t)z128);
  size_t np; uint32_t *pr = primes_upto(sq, &np);           /* pi(n) for every n used is <= sqrt x */
  const i128 a = (i128)count_le(pr, np, y), b = (i128)count_le(pr, np, c3);
  const i128 c = (i128)count_le(pr, np, q4), d = (i128)count_le(pr, np, sz);
  const uint64_t j0 = count_le(pr, np, q4), j1 = count_le(pr, np, c3);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  i128 A = 0, s4 = 0, s5 = 0, s6 = 0;
#pragma omp parallel
  {
    i128 la = 0, l4 = 0, l5 = 0, l6 = 0;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t j = j0; j < j1; j++) {
      const uint64_t p = pr[j];
      const uint64_t xp = (uint64_t)(x / p);
      const uint64_t s = orc_isqrt(0, xp);
      for (uint64_t k = j + 1; k < np && pr[k] <= s; k++) {
        const uint64_t n = xp / pr[k];
        la += (i128)(n < y ? 2 : 1) * (i128)count_le(pr, np, n);
      }
      if (p <= sz) l4 += (i128)count_le(pr, np, xp / y);
      else {
        const uint64_t v = xp / p;
        l5 += (i128)count_le(pr, np, v < y ? v : y);
      }
      const i128 ps = (i128)count_le(pr, np, s);
      l6 -= ps * ps;
    }
#pragma omp critical
    { A += la; s4 += l4; s5 += l5; s6 += l6; }
  }
  s4 *= a;
  const i128 s1 = (a - b) * (a - b - 1) / 2;
  const i128 s2 = a * (c * (c - 3) / 2 - d * (d - 3) / 2);
  const i128 s3 = b * (b - 1) * (2 * b - 1) / 6 - b - c * (c - 1) * (2 * c - 1) / 6 + c;
  const i128 S2 = A + s2 + s3 + s4 + 2 * s5 + s6;
  put128((u128)A, out + 0); put128((u128)s1, out + 2); put128((u128)s2, out + 4);
  put128((u128)s3, out + 6); put128((u128)s4, out + 8); put128((u128)s5, out + 10);
  put128((u128)s6, out + 12); put128((u128)S2, out + 14);
  free(pr);
}

/* ------------------------------------------------------------ assembly */

/* terms: [0..1] S0, [2..3] trivial, [4..5] V, [6..7] S3, [8..9] P2, [10..11] pi(y),
 *        [12..13] pi(x).  Each is a two's-complement 128-bit (lo, hi) pair.
 * pi(x) = S0 + trivial + V + S3 + pi(y) - 1 - P2  (src/Deleglise-Rivat/deleglise-rivat.cpp:77). */
void orc_pi_dr(uint64_t xhi, uint64_t xlo, uint64_t y, uint32_t c, int threads, uint64_t terms[14])
{
  u128 x = mk128(xhi, xlo);
  uint64_t t[2];
  i128 s0, tr, v, s3, p2;
  orc_ordinary(xhi, xlo, y, c, threads, t); s0 = (i128)mk128(t[1], t[0]);
  orc_trivial(xhi, xlo, y, c, t);           tr = (i128)mk128(t[1], t[0]);
  orc_easy(xhi, xlo, y, c, threads, t);     v = (i128)mk128(t[1], t[0]);
  orc_hard(xhi, xlo, y, c, threads, t);     s3 = (i128)mk128(t[1], t[0]);
  orc_p2(xhi, xlo, y, threads, t);          p2 = (i128)mk128(t[1], t[0]);
  size_t np; uint32_t *pr = primes_upto(y, &np); free(pr);
  i128 pix = s0 + tr + v + s3 + (i128)np - 1 - p2;
  put128((u128)s0, terms + 0); put128((u128)tr, terms + 2); put128((u128)v, terms + 4);
  put128((u128)s3, terms + 6); put128((u128)p2, terms + 8); put128((u128)np, terms + 10);
  put128((u128)pix, terms + 12);
  (void)x;
}

/* y = floor(alpha * icbrt(x)), clamped to [icbrt(x), isqrt(x)]. */
uint64_t orc_choose_y(uint64_t xhi, uint64_t xlo, double alpha)
{
  uint64_t cb = orc_icbrt(xhi, xlo), sq = orc_isqrt(xhi, xlo);
  long double yy = (long double)alpha * (long double)cb;
  uint64_t y = (uint64_t)yy;
  if (y < cb) y = cb;
  if (y > sq) y = sq;
  if (y < 1) y = 1;
  return y;
}
