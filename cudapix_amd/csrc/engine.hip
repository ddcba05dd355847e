// engine.hip -- host orchestration of the MI355X pi(x) engine.
//
// One context = one device stream + resident tables (compact pi table,
// prime list, mu*lpf table, wheel patterns) + a batch arena.  A batch of
// queries is planned on the host (bounds by exact integer roots, tile and
// block-descriptor lists, tiles ordered longest-first), uploaded with one copy
// per array, and run as a fixed-shape launch sequence:
//   zero queue -> setup -> ordinary -> trivial -> easy -> segments(persistent)
//   -> hard carry -> combine
// which is captured once into a hipGraph and replayed while the arena keeps
// its capacity.  (The reference instead runs each term as its own host loop
// of kernel + cudaDeviceSynchronize + thrust reduce, with cudaDeviceReset
// between terms: src/Deleglise-Rivat/deleglise-rivat.cpp:50-80,
// src/sieve/S2_hard_host.cu:151-191, src/P2.cu:93-130.)
#include "engine.hpp"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#ifndef CPX_NO_ROCTX
#include <rocprofiler-sdk-roctx/roctx.h>
#endif

// A named range in a rocprofv3 marker trace (--marker-trace) around a host
// phase of a solve: table builds, planning, plan upload, kernel enqueue.
struct CpxRange {
#ifndef CPX_NO_ROCTX
  explicit CpxRange(const char *name) { roctxRangePushA(name); }
  ~CpxRange() { roctxRangePop(); }
#else
  explicit CpxRange(const char *) {}
#endif
};

// Test hook (cpx_debug_fail_alloc): the n-th device allocation from now
// fails with hipErrorOutOfMemory, as a full device would.
static std::atomic<int> g_alloc_fail{0};

static hipError_t cpx_dev_malloc(void **p, size_t bytes)
{
  int n = g_alloc_fail.load();
  while (n > 0 && !g_alloc_fail.compare_exchange_weak(n, n - 1)) {
  }
  if (n == 1) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  return hipMalloc(p, bytes);
}

extern "C" int cpx_debug_fail_alloc(int n)
{
  if (n < 0) return CPX_EINVAL;
  g_alloc_fail.store(n);
  return CPX_OK;
}
#include <thread>
#include <unistd.h>

static double now_ms()
{
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int ctx_fail(cpx_ctx *c, int code, const std::string &msg)
{
  if (c) c->err = msg;
  return code;
}

int hip_fail(cpx_ctx *c, hipError_t e, const char *what)
{
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory) return ctx_fail(c, CPX_ENOMEM, m);
  return ctx_fail(c, CPX_EDEVICE, m);
}

#define HIPCHK(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(c, e_, #call); \
  } while (0)

template <typename T>
static int dbuf_reserve(cpx_ctx *c, DBuf<T> &b, size_t n, uint64_t *acct = nullptr)
{
  if (n <= b.cap && b.p) return CPX_OK;
  if (b.p) {
    // a batch queued in another slot may still read the old table
    (void)hipStreamSynchronize(c->stream);
    if (acct) *acct -= b.bytes();
    (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
  }
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  hipError_t e = cpx_dev_malloc((void **)&b.p, bytes);
  if (e != hipSuccess) { b.p = nullptr; (void)hipGetLastError(); return hip_fail(c, e, "hipMalloc"); }
  b.cap = std::max<size_t>(n, 1);
  if (acct) *acct += b.bytes();
  return CPX_OK;
}

// Frees b and takes its bytes off the account they were charged to.
template <typename T>
static void dbuf_free(DBuf<T> &b, uint64_t *acct = nullptr)
{
  if (b.p) {
    (void)hipFree(b.p);
    if (acct) *acct -= b.bytes();
  }
  b.p = nullptr;
  b.cap = 0;
}

// Frees the size-dependent tables (the bootstrap tables and the wheels stay,
// and stay counted in table_bytes).
void free_tables(cpx_ctx *c)
{
  if (c->stream) (void)hipStreamSynchronize(c->stream);   // no queued batch reads them any more
  uint64_t *a = &c->table_bytes;
  dbuf_free(c->pitab, a); dbuf_free(c->primes, a); dbuf_free(c->prinv, a); dbuf_free(c->mlpf, a);
  dbuf_free(c->scan_tot, a); dbuf_free(c->scan_base, a);
  c->pitab_nent = c->pitab_limit = c->prime_limit = c->nprimes = c->mlpf_limit = 0;
  c->h_primes.clear();
  c->generation++;
}

uint64_t host_pi(const cpx_ctx *c, uint64_t v)
{
  return (uint64_t)(std::upper_bound(c->h_primes.begin(), c->h_primes.end(), v) -
                    c->h_primes.begin());
}

// ------------------------------------------------------------------ alpha
// The engine's own alpha model: y = alpha * cbrt(x) trades hard-leaf sieve
// length x/y against the easy-leaf and mu*lpf work that grows with y.  Fitted
// from MI355X sweeps of this engine (alpha that minimised one-query wall time
// per decade); the reference instead borrows primecount's CPU fit
// (src/main.cpp:128-160).
static double alpha_model(u128 x)
{
  // best alpha per decade from a sweep over alpha in {1..32} at x = 1e10..1e16
  // (one query, resident tables), linear in log10(x) between the points and
  // 1.5 per decade beyond 1e16
  static const double pts[][2] = {{11, 1.0}, {12, 2.0}, {13, 2.0}, {14, 2.5}, {15, 5.0}, {16, 6.0}};
  const int np = (int)(sizeof(pts) / sizeof(pts[0]));
  const double lx = std::log10((double)x);
  double a;
  if (lx <= pts[0][0]) {
    a = pts[0][1];
  } else if (lx >= pts[np - 1][0]) {
    a = pts[np - 1][1] + 1.5 * (lx - pts[np - 1][0]);
  } else {
    int i = 1;
    while (pts[i][0] < lx) i++;
    const double f = (lx - pts[i - 1][0]) / (pts[i][0] - pts[i - 1][0]);
    a = pts[i - 1][1] + f * (pts[i][1] - pts[i - 1][1]);
  }
  return std::min(a, 24.0);
}

uint64_t choose_y_host(const cpx_ctx *c, u128 x)
{
  uint64_t cb = cpx_icbrt128(x), sq = cpx_isqrt128(x);
  double alpha = c->opts.alpha > 0 ? c->opts.alpha : alpha_model(x) * c->alpha_scale;
  long double yy = (long double)alpha * (long double)cb;
  uint64_t y = (uint64_t)yy;
  if (y < cb) y = cb;
  if (y > sq) y = sq;
  if (y < 1) y = 1;
  return y;
}

int make_host_query(cpx_ctx *c, u128 x, uint64_t y, uint32_t cc, HostQuery *hq)
{
  HostQuery h{};
  h.x = x;
  if ((x >> 100) != 0) return ctx_fail(c, CPX_ERANGE, "x above 2^100 is not supported");
  if (cc == 0) cc = c->opts.c;
  if (cc < 1 || cc > 8) return ctx_fail(c, CPX_EINVAL, "c must be in 1..8");
  if (x < CPX_SMALL_X) {
    h.small = true;
    h.y = 1; h.c = 1;
    *hq = h;
    return CPX_OK;
  }
  h.sqrtx = cpx_isqrt128(x);
  h.cbrtx = cpx_icbrt128(x);
  h.qrtx = cpx_isqrt128(h.sqrtx);
  if (y == 0) y = choose_y_host(c, x);
  if (y < h.cbrtx || y > h.sqrtx)
    return ctx_fail(c, CPX_ERANGE, "y must lie in [icbrt(x), isqrt(x)]");
  if (y >= (1ull << 31)) return ctx_fail(c, CPX_ERANGE, "y >= 2^31 is not supported");
  h.y = y;
  u128 z = x / y;
  // leaf quotients x/(m p) < z use the f64-estimate divide, exact below 2^50
  if (z >> 50) return ctx_fail(c, CPX_ERANGE, "x / y must be below 2^50");
  h.z = (uint64_t)z;
  h.zt = (uint64_t)(x / (y + 1));
  h.sqrtz = cpx_isqrt128(z);
  // c <= pi(x^(1/4)) keeps p_c <= x^(1/4) <= sqrt(z) <= cbrt(x)
  static const uint32_t sp[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  uint32_t k = 0;
  while (k < 8 && sp[k] <= h.qrtx) k++;
  h.c = cc < k ? cc : k;
  // the easy leaves divide x / p for p > x^(1/4) as 64-bit values (x < ~2^85);
  // the hard segment tiles hold x / p in 80 bits
  if ((x / h.qrtx) >> 64) return ctx_fail(c, CPX_ERANGE, "x / x^(1/4) must fit 64 bits");
  if ((x / sp[h.c < 8 ? h.c : 7]) >> 80) return ctx_fail(c, CPX_ERANGE, "x / p_(c+1) must be below 2^80");
  *hq = h;
  return CPX_OK;
}

// ---------------------------------------------------------------- tables

static int build_bootstrap(cpx_ctx *c)
{
  if (c->small.p) return CPX_OK;
  int rc;
  // primes below 2^19 (43,390 of them): they sieve a pi table to 2^38
  if ((rc = dbuf_reserve(c, c->small, 8192, &c->table_bytes))) return rc;
  if ((rc = dbuf_reserve(c, c->small_primes, 44000, &c->table_bytes))) return rc;
  HIPCHK(launch_small_sieve(c->small.p, 8192, c->stream));
  HIPCHK(launch_extract_primes(c->small.p, (1u << 19) - 1, c->small_primes.p, 44000, c->stream));
  PiEntry last;
  HIPCHK(hipMemcpyAsync(&last, c->small.p + 8191, sizeof(last), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->n_small_primes = 1 + last.count + (uint32_t)__builtin_popcount(last.bits);
  return CPX_OK;
}

int build_wheels(cpx_ctx *c, uint32_t want)
{
  uint32_t cmax = std::max(c->opts.c, want);
  if (cmax > 8) cmax = 8;
  if (c->wheel_c >= cmax && c->wheel.p) return CPX_OK;
  static const uint32_t sp[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  std::vector<WheelInfo> wi(9);
  uint64_t off = 0, W = 1, tot = 1, period = 1;
  for (uint32_t k = 1; k <= cmax; k++) {
    W *= sp[k - 1];
    tot *= sp[k - 1] - 1;
    if (k >= 2) period *= sp[k - 1];
    wi[k].off = off;
    wi[k].period = (uint32_t)period;
    wi[k].W = (uint32_t)W;
    wi[k].totient = (uint32_t)tot;
    wi[k].pc = sp[k - 1];
    off += period + 1;
  }
  int rc;
  if ((rc = dbuf_reserve(c, c->wheel, off, &c->table_bytes))) return rc;
  if ((rc = dbuf_reserve(c, c->winfo, 9, &c->table_bytes))) return rc;
  size_t maxchunks = (period + 4095) / 4096 + 2;
  if ((rc = dbuf_reserve(c, c->scan_tot, std::max<size_t>(maxchunks, c->scan_tot.cap), &c->table_bytes)))
    return rc;
  if ((rc = dbuf_reserve(c, c->scan_base, std::max<size_t>(maxchunks, c->scan_base.cap), &c->table_bytes)))
    return rc;
  for (uint32_t k = 1; k <= cmax; k++) {
    HIPCHK(launch_wheel(c->wheel.p + wi[k].off, wi[k].period, k, c->stream));
    HIPCHK(launch_scan_entries(c->wheel.p + wi[k].off, wi[k].period, c->scan_tot.p,
                               c->scan_base.p, c->stream));
  }
  HIPCHK(hipMemcpyAsync(c->winfo.p, wi.data(), 9 * sizeof(WheelInfo), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->h_winfo = wi;
  c->wheel_c = cmax;
  c->generation++;
  return CPX_OK;
}

// The table limits that needs (R: pi table, P: prime list, Y: mu*lpf) ask for.
static void table_limits(uint64_t R, uint64_t P, uint64_t Y, uint64_t *Rm_, uint64_t *Pm_, uint64_t *Ym_)
{
  if (Y < 64) Y = 64;
  uint64_t Pm = std::max<uint64_t>(P, cpx_isqrt128((u128)Y + 16384) + 1);
  Pm = std::max<uint64_t>(Pm, Y);
  *Rm_ = std::max<uint64_t>({R, Pm, (uint64_t)1 << 14});
  *Pm_ = Pm;
  *Ym_ = Y;
}

// Largest pi-table reach: entries hold 32-bit prefix counts, and
// pi(1.04e11) is about 4.28e9 < 2^32 (the 8192-entry rounding adds at most
// 2^19 integers); the bootstrap primes below 2^19 sieve to 2^38.  So x <=
// (1.04e11)^2 ~ 1.08e22 for the tables that reach sqrt x.
static const uint64_t PITAB_MAX = 104000000000ull;

bool tables_cover(const cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y)
{
  uint64_t Rm, Pm, Ym;
  table_limits(R, P, Y, &Rm, &Pm, &Ym);
  return c->small.p && c->pitab.p && c->pitab_limit >= Rm && c->prime_limit >= Pm && c->mlpf_limit >= Ym;
}

int ensure_tables(cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y)
{
  CpxRange range("cpx:tables");
  int rc;
  if ((rc = build_bootstrap(c))) return rc;
  if ((rc = build_wheels(c, 0))) return rc;
  uint64_t Rm, Pm;
  table_limits(R, P, Y, &Rm, &Pm, &Y);
  if (!c->opts.fresh_tables && c->pitab_limit >= Rm && c->prime_limit >= Pm && c->mlpf_limit >= Y)
    return CPX_OK;
  if (Rm > PITAB_MAX)
    return ctx_fail(c, CPX_ERANGE, "pi table above 1.04e11 (its 32-bit prefix counts) is not supported");
  if (!c->opts.fresh_tables) {                      // headroom for later, larger queries
    Rm = std::max<uint64_t>(Rm, c->pitab_limit + c->pitab_limit / 4);
    Pm = std::max<uint64_t>(Pm, c->prime_limit + c->prime_limit / 4);
    Y = std::max<uint64_t>(Y, c->mlpf_limit + c->mlpf_limit / 4);
    Rm = std::min<uint64_t>(std::max<uint64_t>({Rm, Pm, Y}), PITAB_MAX);
  }
  if (Pm >= (1ull << 32)) return ctx_fail(c, CPX_ERANGE, "prime table above 2^32 is not supported");
  free_tables(c);
  // pi table over [0, 64 * nent), nent a multiple of the 8192-entry segment
  uint64_t nent = (Rm + 1 + 63) / 64;
  nent = (nent + 8191) / 8192 * 8192;
  if ((rc = dbuf_reserve(c, c->pitab, nent, &c->table_bytes))) return rc;
  size_t nch = (nent + 4095) / 4096 + 2;
  if ((rc = dbuf_reserve(c, c->scan_tot, std::max<size_t>(nch, c->scan_tot.cap), &c->table_bytes))) return rc;
  if ((rc = dbuf_reserve(c, c->scan_base, std::max<size_t>(nch, c->scan_base.cap), &c->table_bytes)))
    return rc;
  HIPCHK(launch_pitab_segments(c->pitab.p, nent, c->small_primes.p, c->n_small_primes,
                               c->scan_tot.p, c->scan_base.p, c->stream));
  c->pitab_nent = nent;
  c->pitab_limit = nent * 64 - 1;
  // prime list <= Pm
  PiEntry e;
  HIPCHK(hipMemcpyAsync(&e, c->pitab.p + (Pm >> 6), sizeof(e), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  uint32_t kk = (uint32_t)(((Pm & 63) + 1) >> 1);
  uint64_t np = 1 + e.count + (uint64_t)__builtin_popcount(e.bits & (uint32_t)((1ull << kk) - 1));
  if ((rc = dbuf_reserve(c, c->primes, np + 1, &c->table_bytes))) return rc;
  HIPCHK(launch_extract_primes(c->pitab.p, Pm, c->primes.p, np + 1, c->stream));
  if ((rc = dbuf_reserve(c, c->prinv, np + 1, &c->table_bytes))) return rc;
  HIPCHK(launch_recip(c->primes.p, c->prinv.p, np, c->stream));
  c->h_primes.resize(np);
  HIPCHK(hipMemcpyAsync(c->h_primes.data(), c->primes.p, np * sizeof(uint32_t),
                        hipMemcpyDeviceToHost, c->stream));
  c->nprimes = np;
  c->prime_limit = Pm;
  // mu * lpf for odd m <= Y
  uint64_t nodd = (Y + 1) / 2;
  uint64_t nodd_alloc = (nodd + 4095) / 4096 * 4096;
  if ((rc = dbuf_reserve(c, c->mlpf, nodd_alloc, &c->table_bytes))) return rc;
  HIPCHK(launch_mlpf(c->mlpf.p, nodd_alloc, c->primes.p, (uint32_t)np, c->stream));
  c->mlpf_limit = 2 * nodd_alloc - 1;
  HIPCHK(hipStreamSynchronize(c->stream));
  c->generation++;
  return CPX_OK;
}

// ---------------------------------------------------------------- planning

static uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

static const uint64_t ORD_CH = 8192, TRIV_CH = 4096, EASY_WORK = 131072, P2_CH = 65536;

static void plan_needs(const cpx_ctx *c, const HostQuery &h, uint32_t mask, uint64_t SB,
                       uint64_t *R, uint64_t *P, uint64_t *Y)
{
  if (h.small) { *R = std::max<uint64_t>(*R, (uint64_t)h.x); return; }
  *R = std::max(*R, std::max(h.sqrtx, h.y));
  *Y = std::max(*Y, h.y);
  uint64_t p = h.y;
  if ((mask & TERM_P2) && h.y < h.sqrtx) {
    if (h.p2tab) {                  // lookups of x/p for the primes p <= sqrt x
      *R = std::max(*R, h.zt);
      p = std::max(p, h.sqrtx);
    } else {                        // sieving primes of the last P2 segment
      uint64_t e = h.sqrtx & ~63ull;
      uint64_t nseg = ceil_div(h.zt + 1 - e, 2 * SB);
      p = std::max<uint64_t>(p, cpx_isqrt128((u128)e + nseg * 2 * SB) + 1);
    }
  }
  *P = std::max(*P, p);
  (void)c;
}

// One query's share of the plan, with query-local offsets (rows, parts,
// counters); build_plan places the shares.
struct QPlan {
  QueryDev d{};
  std::vector<BlockDesc> ord, triv, easy, fin, p2d;
  std::vector<uint32_t> fin_rows;    // rows each carry block scans (its first stage's active rows)
  std::vector<HardRowInfo> rowinfo;  // replicas and M rows are set per group
  std::vector<float> rleaves;        // estimated leaves of each hard row
  std::vector<float> rleafy;         // estimated leafy stages of each hard row
  uint64_t parts = 0, tm = 0, mm = 0;
};

// Tiles of one group of queries (see GroupDev).
struct GPlan {
  std::vector<TileDev> tiles;
  std::vector<uint16_t> tbucket;     // cost bucket of each tile (longest first)
};

// Where one query's descriptors and rows go in the batch arrays.
struct PlanOff { size_t ord, triv, easy, fin, p2d, rows; };

// A batch plan: queries, groups and sorted tiles; the per-query descriptors
// and rows stay in the query shares and are written straight into the
// pinned upload buffer at their offsets.
struct Plan {
  std::vector<QueryDev> q;
  std::vector<TileDev> tiles;
  std::vector<GroupDev> groups;
  std::vector<QPlan> qps;
  std::vector<PlanOff> off;
  std::vector<BlockDesc> fin;        // carry blocks of all queries, longest first
  bool stream = false;               // a streaming cohort: ordinary and trivial as tiles,
                                     // stage table from the host
  Caps base{};                       // where its entries start in each arena array (a
                                     // streaming cohort's slot; zero for a batch)
  size_t nord = 0, ntriv = 0, neasy = 0, nfin = 0, np2d = 0, nrows = 0;
  uint64_t stage_total = 0, hard_rows = 0, hard_tm = 0, p2_rows = 0, parts = 0;
  uint64_t hard_mrows = 0, hard_mm = 0;
  uint64_t hard_segments = 0, p2_segments = 0;
};

// Cost buckets: 1/8 octave of the tile's cost estimate.
static const int PLAN_NB = 512;
static inline uint16_t cost_bucket(double cost)
{
  // floor(8 log2 cost) from the exponent and the top mantissa bits (the
  // eighth-octave thresholds of [1, 2) are 2^(k/8))
  static const double thr[8] = {1.0, 1.0905077326652577, 1.1892071150027210, 1.2968395546510096,
                                1.4142135623730951, 1.5422108254079407, 1.6817928305074290,
                                1.8340080864093424};
  if (!(cost > 1.0)) return 0;
  int e;
  const double m = 2.0 * std::frexp(cost, &e);       // cost = m * 2^(e-1), m in [1, 2)
  int k = 7;
  while (k > 0 && m < thr[k]) k--;
  const int b = 8 * (e - 1) + k;
  return (uint16_t)std::min(std::max(b, 0), PLAN_NB - 1);
}

static void plan_query(const cpx_ctx *c, const HostQuery &h, uint32_t i, uint32_t mask, QPlan &qp)
// (i: the query's index in the arena's query array, which its descriptors name)
{
  const uint64_t SB = c->shape.seg_bits();
  const uint32_t T = (uint32_t)c->shape.threads;
  QueryDev &d = qp.d;
  d.x_lo = (uint64_t)h.x;
  d.x_hi = (uint64_t)(h.x >> 64);
  d.omega = 0;
  d.pad_o = 0;
  if (h.small) {
    d.small = 1;
    return;
  }
  d.y = h.y; d.z = h.z; d.sqrtx = h.sqrtx; d.cbrtx = h.cbrtx; d.qrtx = h.qrtx; d.sqrtz = h.sqrtz;
  d.p2_zt = h.zt;
  // the wheel pattern is word-aligned: segments start on a multiple of 64
  d.p2_e = h.sqrtx & ~63ull;
  d.c = h.c;
  uint64_t B = host_pi(c, h.qrtx);
  d.B = (uint32_t)B;
  d.nstage = ((mask & TERM_HARD) && B > h.c) ? (uint32_t)(B - h.c) : 0;
  d.a = (uint32_t)host_pi(c, h.y);
  d.pi_cbrt = (uint32_t)host_pi(c, std::min(h.cbrtx, h.y));
  d.pi_qrt = (uint32_t)B;
  d.pi_sqrtz = (uint32_t)host_pi(c, h.sqrtz);
  // segments of 2*SB integers from 0: the hard leaves use [0, z]; P2 counts
  // primes in (p2_e, zt] in the segments that hold them (a P2 epilogue of
  // the same tile), so the tiles reach z, or zt when only P2 is asked for --
  // unless P2 comes from pi-table lookups
  const bool p2 = (mask & TERM_P2) && h.y < h.sqrtx;
  d.p2tab = p2 && h.p2tab;
  d.p2_seg0 = d.p2_e / (2 * SB);
  d.p2_nseg = p2 && !h.p2tab ? h.zt / (2 * SB) + 1 - d.p2_seg0 : 0;
  d.hard_nseg = d.nstage ? ceil_div(h.z + 1, 2 * SB) : (d.p2_nseg ? d.p2_seg0 + d.p2_nseg : 0);
  uint32_t part = 0;
  // ordinary: odd n <= y
  if (mask & TERM_ORD) {
    uint64_t nodd = (h.y + 1) / 2;
    for (uint64_t b = 0; b < nodd; b += ORD_CH)
      qp.ord.push_back({i, part++, b, std::min(nodd, b + ORD_CH)});
  }
  d.n_ord_parts = part;
  // trivial U: prime indices [pi(sqrt z), pi(cbrt x))
  if ((mask & TERM_TRIV) && d.pi_cbrt > d.pi_sqrtz) {
    for (uint64_t b = d.pi_sqrtz; b < d.pi_cbrt; b += TRIV_CH)
      qp.triv.push_back({i, part++, b, std::min<uint64_t>(d.pi_cbrt, b + TRIV_CH)});
  }
  d.n_triv_parts = part - d.n_ord_parts;
  // easy V: prime indices [pi(x^1/4), pi(cbrt x)), grouped by q-count
  uint64_t ej0 = d.pi_qrt, ej1 = d.pi_cbrt;
  if (mask & TERM_EASY_A) ej1 = std::min<uint64_t>(ej1, d.pi_sqrtz);
  if (mask & TERM_EASY_B) ej0 = std::max<uint64_t>(ej0, d.pi_sqrtz);
  if ((mask & TERM_EASY) && ej1 > ej0) {
    // work estimate only: q-count pi(min(y, x/p^2)) - j - 1, with x/p^2 in
    // doubles and pi by a pointer that only moves down as p grows
    uint64_t j0 = ej0, work = 0;
    const double fx = (double)h.x;
    uint64_t k1 = d.a;                                  // pi(y)
    for (uint64_t j = ej0; j < ej1; j++) {
      const double p = (double)c->h_primes[j];
      const double qm = std::min((double)h.y, fx / (p * p));
      while (k1 > 0 && (double)c->h_primes[k1 - 1] > qm) k1--;
      work += (k1 > j + 1 ? k1 - j - 1 : 0) + 64;
      if (work >= EASY_WORK || j + 1 == ej1) {
        qp.easy.push_back({i, part++, j0, j + 1});
        j0 = j + 1;
        work = 0;
      }
    }
  }
  d.n_easy_parts = part - d.n_ord_parts - d.n_triv_parts;
  // hard carry: one block per CPX_FIN_STAGES stages
  for (uint64_t b = 0; b < d.nstage; b += CPX_FIN_STAGES)
    qp.fin.push_back({i, part++, b, std::min<uint64_t>(d.nstage, b + CPX_FIN_STAGES)});
  // hard leaves with x/(pq) < p^2: at the sparse stages (p^2 > y) every leaf
  // is a prime q in (p, y], and for q > x/p^3 its phi(x/(pq), pi(p) - 1) is
  // pi(x/(pq)) - pi(p) + 2, a pi-table lookup as V's leaves are.  The segment
  // sieve counts only q <= min(y, x/p^3) (StageDev::xp3); the rest go to
  // easy-leaf tiles whose descriptors sit in the hard part group (the tile
  // negates their sums: the hard term is minus the sum of mu * phi, and
  // mu(q) = -1).  The tiles take x/p in 64 bits, so x/p < 2^64 is required
  // at the first such p.  About half of all hard leaves at 1e13..1e16.
  if (d.nstage) {
    const uint64_t sy = cpx_isqrt128((u128)h.y);                         // p > sy <=> p^2 > y
    const uint64_t js = std::max<uint64_t>(h.c, host_pi(c, sy));         // first such stage prime
    if (js < d.B && ((h.x / c->h_primes[js]) >> 64) == 0) {
      d.omega = 1;
      const uint64_t ky = d.a;                                          // pi(y)
      uint64_t j0 = js, work = 0;
      for (uint64_t j = js; j < d.B; j++) {
        const uint64_t p = c->h_primes[j];
        const u128 t3 = h.x / ((u128)p * p * p);
        const uint64_t k0 = std::max<uint64_t>(j + 1, t3 >= h.y ? ky : host_pi(c, (uint64_t)t3));
        work += (ky > k0 ? ky - k0 : 0) + 64;
        if (work >= EASY_WORK || j + 1 == d.B) {
          qp.easy.push_back({i, part++, j0, j + 1});
          j0 = j + 1;
          work = 0;
        }
      }
    }
  }
  d.n_hard_parts = part - d.n_ord_parts - d.n_triv_parts - d.n_easy_parts;
  // P2 by lookups: prime indices (pi(y), pi(sqrt x)] (primes[k] is the (k+1)-th prime)
  if (d.p2tab) {
    const uint64_t k0 = d.a, k1 = host_pi(c, h.sqrtx);
    for (uint64_t b = k0; b < k1; b += P2_CH) qp.p2d.push_back({i, part++, b, std::min(k1, b + P2_CH)});
  }
  d.n_p2_parts = part - d.n_ord_parts - d.n_triv_parts - d.n_easy_parts - d.n_hard_parts;
  qp.parts = part;
  // hard rows: active stages per segment and, for the low (leaf-dense)
  // segments, an estimate of the leaves to evaluate (replicas are chosen per
  // group, from the members' leaves of the segment)
  std::vector<double> xps;
  if (d.hard_nseg) {
    xps.resize(d.nstage);
    for (uint32_t k = 0; k < d.nstage; k++) xps[k] = (double)(h.x / c->h_primes[h.c + k]);
  }
  const double leaf_target = 256.0 * T;
  uint32_t sparse_run = 0;
  uint64_t toff = 0;
  uint64_t act = d.nstage;                             // non-increasing in the segment
  qp.rowinfo.reserve(d.hard_nseg);
  qp.rleaves.reserve(d.hard_nseg);
  qp.rleafy.reserve(d.hard_nseg);
  for (uint64_t s = 0; s < d.hard_nseg; s++) {
    const uint64_t lo = s * 2 * SB, hi = lo + 2 * SB;
    // active stages: p_b^2 <= x / lo, i.e. p_b^2 * lo <= x (the kernel's test)
    while (act > 0) {
      const u128 p = c->h_primes[h.c + act - 1];
      if (p * p * (u128)lo <= h.x) break;
      act--;
    }
    double leaves = 0, leafy = 0;
    const bool est = sparse_run < 4;
    const double inv_lo = lo ? 1.0 / (double)lo : 0.0, inv_hi = 1.0 / (double)hi;
    const double fy = (double)h.y;
    // (a sample of at most 32 stages, scaled: the estimate only sizes replicas)
    const uint32_t kstep = act > 32 ? (uint32_t)((act + 31) / 32) : 1u;
    for (uint32_t k = 0; est && k < act; k += kstep) {
      const double p = (double)c->h_primes[h.c + k];
      const double xp = (double)xps[k];
      double mhi = lo ? std::min(fy, std::floor(xp * inv_lo)) : fy;
      double mlo = std::max(std::floor(xp * inv_hi), std::floor(fy / p));
      const bool sparse = p * p > fy;
      if (sparse) mlo = std::max(mlo, p);
      if (sparse && d.omega) mhi = std::min(mhi, std::floor(xp / (p * p)));
      if (mhi > mlo) {
        leafy += kstep;
        leaves += kstep * (sparse ? (mhi - mlo) / std::log(mhi) : 0.5 * (mhi - mlo));
      }
    }
    if (est) sparse_run = leaves < leaf_target / 4 ? sparse_run + 1 : 0;
    qp.rowinfo.push_back({0u, 0u, (uint32_t)act, 0u, toff, 0u});
    qp.rleaves.push_back((float)leaves);
    qp.rleafy.push_back((float)leafy);
    toff += act;
  }
  qp.tm = toff;
  // carry blocks: rows with their first stage active (act is non-increasing)
  qp.fin_rows.resize(qp.fin.size());
  uint64_t r = d.hard_nseg;
  for (size_t k = 0; k < qp.fin.size(); k++) {
    while (r > 0 && qp.rowinfo[r - 1].act <= qp.fin[k].begin) r--;
    qp.fin_rows[k] = (uint32_t)r;
  }
}

// Groups of queries that share the segment sieve: queries with hard rows,
// in increasing x, with the same wheel, at most CPX_GROUP, and every
// member's P2 range (from p2_e) above every member's stage primes (segment 0
// loses its stage primes to the sieve, so they must not be counted as P2
// primes of another member).  A query with y >= 2^28 stays alone, which
// keeps a wave's leaf list of a group (at most y/2 candidates per member)
// below 2^30 entries.
static void make_groups(const cpx_ctx *c, const std::vector<HostQuery> &hq,
                        const std::vector<QPlan> &qps, std::vector<GroupDev> &groups)
{
  std::vector<uint32_t> hs;
  for (uint32_t i = 0; i < (uint32_t)qps.size(); i++)
    if (qps[i].d.hard_nseg) hs.push_back(i);
  std::stable_sort(hs.begin(), hs.end(), [&](uint32_t a, uint32_t b) { return hq[a].x < hq[b].x; });
  GroupDev cur{};
  bool cur_alone = false;
  uint64_t smax = 0, p2min = ~0ull;
  for (uint32_t i : hs) {
    const QueryDev &d = qps[i].d;
    const uint64_t si = d.nstage ? c->h_primes[d.c + d.nstage - 1] : 0;
    const uint64_t pi = d.p2_nseg ? d.p2_e : ~0ull;
    const bool alone = d.y >= (1ull << 28);
    if (cur.n && (cur.n == CPX_GROUP || qps[cur.q[0]].d.c != d.c || alone || cur_alone ||
                  std::max(smax, si) > std::min(p2min, pi))) {
      groups.push_back(cur);
      cur = GroupDev{};
      smax = 0;
      p2min = ~0ull;
    }
    cur.q[cur.n++] = i;
    cur_alone = alone;
    smax = std::max(smax, si);
    p2min = std::min(p2min, pi);
  }
  if (cur.n) groups.push_back(cur);
}

// Replicas, M rows and tiles of one group: a segment's replica count comes
// from all members' leaves there; every member with the segment in its hard
// range gets that many M rows.  Tile cost: the shared sieve once, each
// member's leaves, and on replica 0 the P2 sieve and counts.
static void plan_group(const cpx_ctx *c, const GroupDev &g, uint32_t gi, std::vector<QPlan> &qps,
                       GPlan &gp)
{
  const uint64_t SB = c->shape.seg_bits();
  const uint32_t T = (uint32_t)c->shape.threads;
  const double leaf_target = 256.0 * T;
  uint64_t nseg = 0;
  for (uint32_t m = 0; m < g.n; m++) nseg = std::max(nseg, qps[g.q[m]].d.hard_nseg);
  const uint32_t cw = qps[g.q[0]].d.c;
  uint64_t mrows[CPX_GROUP] = {0}, moff[CPX_GROUP] = {0};
  uint64_t nsp_i = 0;                                  // pi(sqrt(hi)), grows with s
  gp.tiles.reserve(nseg);
  gp.tbucket.reserve(nseg);
  for (uint64_t s = 0; s < nseg; s++) {
    const uint64_t hi = (s + 1) * 2 * SB;
    double leaves = 0, lcost = 0;
    uint32_t act = 0, np2 = 0;
    for (uint32_t m = 0; m < g.n; m++) {
      const QPlan &qp = qps[g.q[m]];
      if (s >= qp.d.hard_nseg) continue;
      leaves += qp.rleaves[s];
      act = std::max(act, qp.rowinfo[s].act);
      if (s >= qp.d.p2_seg0 && s < qp.d.p2_seg0 + qp.d.p2_nseg) np2++;
    }
    const uint32_t nrep = (uint32_t)std::min(64.0, std::max(1.0, std::ceil(leaves / leaf_target)));
    for (uint32_t m = 0; m < g.n; m++) {
      QPlan &qp = qps[g.q[m]];
      if (s >= qp.d.hard_nseg) continue;
      HardRowInfo &ri = qp.rowinfo[s];
      ri.nrep = nrep;
      ri.mrow = (uint32_t)mrows[m];
      ri.moff = moff[m];
      mrows[m] += nrep;
      moff[m] += (uint64_t)nrep * ri.act;
      lcost += qp.rleafy[s] * 1.5 + 2.0 * qp.rleaves[s] / ((double)nrep * T);
    }
    const double cost = 1.0 + (double)act * (0.5 + 4.0 * SB / (64.0 * 1024.0 * T)) + lcost;
    // replica 0 of a segment of a P2 range also crosses the primes from
    // p_(c+act+1) to sqrt(hi) and counts the P2 leaves there
    double p2cost = 0;
    if (np2) {
      const double sq = std::min(std::sqrt((double)hi), (double)c->prime_limit);
      while (nsp_i < c->h_primes.size() && (double)c->h_primes[nsp_i] <= sq) nsp_i++;
      const double nsp = (double)nsp_i - (double)(cw + act);
      p2cost = 2.0 * np2 + std::max(nsp, 0.0) * 8.0 / T;
    }
    for (uint32_t r = 0; r < nrep; r++) {
      gp.tiles.push_back({gi, (uint16_t)nrep, (uint16_t)r, s});
      gp.tbucket.push_back(cost_bucket(r == 0 ? cost + p2cost : cost));
    }
  }
  for (uint32_t m = 0; m < g.n; m++) {
    QPlan &qp = qps[g.q[m]];
    qp.d.hard_nmrow = mrows[m];
    qp.mm = moff[m];
  }
}

// Host worker pool for the planner: threads are started once per process
// and woken per job (a planner call costs microseconds, so spawning
// threads per call would dominate it).  run(n, threads, f) runs f(k) for k
// in [0, n) on the caller and up to threads - 1 workers.  Jobs are
// serialised by a mutex: contexts on several host threads plan one at a time.
// The pool is never destroyed (its idle workers end with the process), and
// a forked child, which has no workers, runs every job inline.
class PlanPool {
public:
  static PlanPool &get()
  {
    static PlanPool *pool = new PlanPool();
    return *pool;
  }
  template <typename F>
  void run(size_t n, unsigned threads, F &f)
  {
    if (n == 0) return;
    threads = (unsigned)std::min<size_t>(std::max(1u, threads), n);
    threads = std::min<unsigned>(threads, (unsigned)workers_.size() + 1);
    if (getpid() != pid_) threads = 1;       // a forked child has no workers
    if (threads == 1) {
      for (size_t k = 0; k < n; k++) f(k);
      return;
    }
    std::lock_guard<std::mutex> job_lock(job_mu_);
    std::atomic<size_t> next{0};
    auto body = [&] {
      for (size_t k; (k = next.fetch_add(1)) < n;) f(k);
    };
    std::function<void()> job = body;
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &job;
      want_ = threads - 1;
      done_ = 0;
      gen_++;
    }
    cv_.notify_all();
    body();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return done_ == want_; });
    job_ = nullptr;
  }

private:
  PlanPool() : pid_(getpid())
  {
    unsigned hc = std::thread::hardware_concurrency();
    unsigned nw = std::min<unsigned>(hc ? hc : 4, 16);
    for (unsigned i = 1; i < nw; i++) workers_.emplace_back([this] { loop(); });
  }
  void loop()
  {
    uint64_t seen = 0;
    for (;;) {
      std::function<void()> *job;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (taken_gen_ != gen_) { taken_gen_ = gen_; taken_ = 0; }
        if (taken_ >= want_) continue;     // enough workers joined this job
        taken_++;
        job = job_;
      }
      (*job)();
      {
        std::lock_guard<std::mutex> lk(mu_);
        done_++;
      }
      done_cv_.notify_one();
    }
  }
  const pid_t pid_;
  std::vector<std::thread> workers_;
  std::mutex job_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  std::function<void()> *job_ = nullptr;
  uint64_t gen_ = 0, taken_gen_ = 0;
  unsigned want_ = 0, taken_ = 0, done_ = 0;
};

template <typename F>
static void parallel_for(size_t n, unsigned threads, F f)
{
  PlanPool::get().run(n, threads, f);
}

static unsigned plan_threads(size_t nq)
{
  if (nq < 8) return 1;
  unsigned hc = std::thread::hardware_concurrency();
  return std::min<unsigned>(hc ? hc : 4, 16);
}

static int build_plan(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, Plan &pl,
                      bool stream = false, const Caps *base = nullptr)
{
  CpxRange range("cpx:plan");
  pl.stream = stream;
  pl.base = base ? *base : Caps{};
  const Caps &B = pl.base;
  const size_t nq = hq.size();
  std::vector<QPlan> &qps = pl.qps;
  qps.assign(nq, QPlan());
  const unsigned nt = plan_threads(nq);
  parallel_for(nq, nt, [&](size_t i) { plan_query(c, hq[i], (uint32_t)(B.q + i), mask, qps[i]); });
  pl.groups.clear();
  make_groups(c, hq, qps, pl.groups);
  std::vector<GPlan> gps(pl.groups.size());
  parallel_for(pl.groups.size(), nt,
               [&](size_t g) { plan_group(c, pl.groups[g], (uint32_t)g, qps, gps[g]); });
  // place the shares: global offsets of every query's rows, parts and counters
  std::vector<PlanOff> &off = pl.off;
  off.resize(nq);
  size_t ntiles = 0;
  pl.q.resize(nq);
  for (size_t i = 0; i < nq; i++) {
    QPlan &qp = qps[i];
    QueryDev &d = qp.d;
    d.part_off = B.parts + pl.parts;
    d.stage_off = B.stages + pl.stage_total;
    d.hard_row0 = B.rows + pl.hard_rows;
    d.tm_off = B.tm + pl.hard_tm;
    d.p2_row0 = B.p2rows + pl.p2_rows;
    d.mrow_base = B.mrows + pl.hard_mrows;
    d.m_off = B.mm + pl.hard_mm;
    pl.parts += qp.parts;
    pl.stage_total += d.nstage;
    pl.hard_rows += d.hard_nseg;
    pl.hard_tm += qp.tm;
    pl.p2_rows += d.p2_nseg;
    pl.hard_mrows += d.hard_nmrow;
    pl.hard_mm += qp.mm;
    pl.hard_segments += d.hard_nseg;
    pl.p2_segments += d.p2_nseg;
    pl.q[i] = d;
    off[i] = {pl.nord, pl.ntriv, pl.neasy, pl.nfin, pl.np2d, pl.nrows};
    pl.nord += qp.ord.size(); pl.ntriv += qp.triv.size(); pl.neasy += qp.easy.size();
    pl.np2d += qp.p2d.size(); pl.nfin += qp.fin.size(); pl.nrows += qp.rowinfo.size();
  }
  for (const GPlan &gp : gps) ntiles += gp.tiles.size();
  // longest first: counting sort on the cost buckets (descending, stable in
  // group order within a bucket), O(tiles)
  std::vector<uint32_t> start(PLAN_NB + 1, 0);
  for (const GPlan &gp : gps)
    for (uint16_t b : gp.tbucket) start[PLAN_NB - 1 - b + 1]++;
  for (int b = 0; b < PLAN_NB; b++) start[b + 1] += start[b];
  pl.tiles.resize(ntiles);
  for (const GPlan &gp : gps)
    for (size_t k = 0; k < gp.tiles.size(); k++) pl.tiles[start[PLAN_NB - 1 - gp.tbucket[k]]++] = gp.tiles[k];
  // easy-leaf and P2-lookup tiles (one per descriptor, its global index in
  // `seg`) spread evenly through the segment tiles, so that CUs run gathers
  // beside sieve work
  auto spread = [&](uint32_t kind, uint64_t ne) {
    if (!ne) return;
    std::vector<TileDev> merged;
    merged.reserve(pl.tiles.size() + ne);
    const uint64_t ns = pl.tiles.size();
    uint64_t is = 0, ie = 0;
    while (is < ns || ie < ne) {
      if (ie < ne && (is == ns || ie * ns <= is * ne)) merged.push_back({kind, 1, 0, ie++});
      else merged.push_back(pl.tiles[is++]);
    }
    pl.tiles.swap(merged);
  };
  spread(CPX_TILE_EASY, pl.neasy);
  spread(CPX_TILE_P2, pl.np2d);
  if (stream) {
    spread(CPX_TILE_ORD, pl.nord);
    spread(CPX_TILE_TRIV, pl.ntriv);
    // a cohort of table lookups only still needs one main tile: its end
    // opens the combine phase
    if (pl.tiles.empty()) pl.tiles.push_back({CPX_TILE_NOP, 1, 0, 0});
  }
  // carry blocks longest first too (a block's time is its chain of row
  // chunks), so the grid-stride loop does not end on a few long ones
  std::fill(start.begin(), start.end(), 0u);
  for (const QPlan &qp : pl.qps)
    for (uint32_t r : qp.fin_rows) start[PLAN_NB - 1 - cost_bucket(1.0 + r) + 1]++;
  for (int b = 0; b < PLAN_NB; b++) start[b + 1] += start[b];
  pl.fin.resize(pl.nfin);
  for (const QPlan &qp : pl.qps)
    for (size_t k = 0; k < qp.fin.size(); k++)
      pl.fin[start[PLAN_NB - 1 - cost_bucket(1.0 + qp.fin_rows[k])]++] = qp.fin[k];
  return CPX_OK;
}

// ------------------------------------------------------------- the arena
static size_t al(size_t v) { return (v + 255) & ~(size_t)255; }

static Layout make_layout(const Caps &k)
{
  Layout L{};
  size_t o = 0;
  L.counters = o; o = al(o + 16 * 4);
  L.queries = o;  o = al(o + k.q * sizeof(QueryDev));
  L.stages = o;   o = al(o + k.stages * sizeof(StageDev));
  L.tiles = o;    o = al(o + k.tiles * sizeof(TileDev));
  L.groups = o;   o = al(o + k.groups * sizeof(GroupDev));
  L.ord = o;      o = al(o + k.ord * sizeof(BlockDesc));
  L.triv = o;     o = al(o + k.triv * sizeof(BlockDesc));
  L.easy = o;     o = al(o + k.easy * sizeof(BlockDesc));
  L.fin = o;      o = al(o + k.fin * sizeof(BlockDesc));
  L.p2d = o;      o = al(o + k.p2d * sizeof(BlockDesc));
  L.rowinfo = o;  o = al(o + k.rows * sizeof(HardRowInfo));
  L.plan_bytes = o;
  o = 0;
  L.hard_T = o;    o = al(o + k.tm * 4);
  L.hard_M = o;    o = al(o + k.mm * 4);
  L.hard_bact = o; o = al(o + k.rows * 4);
  L.hard_acc = o;  o = al(o + k.mrows * 8);
  L.hard_base = o; o = al(o + k.base * 8);
  L.p2_cnt = o;    o = al(o + k.p2rows * 4);
  L.p2_np = o;     o = al(o + k.p2rows * 4);
  L.p2_loc = o;    o = al(o + k.p2rows * 8);
  L.parts = o;     o = al(o + k.parts * sizeof(i128));
  L.results = o;   o = al(o + k.q * sizeof(ResultDev));
  L.out_bytes = o;
  return L;
}


static Caps needed_caps(const Plan &pl)
{
  Caps k{};
  k.q = pl.q.size(); k.stages = pl.stage_total; k.tiles = pl.tiles.size(); k.groups = pl.groups.size();
  k.ord = pl.nord; k.triv = pl.ntriv; k.easy = pl.neasy; k.fin = pl.nfin;
  k.p2d = pl.np2d;
  k.tm = pl.hard_tm; k.rows = pl.hard_rows; k.p2rows = pl.p2_rows; k.parts = pl.parts;
  k.mm = pl.hard_mm; k.mrows = pl.hard_mrows;
  return k;
}

// Arena capacities for `need` from the current ones `have`: every field of
// Caps at least its need (and 64 queries, 16 groups), grown by half when it
// grows.  Returns whether any field grew.
static bool grown_caps(const Caps &have, const Caps &need, Caps &k)
{
  static_assert(sizeof(Caps) % sizeof(uint64_t) == 0, "Caps is read as an array of counts");
  k = have;
  Caps want = need;
  want.q = std::max<uint64_t>(want.q, 64);
  want.groups = std::max<uint64_t>(want.groups, 16);
  uint64_t *kk = &k.q;
  const uint64_t *ww = &want.q;
  bool grow = false;
  for (size_t f = 0; f < sizeof(Caps) / sizeof(uint64_t); f++)
    if (ww[f] > kk[f]) {
      kk[f] = std::max<uint64_t>(ww[f], kk[f] + kk[f] / 2);
      grow = true;
    }
  return grow;
}

static int ensure_arena_caps(cpx_ctx *c, BatchSlot &S, const Caps &need, uint64_t budget)
{
  const Caps before = S.caps;
  Caps &k = S.caps;
  bool grow = grown_caps(before, need, k);
  if (!S.d_plan || !S.d_out) grow = true;
  if (!grow) return CPX_OK;
  Layout L = make_layout(k);
  if (budget && L.plan_bytes + L.out_bytes > budget) {
    // without the growth headroom, if that fits
    uint64_t *kk = &k.q;
    const uint64_t *nn = &need.q, *oo = &before.q;
    for (size_t f = 0; f < sizeof(Caps) / sizeof(uint64_t); f++) kk[f] = std::max(nn[f], oo[f]);
    k.q = std::max<uint64_t>(k.q, 64);
    k.groups = std::max<uint64_t>(k.groups, 16);
    L = make_layout(k);
  }
  // this slot's last batch was collected; the other slots' buffers are untouched
  if (S.d_plan) (void)hipFree(S.d_plan);
  if (S.d_out) (void)hipFree(S.d_out);
  if (S.h_plan) (void)hipHostFree(S.h_plan);
  if (S.h_res) (void)hipHostFree(S.h_res);
  S.d_plan = S.d_out = S.h_plan = S.h_res = nullptr;
  S.L = Layout{};
  S.arena_gen++;
  const Caps kk = k;
  auto fail = [&](int rc) { S.caps = Caps{}; return rc; };
  if (budget && L.plan_bytes + L.out_bytes > budget)
    return fail(ctx_fail(c, CPX_ENOMEM, "batch state exceeds the HBM budget"));
  hipError_t e;
  if ((e = cpx_dev_malloc(&S.d_plan, L.plan_bytes)) != hipSuccess) { S.d_plan = nullptr; (void)hipGetLastError(); return fail(hip_fail(c, e, "hipMalloc(plan)")); }
  if ((e = cpx_dev_malloc(&S.d_out, L.out_bytes)) != hipSuccess) { S.d_out = nullptr; (void)hipGetLastError(); return fail(hip_fail(c, e, "hipMalloc(out)")); }
  if ((e = hipHostMalloc(&S.h_plan, L.plan_bytes, hipHostMallocDefault)) != hipSuccess) { S.h_plan = nullptr; return fail(hip_fail(c, e, "hipHostMalloc")); }
  if ((e = hipHostMalloc(&S.h_res, kk.q * sizeof(ResultDev), hipHostMallocDefault)) != hipSuccess) { S.h_res = nullptr; return fail(hip_fail(c, e, "hipHostMalloc")); }
  S.L = L;
  return CPX_OK;
}

static int ensure_arena(cpx_ctx *c, BatchSlot &S, const Caps &need)
{
  const uint64_t budget = c->opts.hbm_budget;
  uint64_t other = c->table_bytes + slot_bytes(c);
  if (S.d_plan && S.d_out) other -= S.L.plan_bytes + S.L.out_bytes;
  if (budget && other >= budget)
    return ctx_fail(c, CPX_ENOMEM, "batch state exceeds the context HBM budget");
  return ensure_arena_caps(c, S, need, budget ? budget - other : 0);
}

uint64_t slot_bytes(const cpx_ctx *c)
{
  uint64_t b = 0;
  for (const BatchSlot &S : c->slot)
    if (S.d_plan && S.d_out) b += S.L.plan_bytes + S.L.out_bytes;
  return b;
}

void free_slots(cpx_ctx *c)
{
  for (BatchSlot &S : c->slot) {
    for (auto &g : S.gexec)
      if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
    if (S.d_plan) (void)hipFree(S.d_plan);
    if (S.d_out) (void)hipFree(S.d_out);
    if (S.h_plan) (void)hipHostFree(S.h_plan);
    if (S.h_res) (void)hipHostFree(S.h_res);
    S.d_plan = S.d_out = S.h_plan = S.h_res = nullptr;
    for (auto &ev : S.ev)
      if (ev) { (void)hipEventDestroy(ev); ev = nullptr; }
    S.busy = false;
  }
}

// The launch record of a plan laid out as L, its plan arrays at P and its
// output arrays at O (device addresses), over the context's tables.
static BatchDev batchdev_at(const cpx_ctx *c, const Layout &L, char *P, char *O)
{
  BatchDev b{};
  b.pitab = c->pitab.p; b.pitab_n = c->pitab_nent;
  b.primes = c->primes.p; b.nprimes = c->nprimes; b.prinv = c->prinv.p;
  b.mlpf = c->mlpf.p; b.mlpf_n = c->mlpf.cap;
  b.wheel = c->wheel.p; b.winfo = c->winfo.p;
  uint32_t *cnt = (uint32_t *)(P + L.counters);
  b.nqueries = cnt + 0; b.ntiles = cnt + 1; b.tile_head = cnt + 2;
  b.n_ord = cnt + 3; b.n_triv = cnt + 4; b.n_easy = cnt + 5; b.n_fin = cnt + 6; b.n_p2 = cnt + 7;
  b.queries = (QueryDev *)(P + L.queries);
  b.stages = (StageDev *)(P + L.stages);
  b.tiles = (TileDev *)(P + L.tiles);
  b.groups = (GroupDev *)(P + L.groups);
  b.ord_desc = (BlockDesc *)(P + L.ord);
  b.triv_desc = (BlockDesc *)(P + L.triv);
  b.easy_desc = (BlockDesc *)(P + L.easy);
  b.p2_desc = (BlockDesc *)(P + L.p2d);
  b.fin_desc = (BlockDesc *)(P + L.fin);
  b.rowinfo = (HardRowInfo *)(P + L.rowinfo);
  b.hard_T = (uint32_t *)(O + L.hard_T);
  b.hard_M = (int32_t *)(O + L.hard_M);
  b.hard_bact = (uint32_t *)(O + L.hard_bact);
  b.hard_acc = (int64_t *)(O + L.hard_acc);
  b.hard_base = nullptr;                 // (a run in passes points it at L.hard_base)
  b.pass_r0 = 0; b.pass_r1 = ~0ull; b.pass_toff = 0; b.pass_moff = 0; b.pass_add = 0; b.pass_pad = 0;
  b.p2_cnt = (uint32_t *)(O + L.p2_cnt);
  b.p2_np = (uint32_t *)(O + L.p2_np);
  b.p2_loc = (uint64_t *)(O + L.p2_loc);
  b.parts = (i128 *)(O + L.parts);
  b.results = (ResultDev *)(O + L.results);
  return b;
}

static BatchDev make_batchdev(cpx_ctx *c, const BatchSlot &S)
{
  return batchdev_at(c, S.L, (char *)S.d_plan, (char *)S.d_out);
}

// Writes the plan into a host image of the plan arena (layout L).
static void pack_plan(const Layout &L, const Plan &pl, char *H)
{
  uint32_t *cnt = (uint32_t *)(H + L.counters);
  memset(cnt, 0, 64);
  cnt[0] = (uint32_t)pl.q.size(); cnt[1] = (uint32_t)pl.tiles.size(); cnt[2] = 0;
  cnt[3] = (uint32_t)pl.nord; cnt[4] = (uint32_t)pl.ntriv;
  cnt[5] = (uint32_t)pl.neasy; cnt[6] = (uint32_t)pl.nfin;
  cnt[7] = (uint32_t)pl.np2d;
  auto put = [&](size_t off, const void *src, size_t bytes) {
    if (bytes) memcpy(H + off, src, bytes);
  };
  const Caps &B = pl.base;
  put(L.queries + B.q * sizeof(QueryDev), pl.q.data(), pl.q.size() * sizeof(QueryDev));
  put(L.fin + B.fin * sizeof(BlockDesc), pl.fin.data(), pl.fin.size() * sizeof(BlockDesc));
  // tiles and groups name groups, descriptors and queries: rebased to the
  // cohort's entries (a batch's base is zero)
  TileDev *td = (TileDev *)(H + L.tiles) + B.tiles;
  for (size_t k = 0; k < pl.tiles.size(); k++) {
    TileDev t = pl.tiles[k];
    switch (t.group) {
      case CPX_TILE_EASY: t.seg += B.easy; break;
      case CPX_TILE_P2: t.seg += B.p2d; break;
      case CPX_TILE_ORD: t.seg += B.ord; break;
      case CPX_TILE_TRIV: t.seg += B.triv; break;
      case CPX_TILE_NOP: break;
      default: t.group += (uint32_t)B.groups; break;
    }
    td[k] = t;
  }
  GroupDev *gd = (GroupDev *)(H + L.groups) + B.groups;
  for (size_t k = 0; k < pl.groups.size(); k++) {
    GroupDev g = pl.groups[k];
    for (uint32_t m = 0; m < g.n; m++) g.q[m] += (uint32_t)B.q;
    gd[k] = g;
  }
  // every query's descriptors and rows straight to their places
  parallel_for(pl.qps.size(), plan_threads(pl.qps.size()), [&](size_t i) {
    const QPlan &qp = pl.qps[i];
    const PlanOff &o = pl.off[i];
    put(L.ord + (B.ord + o.ord) * sizeof(BlockDesc), qp.ord.data(), qp.ord.size() * sizeof(BlockDesc));
    put(L.triv + (B.triv + o.triv) * sizeof(BlockDesc), qp.triv.data(), qp.triv.size() * sizeof(BlockDesc));
    put(L.easy + (B.easy + o.easy) * sizeof(BlockDesc), qp.easy.data(), qp.easy.size() * sizeof(BlockDesc));
    put(L.p2d + (B.p2d + o.p2d) * sizeof(BlockDesc), qp.p2d.data(), qp.p2d.size() * sizeof(BlockDesc));
    put(L.rowinfo + (B.rows + o.rows) * sizeof(HardRowInfo), qp.rowinfo.data(),
        qp.rowinfo.size() * sizeof(HardRowInfo));
  });
}

// The stage table of a streaming cohort, as k_setup writes it for a batch:
// {floor(x / p), 1 / p, ceil(2^64 / p), p, floor(y / p)} for the hard stages.
static void pack_stages(const cpx_ctx *c, const Layout &L, const Plan &pl, char *H)
{
  StageDev *st = (StageDev *)(H + L.stages);
  parallel_for(pl.q.size(), plan_threads(pl.q.size()), [&](size_t i) {
    const QueryDev &q = pl.q[i];
    const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
    for (uint32_t k = 0; k < q.nstage; k++) {
      const uint32_t p = c->h_primes[q.c + k];
      StageDev d;
      d.p = p;
      d.ylo = (uint32_t)(q.y / p);
      d.inv_p = 1.0 / (double)p;
      d.magic = ~0ull / p + 1;
      const u128 xq = x / p;
      d.xp = (uint64_t)xq;
      d.xph = (uint32_t)(xq >> 64);
      d.pad = 0;
      d.xp3 = ~0ull;
      if (q.omega && (uint64_t)p * p > q.y) {
        const u128 t = xq / ((uint64_t)p * p);
        d.xp3 = t < (u128)q.y ? (uint64_t)t : q.y;
      }
      st[q.stage_off + k] = d;
    }
  });
}

// Writes the plan into the slot's pinned staging and queues its copy on the
// upload stream; the batch stream waits for it (event ev[4]).
static int upload_plan(cpx_ctx *c, BatchSlot &S, const Plan &pl)
{
  CpxRange range("cpx:upload");
  const Layout &L = S.L;
  char *H = (char *)S.h_plan;
  pack_plan(L, pl, H);
  // copy only the used prefix of each array
  auto up = [&](size_t off, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    return hipMemcpyAsync((char *)S.d_plan + off, H + off, bytes, hipMemcpyHostToDevice, c->up_stream);
  };
  HIPCHK(up(L.counters, 64));
  HIPCHK(up(L.queries, pl.q.size() * sizeof(QueryDev)));
  HIPCHK(up(L.tiles, pl.tiles.size() * sizeof(TileDev)));
  HIPCHK(up(L.groups, pl.groups.size() * sizeof(GroupDev)));
  HIPCHK(up(L.ord, pl.nord * sizeof(BlockDesc)));
  HIPCHK(up(L.triv, pl.ntriv * sizeof(BlockDesc)));
  HIPCHK(up(L.easy, pl.neasy * sizeof(BlockDesc)));
  HIPCHK(up(L.p2d, pl.np2d * sizeof(BlockDesc)));
  HIPCHK(up(L.fin, pl.nfin * sizeof(BlockDesc)));
  HIPCHK(up(L.rowinfo, pl.nrows * sizeof(HardRowInfo)));
  HIPCHK(hipEventRecord(S.ev[4], c->up_stream));
  HIPCHK(hipStreamWaitEvent(c->stream, S.ev[4], 0));
  return CPX_OK;
}

// Occupancy query (and the >64 KiB LDS attribute) happen here, outside any
// graph capture; the grid is cached per shape.
static int seg_grid(cpx_ctx *c)
{
  static thread_local SegShape last{0, 0};
  static thread_local int last_nb = 0;
  if (last.threads != c->shape.threads || last.kwords != c->shape.kwords) {
    last_nb = seg_blocks_per_cu(c->shape);
    last = c->shape;
  }
  int nb = last_nb < 1 ? 1 : last_nb;
  return nb * c->cu_count;
}

static const uint32_t GRID_Q = 256, GRID_T = 2048;
// carry kernel: 128-thread blocks whose time is a latency chain, so many of
// them in flight (extra blocks of a small batch exit at once)
static const uint32_t GRID_FIN = 16384;

// The launch sequence in three parts; events between the parts time the
// persistent segment kernel on its own stream (also when replayed as graphs).
//   part 0: clear results + queue head, setup, ordinary, trivial
//   part 1: persistent segment kernel (hard + P2 segment tiles, easy-leaf
//           and P2-lookup tiles)
//   part 2: hard-leaf carry, combine
static int enqueue_part(cpx_ctx *c, const BatchDev &b, int part, uint32_t nres_words)
{
  CpxRange range(part == 0 ? "cpx:enqueue setup, ordinary, trivial"
                 : part == 1 ? "cpx:enqueue segment kernel" : "cpx:enqueue carry, combine");
  if (part == 0) {
    HIPCHK(launch_zero_u32((uint32_t *)b.results, nres_words, c->stream));
    HIPCHK(launch_zero_u32(b.tile_head, 1, c->stream));
    HIPCHK(launch_setup(b, GRID_Q, 0, c->stream));
    HIPCHK(launch_ordinary(b, GRID_T, c->stream));
    HIPCHK(launch_trivial(b, GRID_T, c->stream));
  } else if (part == 1) {
    HIPCHK(launch_segments(b, c->shape, (uint32_t)seg_grid(c), c->stream));
  } else {
    HIPCHK(launch_hard_finalize(b, GRID_FIN, c->stream));
    HIPCHK(launch_combine(b, GRID_Q, c->stream));
  }
  return CPX_OK;
}

static int ensure_graph(cpx_ctx *c, BatchSlot &S, const BatchDev &b, uint32_t nres_words)
{
  if (S.gexec[0] && S.graph_gen == c->generation && S.graph_arena_gen == S.arena_gen &&
      S.graph_shape.threads == c->shape.threads && S.graph_shape.kwords == c->shape.kwords)
    return CPX_OK;
  for (auto &g : S.gexec)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
  // kernel attributes and the occupancy queries must happen outside capture
  (void)seg_grid(c);
  for (int part = 0; part < 3; part++) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_part(c, b, part, nres_words);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (e2 != hipSuccess) return hip_fail(c, e2, "hipStreamEndCapture");
    hipError_t e3 = hipGraphInstantiate(&S.gexec[part], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e3 != hipSuccess) { S.gexec[part] = nullptr; return hip_fail(c, e3, "hipGraphInstantiate"); }
  }
  S.graph_gen = c->generation;
  S.graph_arena_gen = S.arena_gen;
  S.graph_shape = c->shape;
  return CPX_OK;
}

uint64_t plan_bytes_estimate(const cpx_ctx *c, const HostQuery &h)
{
  if (h.small) return 1024;
  const uint64_t SB = c->shape.seg_bits();
  // stages ~ pi(x^(1/4)) <= 1.26 n / ln n
  double q = (double)h.qrtx;
  double B = q < 3 ? 1 : 1.26 * q / std::log(q);
  double hseg = std::ceil((double)(h.z + 1) / (2.0 * SB));
  double p2seg = std::ceil((double)(h.zt + 1) / (2.0 * SB));
  // active (segment, stage) pairs: segment r >= 1 (lo = 2 SB r) has about
  // min(B, pi(sqrt(x / lo))) active stages; summed over log-spaced buckets
  const double X = (double)h.x;
  double pairs = B;
  for (double r = 1; r < hseg;) {
    const double r1 = std::min(hseg, std::max(r + 1, std::floor(r * 1.25)));
    const double v = std::sqrt(X / (2.0 * SB * std::sqrt(r * r1)));
    const double pv = v < 3 ? 0 : 1.26 * v / std::log(v);
    pairs += (r1 - r) * std::min(B, pv);
    r = r1;
  }
  // T and M per pair (M rows of replicated segments add some: a group's
  // leaf-dense segments give every member the group's replicas), rowinfo,
  // bact and the P2 rows, stage and carry descriptors, easy descriptors
  double bytes = pairs * 8 * 1.5 + hseg * 48 + p2seg * 16 + 4096 + B * 80 + (h.y / 8192.0 + 16) * 48;
  return (uint64_t)bytes;
}

// P2 by lookups in a resident pi table to x/(y+1) (instead of sieving the
// P2 segments) for x < 2^64 with x/(y+1) < P2TAB_LIMIT: chosen for a batch
// of at least P2TAB_MIN such queries, whose table build (about the cost of
// sieving one query's P2 range) is shared, or when the resident table
// already covers the queries.
static const uint64_t P2TAB_LIMIT = 1ull << 35;
static const size_t P2TAB_MIN = 8;

static void choose_p2tab(const cpx_ctx *c, std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode)
{
  size_t n = 0;
  auto eligible = [&](const HostQuery &h) {
    return !h.small && (mask & TERM_P2) && h.y < h.sqrtx && h.sqrtx < (1ull << 32) && h.zt < P2TAB_LIMIT;
  };
  for (const HostQuery &h : hq)
    if (eligible(h)) n++;
  const bool batch = p2mode == 0 && n >= P2TAB_MIN;
  for (HostQuery &h : hq)
    h.p2tab = p2mode != 2 && eligible(h) &&
              (batch || (!c->opts.fresh_tables && c->pitab_limit >= h.zt && c->prime_limit >= h.sqrtx));
}

// ------------------------------------------------ hard segments in passes
// A solve of one query whose hard rows (T per active stage of every
// segment, sum(mu) per replica and active stage) exceed the pass size runs
// its hard segments in passes: ranges of consecutive segments whose rows
// fit (a segment's rows are never split).  Every pass's tiles fill the
// persistent segment kernel on their own (the easy-leaf and P2-lookup
// tiles join the first), its rows land at the start of the T and M arrays,
// and its carry blocks continue each stage's survivor prefix from the
// earlier passes (BatchDev::hard_base).  The other outputs (replica sums,
// P2 rows, partial sums) are held whole.  The same multi-range carry is
// what src/sieve/S2_hard_host.cu:115-191 does per block of segments.
static const uint64_t PASS_BYTES_DEFAULT = 48ull << 30;

struct PassPlan {
  std::vector<uint64_t> s;                 // segment boundaries: pass k is [s[k], s[k+1])
  std::vector<uint64_t> toff, moff;        // its first T and M element
  std::vector<uint32_t> tstart, tcnt;      // its tiles
  uint64_t tm_max = 0, mm_max = 0;
  size_t n() const { return s.size() - 1; }
};

static bool plan_passes(const cpx_ctx *c, Plan &pl, PassPlan &pp)
{
  const uint64_t limit = c->opts.pass_bytes ? c->opts.pass_bytes : PASS_BYTES_DEFAULT;
  if (pl.qps.size() != 1 || !pl.qps[0].d.nstage) return false;
  const QPlan &qp = pl.qps[0];
  if (4 * (qp.tm + qp.mm) <= limit) return false;
  const std::vector<HardRowInfo> &R = qp.rowinfo;
  const uint64_t n = R.size();
  auto tend = [&](uint64_t s) { return R[s].toff + R[s].act; };
  auto mend = [&](uint64_t s) { return R[s].moff + (uint64_t)R[s].nrep * R[s].act; };
  pp = PassPlan{};
  pp.s.push_back(0);
  for (uint64_t s = 0, a = 0; s < n; s++) {
    if (s > a && 4 * (tend(s) - R[a].toff + mend(s) - R[a].moff) > limit) {
      pp.s.push_back(s);
      a = s;
    }
  }
  pp.s.push_back(n);
  const size_t np = pp.n();
  for (size_t k = 0; k < np; k++) {
    const uint64_t a = pp.s[k], b = pp.s[k + 1] - 1;
    pp.toff.push_back(R[a].toff);
    pp.moff.push_back(R[a].moff);
    pp.tm_max = std::max(pp.tm_max, tend(b) - R[a].toff);
    pp.mm_max = std::max(pp.mm_max, mend(b) - R[a].moff);
  }
  // tiles regrouped by pass, in their order within a pass (a stable counting sort)
  std::vector<uint32_t> pid(pl.tiles.size());
  pp.tcnt.assign(np, 0);
  for (size_t i = 0; i < pl.tiles.size(); i++) {
    const TileDev &t = pl.tiles[i];
    uint32_t k = 0;
    if (t.group < CPX_TILE_P2)
      k = (uint32_t)(std::upper_bound(pp.s.begin(), pp.s.end(), t.seg) - pp.s.begin()) - 1;
    pid[i] = k;
    pp.tcnt[k]++;
  }
  pp.tstart.assign(np, 0);
  for (size_t k = 1; k < np; k++) pp.tstart[k] = pp.tstart[k - 1] + pp.tcnt[k - 1];
  std::vector<uint32_t> at(pp.tstart);
  std::vector<TileDev> tiles(pl.tiles.size());
  for (size_t i = 0; i < pl.tiles.size(); i++) tiles[at[pid[i]]++] = pl.tiles[i];
  pl.tiles.swap(tiles);
  return true;
}

// The launches of a run in passes (no graphs): setup, ordinary and trivial
// once; per pass its tile count (staged in the slot's pinned plan image,
// so the host waits for the previous pass before it stages the next), the
// segment kernel over its tiles and the carry blocks over its rows; then the
// combine.
static int launch_passes(cpx_ctx *c, BatchSlot &S, const Plan &pl, const PassPlan &pp, const BatchDev &b0,
                         uint32_t nres_words)
{
  HIPCHK(hipEventRecord(S.ev[0], c->stream));
  int rc = enqueue_part(c, b0, 0, nres_words);
  if (rc) return rc;
  BatchDev b = b0;
  b.hard_base = (uint64_t *)((char *)S.d_out + S.L.hard_base);
  HIPCHK(launch_zero_u32((uint32_t *)b.hard_base, 2 * (size_t)pl.q[0].nstage, c->stream));
  uint32_t *cnt = (uint32_t *)((char *)S.h_plan + S.L.counters);
  HIPCHK(hipStreamSynchronize(c->up_stream));       // the plan upload has read the image
  HIPCHK(hipEventRecord(S.ev[2], c->stream));
  const bool progress = getenv("CPX_PROGRESS") != nullptr;
  const double t0 = now_ms();
  for (size_t k = 0; k < pp.n(); k++) {
    if (k) {
      HIPCHK(hipStreamSynchronize(c->stream));      // the previous pass's counter copy is done
      if (progress)
        fprintf(stderr, "cpx: pass %zu of %zu done (segments to %llu of %llu) at %.1f s\n", k, pp.n(),
                (unsigned long long)pp.s[k], (unsigned long long)pp.s.back(), (now_ms() - t0) / 1e3);
    }
    cnt[1] = pp.tcnt[k];
    cnt[2] = 0;                                      // queue head
    HIPCHK(hipMemcpyAsync((char *)S.d_plan + S.L.counters, cnt, 64, hipMemcpyHostToDevice, c->stream));
    b.tiles = b0.tiles + pp.tstart[k];
    b.pass_r0 = pp.s[k];
    b.pass_r1 = pp.s[k + 1];
    b.pass_toff = pp.toff[k];
    b.pass_moff = pp.moff[k];
    b.pass_add = k ? 1u : 0u;
    HIPCHK(launch_segments(b, c->shape, (uint32_t)seg_grid(c), c->stream));
    HIPCHK(launch_hard_finalize(b, GRID_FIN, c->stream));
  }
  HIPCHK(hipEventRecord(S.ev[3], c->stream));
  HIPCHK(launch_combine(b0, GRID_Q, c->stream));
  return CPX_OK;
}

static int submit_impl(cpx_ctx *c, const std::vector<HostQuery> &hq_in, uint32_t mask, uint32_t p2mode,
                       bool use_graph, uint32_t *slot_out)
{
  const double t0 = now_ms();
  uint32_t si = CPX_SLOTS;
  for (uint32_t i = 0; i < CPX_SLOTS; i++)
    if (!c->slot[i].busy) { si = i; break; }
  if (si == CPX_SLOTS) return ctx_fail(c, CPX_ESTATE, "every batch slot holds an uncollected batch");
  BatchSlot &S = c->slot[si];
  for (auto &ev : S.ev)
    if (!ev) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDefault));
  std::vector<HostQuery> hq(hq_in);
  choose_p2tab(c, hq, mask, p2mode);
  // tables first (planning needs the host prime list); a table that grows is
  // freed only after the stream has drained (dbuf_reserve), so a batch still
  // in flight in the other slot keeps reading valid tables
  uint64_t R = 0, P = 0, Y = 0;
  const uint64_t SB = c->shape.seg_bits();
  for (const HostQuery &h : hq) plan_needs(c, h, mask, SB, &R, &P, &Y);
  const double tt0 = now_ms();
  int rc = ensure_tables(c, R, P, Y);
  if (rc) return rc;
  uint32_t cmax = 1;
  for (const HostQuery &h : hq) cmax = std::max(cmax, h.c);
  if ((rc = build_wheels(c, cmax))) return rc;
  const double tt1 = now_ms();
  Plan pl;
  if ((rc = build_plan(c, hq, mask, pl))) return rc;
  PassPlan pp;
  const bool passes = plan_passes(c, pl, pp);
  Caps need = needed_caps(pl);
  if (passes) {
    need.tm = pp.tm_max;
    need.mm = pp.mm_max;
    need.base = pl.q[0].nstage;
  }
  if ((rc = ensure_arena(c, S, need))) return rc;
  const double tp = now_ms();
  if ((rc = upload_plan(c, S, pl))) return rc;
  BatchDev b = make_batchdev(c, S);
  const bool graph = use_graph && c->opts.use_graphs && !passes;
  // results region of the arena, in 32-bit words (capacity, so graphs stay valid)
  const uint32_t nres_words = (uint32_t)(S.caps.q * sizeof(ResultDev) / 4);
  if (graph && (rc = ensure_graph(c, S, b, nres_words))) return rc;
  if (passes) {
    if ((rc = launch_passes(c, S, pl, pp, b, nres_words))) return rc;
  } else {
    HIPCHK(hipEventRecord(S.ev[0], c->stream));
    for (int part = 0; part < 3; part++) {
      if (part == 1) HIPCHK(hipEventRecord(S.ev[2], c->stream));
      if (graph) HIPCHK(hipGraphLaunch(S.gexec[part], c->stream));
      else if ((rc = enqueue_part(c, b, part, nres_words))) return rc;
      if (part == 1) HIPCHK(hipEventRecord(S.ev[3], c->stream));
    }
  }
  HIPCHK(hipMemcpyAsync(S.h_res, (char *)S.d_out + S.L.results, hq.size() * sizeof(ResultDev),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipEventRecord(S.ev[1], c->stream));
  S.busy = true;
  S.ticket = c->next_ticket++;
  S.nq = hq.size();
  S.t_submit = t0;
  S.ms_tables = tt1 - tt0;
  S.ms_plan = (tp - t0) - (tt1 - tt0);
  S.tiles = pl.tiles.size();
  S.passes = passes ? (uint32_t)pp.n() : 1u;
  S.hard_segments = pl.hard_segments;
  S.p2_segments = pl.p2_segments;
  *slot_out = si;
  return CPX_OK;
}

int submit_queries(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode,
                   bool use_graph, uint32_t *slot)
{
  if (hq.empty()) return ctx_fail(c, CPX_EINVAL, "empty batch");
  const int rc = submit_impl(c, hq, mask, p2mode, use_graph, slot);
  if (rc) {
    // a failed submission may have queued part of its upload or launches:
    // let them drain before the slot is reused
    (void)hipStreamSynchronize(c->up_stream);
    (void)hipStreamSynchronize(c->stream);
  }
  return rc;
}

int collect_queries(cpx_ctx *c, uint32_t si, std::vector<ResultDev> &res, cpx_batch_stats *st)
{
  if (si >= CPX_SLOTS || !c->slot[si].busy) return ctx_fail(c, CPX_ESTATE, "no batch in this slot");
  BatchSlot &S = c->slot[si];
  S.busy = false;
  HIPCHK(hipEventSynchronize(S.ev[1]));
  res.assign((ResultDev *)S.h_res, (ResultDev *)S.h_res + S.nq);
  for (const ResultDev &r : res)
    if (r.status != 1) return ctx_fail(c, CPX_EDEVICE, "a query produced no result");
  if (st) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, S.ev[0], S.ev[1]);
    st->ms_device = ms;
    float s2 = 0;
    (void)hipEventElapsedTime(&s2, S.ev[2], S.ev[3]);
    st->ms_segments = s2;
    st->ms_tables = S.ms_tables;
    st->ms_plan = S.ms_plan;
    st->ms_total = now_ms() - S.t_submit;
    st->tiles = S.tiles;
    st->hard_segments = S.hard_segments;
    st->p2_segments = S.p2_segments;
    st->seg_threads = (uint32_t)c->shape.threads;
    st->seg_kwords = (uint32_t)c->shape.kwords;
    st->passes = S.passes;
  }
  return CPX_OK;
}

int run_queries(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode,
                std::vector<ResultDev> &res, cpx_batch_stats *st, bool use_graph)
{
  if (hq.empty()) { res.clear(); return CPX_OK; }
  uint32_t si = 0;
  const int rc = submit_queries(c, hq, mask, p2mode, use_graph, &si);
  if (rc) return rc;
  return collect_queries(c, si, res, st);
}

// ------------------------------------------- batches in a caller's region

int run_queries_at(cpx_ctx *c, std::vector<HostQuery> &hq, void *d_region, uint64_t region_bytes,
                   std::vector<ResultDev> &res)
{
  if (hq.empty()) { res.clear(); return CPX_OK; }
  // P2 by lookups only where the resident pi table already reaches
  choose_p2tab(c, hq, TERM_ALL, 1);
  uint64_t R = 0, P = 0, Y = 0;
  uint32_t cmax = 1;
  for (const HostQuery &h : hq) {
    plan_needs(c, h, TERM_ALL, c->shape.seg_bits(), &R, &P, &Y);
    cmax = std::max(cmax, h.c);
  }
  if (!tables_cover(c, R, P, Y) || !c->wheel.p || cmax > c->wheel_c)
    return ctx_fail(c, CPX_EAGAIN, "the resident tables do not cover this batch (an allocating call at this size "
                                   "builds them)");
  Plan pl;
  int rc = build_plan(c, hq, TERM_ALL, pl);
  if (rc) return rc;
  Caps k = needed_caps(pl);
  k.q = std::max<uint64_t>(k.q, 64);
  k.groups = std::max<uint64_t>(k.groups, 16);
  const Layout L = make_layout(k);
  const uint64_t plan_al = (L.plan_bytes + 255) & ~255ull;
  if (plan_al + L.out_bytes > region_bytes)
    return ctx_fail(c, CPX_ENOMEM, "batch state exceeds the caller's device region");
  char *P0 = (char *)d_region, *O0 = P0 + plan_al;
  // the plan from a host image, used prefix of each array (stream order: the
  // previous call's kernels on this stream have read the region before)
  std::vector<char> H(L.plan_bytes);
  pack_plan(L, pl, H.data());
  auto up = [&](size_t off, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    return hipMemcpyAsync(P0 + off, H.data() + off, bytes, hipMemcpyHostToDevice, c->stream);
  };
  HIPCHK(up(L.counters, 64));
  HIPCHK(up(L.queries, pl.q.size() * sizeof(QueryDev)));
  HIPCHK(up(L.tiles, pl.tiles.size() * sizeof(TileDev)));
  HIPCHK(up(L.groups, pl.groups.size() * sizeof(GroupDev)));
  HIPCHK(up(L.ord, pl.nord * sizeof(BlockDesc)));
  HIPCHK(up(L.triv, pl.ntriv * sizeof(BlockDesc)));
  HIPCHK(up(L.easy, pl.neasy * sizeof(BlockDesc)));
  HIPCHK(up(L.p2d, pl.np2d * sizeof(BlockDesc)));
  HIPCHK(up(L.fin, pl.nfin * sizeof(BlockDesc)));
  HIPCHK(up(L.rowinfo, pl.nrows * sizeof(HardRowInfo)));
  const BatchDev b = batchdev_at(c, L, P0, O0);
  const uint32_t nres_words = (uint32_t)(k.q * sizeof(ResultDev) / 4);
  for (int part = 0; part < 3; part++)
    if ((rc = enqueue_part(c, b, part, nres_words))) return rc;
  res.resize(hq.size());
  HIPCHK(hipMemcpyAsync(res.data(), O0 + L.results, hq.size() * sizeof(ResultDev), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (const ResultDev &r : res)
    if (r.status != 1) return ctx_fail(c, CPX_EDEVICE, "a query produced no result");
  return CPX_OK;
}

// ------------------------------------------------------ streaming cohorts

// Table needs of a cohort (P2 from the resident pi table where it reaches,
// as every streaming cohort takes it).
void stream_table_needs(const cpx_ctx *c, std::vector<HostQuery> &hq, uint64_t *R, uint64_t *P, uint64_t *Y,
                        bool p2_lookups)
{
  choose_p2tab(c, hq, TERM_ALL, p2_lookups ? 0 : 1);
  if (p2_lookups)
    for (HostQuery &h : hq)                    // every eligible query, even alone
      h.p2tab = !h.small && h.y < h.sqrtx && h.sqrtx < (1ull << 32) && h.zt < P2TAB_LIMIT;
  *R = *P = *Y = 0;
  for (const HostQuery &h : hq) plan_needs(c, h, TERM_ALL, c->shape.seg_bits(), R, P, Y);
}

int stream_plan(cpx_ctx *c, std::vector<HostQuery> &hq, BatchSlot &S, const Caps &cursor, bool grow,
                uint64_t budget, CohortInfo *ci, double mult)
{
  *ci = CohortInfo{};
  uint64_t R, P, Y;
  stream_table_needs(c, hq, &R, &P, &Y);
  uint32_t cmax = 1;
  for (const HostQuery &h : hq) cmax = std::max(cmax, h.c);
  if (!tables_cover(c, R, P, Y) || cmax > c->wheel_c) {
    ci->why = CohortInfo::TABLES;
    return CPX_EAGAIN;
  }
  const size_t NF = sizeof(Caps) / sizeof(uint64_t);
  // the cohort's entries start at the cursor, rounded up to CPX_STREAM_ALIGN
  Caps base = cursor;
  {
    uint64_t *b = &base.q;
    for (size_t f = 0; f < NF; f++) b[f] = (b[f] + CPX_STREAM_ALIGN - 1) / CPX_STREAM_ALIGN * CPX_STREAM_ALIGN;
  }
  Plan pl;
  int rc = build_plan(c, hq, TERM_ALL, pl, true, &base);
  if (rc) return rc;
  const Caps need = needed_caps(pl);
  const Layout Ln = make_layout(need);
  ci->need_bytes = Ln.plan_bytes + Ln.out_bytes;
  if (budget && ci->need_bytes > budget) {
    ci->why = CohortInfo::TOO_BIG;
    return ctx_fail(c, CPX_ENOMEM, "cohort state exceeds the server's HBM budget");
  }
  auto fits = [&](const Caps &b) {
    if (!S.d_plan || !S.d_out) return false;
    const uint64_t *bb = &b.q, *n = &need.q, *k = &S.caps.q;
    for (size_t f = 0; f < NF; f++)
      if (bb[f] + n[f] > k[f]) return false;
    return true;
  };
  if (!fits(base)) {
    if (!grow) {
      ci->why = CohortInfo::CAPS;
      return CPX_EAGAIN;
    }
    // an empty arena that holds mult cohorts like this one (within the budget)
    Caps want{};
    {
      uint64_t *w = &want.q;
      const uint64_t *n = &need.q, *k = &S.caps.q;
      for (size_t f = 0; f < NF; f++)
        w[f] = std::max(k[f], (uint64_t)((n[f] + CPX_STREAM_ALIGN) * std::max(1.0, mult)));
    }
    if ((rc = ensure_arena_caps(c, S, want, budget))) {
      ci->why = CohortInfo::TOO_BIG;
      return rc;
    }
    base = Caps{};
    if (!fits(base)) {
      ci->why = CohortInfo::TOO_BIG;
      return ctx_fail(c, CPX_ENOMEM, "cohort does not fit the arena");
    }
    pl = Plan();
    if ((rc = build_plan(c, hq, TERM_ALL, pl, true, &base))) return rc;
  }
  const Layout &L = S.L;
  char *H = (char *)S.h_plan;
  pack_plan(L, pl, H);
  pack_stages(c, L, pl, H);
  ci->nq = (uint32_t)pl.q.size();
  ci->n_main = (uint32_t)pl.tiles.size();
  ci->n_fin = (uint32_t)pl.nfin;
  ci->q_base = base.q;
  ci->tile_base = base.tiles;
  ci->fin_base = base.fin;
  {
    uint64_t *nx = &ci->next.q;
    const uint64_t *b = &base.q, *n = &need.q;
    for (size_t f = 0; f < NF; f++) nx[f] = b[f] + n[f];
  }
  // the cohort's entries of every plan array (the pinned image has the arena's layout)
  const size_t rg[][2] = {{L.queries + base.q * sizeof(QueryDev), pl.q.size() * sizeof(QueryDev)},
                          {L.stages + base.stages * sizeof(StageDev), pl.stage_total * sizeof(StageDev)},
                          {L.tiles + base.tiles * sizeof(TileDev), pl.tiles.size() * sizeof(TileDev)},
                          {L.groups + base.groups * sizeof(GroupDev), pl.groups.size() * sizeof(GroupDev)},
                          {L.ord + base.ord * sizeof(BlockDesc), pl.nord * sizeof(BlockDesc)},
                          {L.triv + base.triv * sizeof(BlockDesc), pl.ntriv * sizeof(BlockDesc)},
                          {L.easy + base.easy * sizeof(BlockDesc), pl.neasy * sizeof(BlockDesc)},
                          {L.fin + base.fin * sizeof(BlockDesc), pl.nfin * sizeof(BlockDesc)},
                          {L.p2d + base.p2d * sizeof(BlockDesc), pl.np2d * sizeof(BlockDesc)},
                          {L.rowinfo + base.rows * sizeof(HardRowInfo), pl.nrows * sizeof(HardRowInfo)}};
  ci->nranges = 0;
  ci->copy_bytes = 0;
  for (const auto &r : rg) {
    if (!r[1]) continue;
    ci->off[ci->nranges] = r[0];
    ci->len[ci->nranges] = r[1];
    ci->copy_bytes += r[1];
    ci->nranges++;
  }
  ci->hard_segments = pl.hard_segments;
  return CPX_OK;
}

int stream_grow_tables(cpx_ctx *c, std::vector<HostQuery> &hq, bool p2_lookups)
{
  uint64_t R, P, Y;
  stream_table_needs(c, hq, &R, &P, &Y, p2_lookups);
  uint32_t cmax = 1;
  for (const HostQuery &h : hq) cmax = std::max(cmax, h.c);
  int rc = ensure_tables(c, R, P, Y);
  if (!rc) rc = build_wheels(c, cmax);
  return rc;
}

int stream_need_bytes(cpx_ctx *c, std::vector<HostQuery> &hq, uint64_t *bytes)
{
  int rc = stream_grow_tables(c, hq);
  if (rc) return rc;
  Plan pl;
  if ((rc = build_plan(c, hq, TERM_ALL, pl, true))) return rc;
  const Layout L = make_layout(needed_caps(pl));
  *bytes = L.plan_bytes + L.out_bytes;
  return CPX_OK;
}

BatchDev stream_batchdev(cpx_ctx *c, BatchSlot &S, ResultDev *d_results)
{
  BatchDev b = make_batchdev(c, S);
  b.results = d_results;
  return b;
}
