// cpx_internal.hpp -- device data layouts and kernel launchers of the
// MI355X pi(x) engine (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "cpx_arith.hpp"

// ---------------------------------------------------------------- layouts

// Compact pi table: one 8-byte entry per 64 integers [64w, 64w+64).
//   bits : bit i set <=> 64w + 2i + 1 is an odd prime
//   count: number of odd primes < 64w
// pi(n) = (n >= 2) + count + popc(bits & lowmask(((n & 63) + 1) >> 1)).
// (Replaces the u32-per-odd-n table of src/pitable.cu:172-218, which costs
// 2 bytes per integer; this one costs 1/8 byte per integer.)
struct PiEntry {
  uint32_t bits;
  uint32_t count;
};

__device__ __forceinline__ uint64_t pi_lookup(const PiEntry *__restrict__ t, uint64_t n)
{
  PiEntry e = t[n >> 6];
  uint32_t k = (uint32_t)(((n & 63) + 1) >> 1);
  return (uint64_t)(n >= 2) + e.count + cpx_popc32(e.bits & cpx_lowmask32(k));
}

// Wheel pattern of the first c primes over the odd-only bitmap: word w of
// any segment equals pat[w mod P] with P = 3*5*..*p_c words (period of the
// pattern replicated 32 times so that 32-bit words line up).  A second word
// per entry holds the exclusive prefix popcount, so phi(r, c) for r < W_c is
// prefix[k>>5] + popc(bits[k>>5] & lowmask(k & 31)), k = (r + 1) / 2.
struct WheelInfo {
  uint64_t off;        // offset of this c's pattern in the wheel array (entries)
  uint32_t period;     // P (words)
  uint32_t W;          // primorial W_c = 2 * 3 * ... * p_c
  uint32_t totient;    // Euler totient of W_c
  uint32_t pc;         // p_c
};

// Per-query record (device resident; host fills the static part).
struct QueryDev {
  uint64_t x_lo, x_hi;
  uint64_t y, z;          // z = floor(x / y)
  uint64_t sqrtx, cbrtx, qrtx, sqrtz;
  uint64_t p2_zt;         // floor(x / (y + 1)) : largest x/p over P2 primes
  uint64_t p2_e;          // even start of the P2 count range (sqrtx & ~1)
  uint64_t tm_off;        // first element of this query in hard_T
  uint32_t c, B;          // hard stages b = c+1 .. B (1-based prime indices)
  uint32_t nstage;        // B - c
  uint32_t a;             // pi(y)
  uint32_t pi_cbrt, pi_qrt, pi_sqrtz, small;   // small: answered by one pi-table lookup
  uint64_t hard_nseg, hard_row0;      // hard segments and first output row (T, bact, rowinfo)
  uint64_t m_off;                     // first element of this query in hard_M
  uint64_t mrow_base, hard_nmrow;     // first M-row (acc slot) and M-row count (>= segments)
  uint64_t p2_nseg, p2_row0;          // segments holding P2 leaves and their first output row
  uint64_t p2_seg0;                   // first of those segments (they run to the one holding zt)
  uint64_t stage_off;                 // first entry of this query in the stage table
  uint64_t part_off;                  // first slot in the partial-sum array
  uint32_t n_ord_parts, n_triv_parts, n_easy_parts, n_hard_parts;
  uint32_t n_p2_parts;                // P2 by table lookups: partial sums of pi(x/p) (after the hard ones)
  uint32_t p2tab;                     // P2 from table lookups instead of P2 segment rows
  uint32_t omega;                     // the hard leaves with x/(pq) < p^2 go to lookups (below)
  uint32_t pad_o;
  uint64_t pi_e;                      // written by setup: pi(p2_e)
  uint64_t pi_sqrtx;                  // written by setup: pi(sqrt x)
};

// One stage of the hard-leaf sieve: p_b, floor(x / p_b) and constants that
// turn the per-stage divisions into multiplications (computed once per query).
struct StageDev {
  uint64_t xp;         // floor(x / p), low 64 bits
  double inv_p;        // 1 / p rounded to f64
  uint64_t magic;      // ceil(2^64 / p): a mod p for 32-bit a (Lemire, Kaser, Kurz 2019)
  uint32_t p;
  uint32_t ylo;        // floor(y / p)
  uint32_t xph;        // floor(x / p) >> 64 (nonzero only for the first stages of x above 2^68)
  uint32_t pad;
  // the largest leaf prime q the segment sieve counts at a sparse stage (p^2 > y)
  // of a query with `omega`: min(y, floor(x / p^3)); the leaves with q above it
  // have x/(pq) < p^2 and go to pi-table lookups; ~0 otherwise
  uint64_t xp3;
};

// floor(a / d) from a precomputed inv ~ 1/d: exact when a < 2^53 or the
// quotient is < 2^50 (one signed remainder correction).
__host__ __device__ __forceinline__ uint64_t cpx_div_recip(uint64_t a, uint64_t d, double inv)
{
  uint64_t q = (uint64_t)((double)a * inv);
  int64_t r = (int64_t)(a - q * d);
  if (r < 0) q--;
  else if ((uint64_t)r >= d) q++;
  return q;
}

// floor(a / d) for a = ah * 2^64 + al (ah < 2^32) whose quotient is below
// 2^50 (checked by the caller with cpx_quot_below_2_50), 2 <= d < 2^62: the
// f64 estimate of a * inv is within one of the quotient, and the remainder
// al - q * d taken modulo 2^64 is the true remainder (it lies in (-d, 2d)).
__host__ __device__ __forceinline__ uint64_t cpx_div_recip_w(uint64_t al, uint32_t ah, uint64_t d, double inv)
{
  const double a = (double)al + (double)ah * 0x1p64;
  uint64_t q = (uint64_t)(a * inv);
  int64_t r = (int64_t)(al - q * d);
  if (r < 0) q--;
  else if ((uint64_t)r >= d) q++;
  return q;
}

// (ah * 2^64 + al) / d < 2^50, for d > 0.
__host__ __device__ __forceinline__ bool cpx_quot_below_2_50(uint64_t al, uint32_t ah, uint64_t d)
{
  return ((((uint64_t)ah) << 14) | (al >> 50)) < d;
}

// a mod p for 32-bit a with magic = ceil(2^64 / p).
__host__ __device__ __forceinline__ uint32_t cpx_fastmod32(uint32_t a, uint32_t p, uint64_t magic)
{
  uint64_t low = magic * a;
  return (uint32_t)(((u128)low * p) >> 64);
}

// Queries of a batch that share one sieve of each segment: a tile sieves its
// segment once for up to CPX_GROUP queries of similar x (same wheel), and
// evaluates every member's leaves, counters and P2 counts against it.
#ifndef CPX_GROUP
#define CPX_GROUP 8
#endif
struct GroupDev {
  uint32_t n;                   // members, sorted by x
  uint32_t q[CPX_GROUP];        // query indices
};

// Work tile of the persistent segment kernel: one segment of one group.  A
// leaf-heavy segment is split into nrep replica tiles: each sieves the whole
// segment but evaluates only the leaves with index = rep (mod nrep), writing
// its own sum(mu) rows.
struct TileDev {
  uint32_t group;        // CPX_TILE_EASY / CPX_TILE_P2: term tile of descriptor `seg`
  uint16_t nrep;
  uint16_t rep;
  uint64_t seg;
};

// TileDev.group of an easy-leaf (V) tile and of a P2-lookup tile (`seg` is
// the descriptor index): the persistent kernel interleaves them with the
// segment tiles, so their gathers overlap the sieve's ALU work.
#define CPX_TILE_EASY 0xffffffffu
#define CPX_TILE_P2 0xfffffffeu

// Per hard segment (query-relative): its first M-row, number of replicas,
// number of active stages `act` (p^2 <= x/lo) and where its counters live:
// T at hard_T[tm_off + toff + b], replica k's sum(mu) at
// hard_M[m_off + moff + k*act + b], for stages b < act.
struct HardRowInfo {
  uint32_t mrow;
  uint32_t nrep;
  uint32_t act;
  uint32_t pad;
  uint64_t toff;
  uint64_t moff;
};

// Hard-leaf stages per carry block (k_hard_fin): a row's T and sum(mu) of
// consecutive stages are adjacent, so a block that takes several stages
// reads them from the same cache lines.
#define CPX_FIN_STAGES 8

// Block descriptor of the grid-stride term kernels (ordinary, trivial, easy).
struct BlockDesc {
  uint32_t query;
  uint32_t part;         // partial-sum slot relative to the query's part_off
  uint64_t begin, end;   // item range [begin, end)
};

// Final per-query output (two's-complement 128-bit values as lo/hi pairs).
struct ResultDev {
  uint64_t pi[2];
  uint64_t ordinary[2], trivial[2], easy[2], hard[2], p2[2];
  uint64_t pi_y;
  uint64_t status;
};

// Everything a batch needs on the device; pointers are stable for a plan
// capacity so a captured graph stays valid across batches.
struct BatchDev {
  // shared resident tables
  const PiEntry *pitab;  uint64_t pitab_n;
  const uint32_t *primes; uint64_t nprimes;
  const double *prinv;                        // 1/p (nearest) of primes[i]
  const int32_t *mlpf;   uint64_t mlpf_n;     // mu(m)*lpf(m) for odd m, index m>>1
  const PiEntry *wheel;  const WheelInfo *winfo;
  // plan
  QueryDev *queries;     uint32_t *nqueries;
  StageDev *stages;
  TileDev *tiles;        uint32_t *ntiles;    uint32_t *tile_head;
  GroupDev *groups;
  BlockDesc *ord_desc;   uint32_t *n_ord;
  BlockDesc *triv_desc;  uint32_t *n_triv;
  BlockDesc *easy_desc;  uint32_t *n_easy;
  BlockDesc *fin_desc;   uint32_t *n_fin;     // hard carry blocks (begin = the stage)
  BlockDesc *p2_desc;    uint32_t *n_p2;      // P2 lookups (prime index ranges)
  HardRowInfo *rowinfo;                       // [row]
  // outputs
  uint32_t *hard_T;                           // per row, active stages: survivors before stage
  int32_t *hard_M;                            // per mrow, active stages: sum of mu of the leaves
  uint32_t *hard_bact;                        // [row]   active stages
  int64_t *hard_acc;                          // [mrow]  sum of mu * local phi
  uint32_t *p2_cnt;      uint32_t *p2_np;     uint64_t *p2_loc;   // [row]
  i128 *parts;           // partial sums
  ResultDev *results;
  // one query's hard segments in passes over segment ranges (its rows of
  // T and sum(mu) held for one range at a time): this launch's rows are the
  // segments [pass_r0, pass_r1) of the batch's only query, stored from
  // element pass_toff of its T and pass_moff of its M; hard_base holds each
  // stage's survivors in the earlier ranges (the carry blocks read and
  // advance it), and pass_add makes the carry blocks add to their partial
  // sums.  Outside passes: r0 = 0, r1 = ~0, offsets 0, hard_base null.
  uint64_t *hard_base;
  uint64_t pass_r0, pass_r1, pass_toff, pass_moff;
  uint32_t pass_add, pass_pad;
};

// Segment-kernel shape: T threads, K words (32 bits) per thread, so a segment
// holds 32*K*T odd numbers.
struct SegShape {
  int threads;
  int kwords;
  uint64_t seg_bits() const { return 32ull * kwords * threads; }
};

// ------------------------------------------------------------- launchers
// tables.hip
hipError_t launch_small_sieve(PiEntry *out, uint64_t nent, hipStream_t s);
hipError_t launch_pitab_segments(PiEntry *tab, uint64_t nent, const uint32_t *sprimes,
                                 uint32_t nsprimes, uint32_t *seg_tot, uint64_t *seg_base,
                                 hipStream_t s);
hipError_t launch_recip(const uint32_t *primes, double *rinv, uint64_t n, hipStream_t s);
hipError_t launch_extract_primes(const PiEntry *tab, uint64_t limit, uint32_t *primes, uint64_t cap,
                                 hipStream_t s);
hipError_t launch_mlpf(int32_t *mlpf, uint64_t nodd, const uint32_t *primes, uint32_t nsp,
                       hipStream_t s);
hipError_t launch_wheel(PiEntry *wheel, uint32_t period, uint32_t c, hipStream_t s);
hipError_t launch_expand_pitab_u32(const PiEntry *tab, uint32_t *out, uint64_t nodd_lo,
                                   uint64_t count, uint64_t base_pi, hipStream_t s);
hipError_t launch_phi_count(const int32_t *mlpf, uint64_t lo_idx, uint64_t hi_idx,
                            uint32_t pa, unsigned long long *out, hipStream_t s);
hipError_t launch_count_range(const PiEntry *tab, uint64_t lo, uint64_t hi,
                              unsigned long long *out, hipStream_t s);
hipError_t launch_count_primes_sieve(uint64_t lo, uint64_t hi, const uint32_t *sprimes,
                                     uint32_t nsp, unsigned long long *out, hipStream_t s);
hipError_t launch_div_x_by_y(uint64_t x_lo, uint64_t x_hi, uint64_t *d_y, size_t len,
                             hipStream_t s);

// terms.hip
hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s);
hipError_t launch_sigma(u128 x, uint64_t y, uint64_t sz, const uint32_t *primes, const PiEntry *pitab,
                        uint64_t j0, uint64_t np, i128 *part, hipStream_t s);
hipError_t launch_setup(BatchDev b, uint32_t maxq, uint32_t max_stage_total, hipStream_t s);
hipError_t launch_ordinary(BatchDev b, uint32_t grid, hipStream_t s);
hipError_t launch_trivial(BatchDev b, uint32_t grid, hipStream_t s);
hipError_t launch_hard_finalize(BatchDev b, uint32_t grid, hipStream_t s);
hipError_t launch_combine(BatchDev b, uint32_t maxq, hipStream_t s);

// segsieve.hip
hipError_t launch_segments(BatchDev b, SegShape shape, uint32_t grid, hipStream_t s);
hipError_t launch_selftest_wave(const uint32_t *in, uint32_t n, uint32_t *bad, hipStream_t s);
size_t seg_lds_bytes(SegShape shape);
bool seg_shape_supported(SegShape shape);
hipError_t launch_scan_entries(PiEntry *tab, uint64_t nent, uint32_t *tot, uint64_t *base,
                               hipStream_t s);
hipError_t launch_mu_lpf64_range(uint64_t *lpf, int8_t *mu, uint64_t ibase, uint64_t nodd,
                                 const uint32_t *primes, uint32_t nsp, hipStream_t s);
hipError_t launch_mu_lpf_split(uint32_t *lpf, int8_t *mu, uint64_t nodd, const uint32_t *primes,
                               uint32_t nsp, hipStream_t s);
hipError_t launch_phi_count_lpf(const uint32_t *lpf, uint64_t i0, uint64_t i1, uint32_t pa,
                                unsigned long long *out, hipStream_t s);
int seg_blocks_per_cu(SegShape shape);
hipError_t launch_mu_lpf_range(uint32_t *lpf, int8_t *mu, uint64_t ibase, uint64_t nodd,
                               const uint32_t *primes, uint32_t nsp, hipStream_t s);
hipError_t launch_phi_prefix(const uint32_t *lpf, uint64_t n, uint32_t pa, uint32_t *out,
                             hipStream_t s);
// odd-only prime bit sieve of [base, hi) (base a multiple of 64) in
// segments of 2^20 integers, with per-segment prime counts; extraction of
// the primes from such words (prime = 1) at per-segment output offsets
uint64_t range_sieve_segments(uint64_t base, uint64_t hi);
hipError_t launch_bit_sieve(uint64_t base, uint64_t lo, uint64_t hi, uint64_t nwords, const uint32_t *sp,
                            uint32_t nsp, bool composite_ones, uint32_t *words, uint32_t *seg_cnt,
                            hipStream_t s);
hipError_t launch_bit_extract(uint64_t base, uint64_t hi, const uint32_t *words, uint64_t nwords,
                              const uint64_t *seg_off, void *out, bool out32, hipStream_t s);
hipError_t launch_extract_range_u64(const PiEntry *tab, uint64_t lo, uint64_t hi, uint64_t *out,
                                    uint64_t cap, hipStream_t s);
