/*
 * cudapix.h -- C ABI of the MI355X combinatorial pi(x) engine
 * (libcudapix_amd.so).
 *
 * The library replaces the pi(x) path of CUDApix: the entry points declared
 * in include/P2.cuh, ordinary.cuh, trivial.cuh, phi.cuh, pitable.cuh, S3.cuh
 * and V.cuh of the reference, plus the CUDASieve prime generation/counting and
 * the LPF/Moebius sieves they call.  Each declaration below names the
 * reference interface (file:line) it stands in for.  The reference API is
 * overloaded C++ with no status codes and global device state; this ABI is
 * plain C: every call takes a context, returns 0 (CPX_OK) or a negative
 * CPX_E* code, never exits or resets the device, and is safe to use from
 * different threads on different contexts.  128-bit values are passed as
 * {lo, hi} pairs (signed ones in two's complement).
 *
 * Semantics are the mathematical values of each term of
 *   pi(x) = S0 + (S1 + U) + V + S3 + pi(y) - 1 - P2      (src/Deleglise-Rivat/deleglise-rivat.cpp:77)
 * with every bound taken by exact integer roots; where the reference's code
 * is known to be off (its README.md:4,7) this library returns the exact value.
 * C++ callers can include cudapix.hpp for overloads with the reference's
 * names.
 */
#ifndef CUDAPIX_H
#define CUDAPIX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t lo, hi; } cpx_u128;   /* unsigned 128-bit */
typedef struct { uint64_t lo, hi; } cpx_i128;   /* signed 128-bit, two's complement */

typedef struct cpx_ctx cpx_ctx;
typedef struct cpx_pitable cpx_pitable;
typedef struct cpx_phi cpx_phi;
typedef struct cpx_server cpx_server;

enum {
  CPX_OK = 0,
  CPX_EINVAL = -1,     /* bad argument (null pointer, malformed x, c out of 1..8) */
  CPX_ENOMEM = -2,     /* device allocation failed or the HBM budget is exceeded */
  CPX_EDEVICE = -3,    /* a HIP call failed; cpx_ctx_last_error() has the text */
  CPX_ERANGE = -4,     /* x or y outside what the path supports */
  CPX_EAGAIN = -5,     /* server: no result ready yet */
  CPX_ECANCELED = -6,  /* server: request cancelled */
  CPX_ENOTFOUND = -7,  /* server: unknown ticket */
  CPX_ESTATE = -8      /* object used in a state that does not allow the call */
};

/* Context options.  Zero / negative fields take the defaults noted. */
typedef struct {
  int device;              /* HIP device ordinal (default 0) */
  uint32_t c;              /* primes in the phi wheel, 1..8 (default 6) */
  double alpha;            /* y = floor(alpha * cbrt(x)); <= 0: built-in alpha(x) model */
  uint32_t seg_threads;    /* segment-kernel threads per block (0: tuned) */
  uint32_t seg_kwords;     /* 32-bit words per thread per segment (0: tuned) */
  uint64_t hbm_budget;     /* bytes the context may hold on the device (0: 90% of free) */
  uint32_t fresh_tables;   /* 1: rebuild every table for each solve (one-at-a-time plain mode) */
  uint32_t use_graphs;     /* 1 (default): replay batch launches from a captured hipGraph */
  uint64_t pass_bytes;     /* a single solve whose hard-leaf rows (4 bytes per active stage per
                              segment, and per replica) exceed this runs its hard segments in
                              passes over segment ranges, each range's rows within it
                              (0: 48 GiB); results are the same.  Such a solve waits for
                              each pass before it queues the next, so cpx_pi_batch_submit
                              returns only when its last pass is queued */
} cpx_opts;

/* Per-solve breakdown (all exact). */
typedef struct {
  cpx_u128 pi;
  cpx_i128 ordinary;       /* S0 */
  cpx_i128 trivial;        /* S1 + U */
  cpx_i128 easy;           /* V */
  cpx_i128 hard;           /* S3 */
  cpx_i128 p2;             /* P2 */
  uint64_t y, z, c, pi_y;
  double alpha;            /* alpha actually used (y / cbrt(x)) */
  double ms;               /* wall time of the solve */
} cpx_terms;

void cpx_opts_default(cpx_opts *o);
int cpx_ctx_create(const cpx_opts *o, cpx_ctx **out);
int cpx_ctx_destroy(cpx_ctx *ctx);
const char *cpx_strerror(int code);
const char *cpx_ctx_last_error(const cpx_ctx *ctx);
/* y the engine would use for x under the context options. */
int cpx_choose_y(cpx_ctx *ctx, cpx_u128 x, uint64_t *y);
/* Device bytes currently held by the context (tables + batch buffers). */
uint64_t cpx_ctx_device_bytes(const cpx_ctx *ctx);

/* ---- full pi(x) -------------------------------------------------------- */
/* deleglise_rivat64::pi_deleglise_rivat (src/Deleglise-Rivat/deleglise-rivat.cpp:50-80);
 * y = 0 picks y from alpha. terms may be NULL. */
int cpx_pi_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, uint32_t c, uint64_t *out, cpx_terms *terms);
/* deleglise_rivat128::pi_deleglise_rivat (src/Deleglise-Rivat/deleglise-rivat.cpp:124-160) */
int cpx_pi_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, uint32_t c, cpx_u128 *out, cpx_terms *terms);

/* ---- terms (drop-in for the seven headers) ------------------------------ */
/* uint64_t P2(uint64_t x, uint64_t y)            include/P2.cuh:16 */
int cpx_p2_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, uint64_t *out);
/* uint128_t P2(uint128_t x, uint64_t y)          include/P2.cuh:17 */
int cpx_p2_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, cpx_u128 *out);
/* void divXbyY(uint128_t x, uint64_t *y, size_t len)  include/P2.cuh:21 (device pointer, in place) */
int cpx_div_x_by_y(cpx_ctx *ctx, cpx_u128 x, uint64_t *d_y, size_t len);
/* uint64_t Ordinary(uint64_t x, uint64_t y, uint16_t c)    include/ordinary.cuh:6 */
int cpx_ordinary_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, uint16_t c, int64_t *out);
/* uint128_t Ordinary(uint128_t x, uint64_t y, uint16_t c)  include/ordinary.cuh:7 */
int cpx_ordinary_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, uint16_t c, cpx_i128 *out);
/* uint64_t S2_trivial(uint64_t x, uint64_t y)    include/trivial.cuh:8  (S1 + U) */
int cpx_s2_trivial_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, uint64_t *out);
/* uint128_t S2_trivial(uint128_t x, uint64_t y)  include/trivial.cuh:12 */
int cpx_s2_trivial_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, cpx_u128 *out);
/* uint64_t S3(uint64_t x, uint64_t y, uint32_t c)  include/S3.cuh:3 -- the DR hard
 * leaves p_c < p <= x^(1/4) (the meaning deleglise_rivat64::S3 gives it,
 * src/Deleglise-Rivat/S3.cu:6-9); signed result. */
int cpx_s3_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, uint32_t c, int64_t *out);
int cpx_s3_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, uint32_t c, cpx_i128 *out);
/* int64_t V(uint64_t x, uint64_t y)               include/V.cuh:6 -- DR easy leaves */
int cpx_v_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, int64_t *out);
int cpx_v_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, cpx_i128 *out);
/* int64_t V_a / V_b(x, pi, p, q)                  include/V.cuh:7-8: V split at p = sqrt(x/y) */
int cpx_v_split_u64(cpx_ctx *ctx, uint64_t x, uint64_t y, int64_t *v_a, int64_t *v_b);
/* The reference's second form of the easy leaves: deleglise_rivat64::S2()
 * (src/Deleglise-Rivat/S2.cu:30-70) = A + sigma2 + sigma3 + sigma4 + 2 sigma5
 * + sigma6, with A the weighted sum of S2_a (S2.cu:72-110) and sigma1..sigma6
 * of src/Deleglise-Rivat/sigma.cu:25-112 (the sums taken with the offsets of
 * their 128-bit twins, sigma.cu:118-214, and sigma5 capped at pi(y)):
 *   A      = sum_{x^1/4 < p <= x^1/3} sum_{p < q <= sqrt(x/p)} w pi(x/pq),
 *            w = 2 if x/pq < y else 1
 *   sigma1 = C(pi(y) - pi(x^1/3), 2)   (part of the trivial leaves, not of S2)
 *   sigma2 = pi(y) (pi4 (pi4 - 3) / 2 - piz (piz - 3) / 2)
 *   sigma3 = pi3 (pi3 - 1)(2 pi3 - 1) / 6 - pi3 - pi4 (pi4 - 1)(2 pi4 - 1) / 6 + pi4
 *   sigma4 = pi(y) sum_{x^1/4 < p <= sqrt z} pi(x / (p y))
 *   sigma5 = sum_{sqrt z < p <= x^1/3} pi(min(y, x / p^2))
 *   sigma6 = - sum_{x^1/4 < p <= x^1/3} pi(sqrt(x / p))^2
 * (pi3 = pi(x^1/3), pi4 = pi(x^1/4), piz = pi(sqrt z), z = x / y, every root
 * an exact integer root).  s2 equals the direct V (cpx_v_*) exactly when no
 * prime lies in (x^1/4, x / y^2] (exact = 1); otherwise A also counts the
 * pairs y < q <= sqrt(x/p) and s2 exceeds V.  y = 0 picks y from alpha;
 * x / x^(1/4) must fit 64 bits (CPX_ERANGE otherwise). */
typedef struct {
  cpx_i128 a;
  cpx_i128 sigma[6];       /* sigma1 .. sigma6 */
  cpx_i128 s2;             /* A + sigma2 + sigma3 + sigma4 + 2 sigma5 + sigma6 */
  uint64_t y;              /* y used */
  uint32_t exact;          /* 1: s2 equals V for this x and y */
  uint32_t pad;
} cpx_sigma_terms;
int cpx_s2_sigma_u128(cpx_ctx *ctx, cpx_u128 x, uint64_t y, cpx_sigma_terms *out);

/* ---- primes, pi tables, phi (CUDASieve / pitable.cuh / phi.cuh) --------- */
/* CudaSieve::countPrimes(bottom, top): primes in [lo, hi) */
int cpx_count_primes(cpx_ctx *ctx, uint64_t lo, uint64_t hi, uint64_t *out);
/* CudaSieve::getDevicePrimes(bottom, top, len): device array of the primes in
 * [lo, hi) (uint64), freed with cpx_free. */
int cpx_get_device_primes(cpx_ctx *ctx, uint64_t lo, uint64_t hi, uint64_t **d_out, size_t *len);
/* CudaSieve::getDevicePrimes32(bottom, top, len)  (call site src/pitable.cu:13,
 * src/Deleglise-Rivat/omega12_host.cu:62): the same as uint32, hi <= 2^32. */
int cpx_get_device_primes32(cpx_ctx *ctx, uint64_t lo, uint64_t hi, uint32_t **d_out, size_t *len);
/* CudaSieve::getHostPrimes(bottom, top, len)  (call site src/S3.cu:27): the
 * primes of [lo, hi) into the caller's buffer out[cap]; *len = their number.
 * out == NULL only counts; cap < *len gives CPX_EINVAL (nothing written). */
int cpx_get_host_primes(cpx_ctx *ctx, uint64_t lo, uint64_t hi, uint64_t *out, size_t cap, size_t *len);
/* CudaSieve::genDeviceBitSieve(bottom, top)  (call sites
 * src/sieve/S2_hard_host.cu:65, src/Deleglise-Rivat/omega12_host.cu:63):
 * odd-only bit sieve of [lo, hi), lo a multiple of 64: bit i of word j stands
 * for n = lo + 64 j + 2 i + 1 and is CLEAR iff n is an odd prime in [lo, hi)
 * (the reader at src/sieve/S2_hard_device.cu:248-249 takes a clear bit as
 * prime); bits of 1, composites and n >= hi are set.  *nwords =
 * ceil((hi - lo) / 64); freed with cpx_free.  Ranges of any size to
 * 2^64 - 2^34 are sieved in 2^20-integer segments with sieving primes to
 * sqrt(hi) only (so are cpx_get_device_primes above 2^24, _primes32 and
 * cpx_get_host_primes, unless the pi table already reaches hi). */
int cpx_gen_device_bit_sieve(cpx_ctx *ctx, uint64_t lo, uint64_t hi, uint32_t **d_words, size_t *nwords);
/* uint32_t* get_d_piTable(uint32_t hi)   include/pitable.cuh:7: u32 pi(n) at
 * index (n+1)/2 for n < hi; freed with cpx_free. */
int cpx_get_d_pitable(cpx_ctx *ctx, uint32_t hi, uint32_t **d_out);
/* class PiTable  include/pitable.cuh:12-51: sliding window [base, base+range)
 * of u32 values pi(n) - pi_base at index (n + 1 - base) / 2 (base even). */
int cpx_pitable_create(cpx_ctx *ctx, uint64_t base, uint64_t bottom, uint64_t range, cpx_pitable **out);
int cpx_pitable_destroy(cpx_pitable *t);
int cpx_pitable_current(cpx_pitable *t, uint32_t **d_table);   /* getCurrent */
int cpx_pitable_next_up(cpx_pitable *t, uint32_t **d_table);   /* getNextUp */
int cpx_pitable_next_down(cpx_pitable *t, uint32_t **d_table); /* getNextDown */
int cpx_pitable_info(cpx_pitable *t, uint64_t *base, uint64_t *range, uint64_t *pi_base);
/* PiTable::reallocate after set_range  include/pitable.cuh:41 (a table for
 * the new range; filled on the next current/next_up/next_down). */
int cpx_pitable_set_range(cpx_pitable *t, uint64_t range);
/* PiTable::set_pi_base  include/pitable.cuh:34: pi(base) given by the caller
 * (not counted); next_up / next_down carry it by the windows' prime counts. */
int cpx_pitable_set_pi_base(cpx_pitable *t, uint64_t pi_base);
/* PiTable::getMaxRange  include/pitable.cuh:38: the range whose table takes
 * 80 % of the device's free bytes (free / 2.5, as src/pitable.cu:45-51). */
int cpx_pitable_max_range(cpx_ctx *ctx, uint64_t *range);
/* class Phi  include/phi.cuh:17-36: phi(x, a) for x < max from an LPF table. */
int cpx_phi_create(cpx_ctx *ctx, uint32_t max, cpx_phi **out);
/* Phi(max, p_max)  include/phi.cuh:28: the prime list ends at p_max (0: sqrt(max)). */
int cpx_phi_create_pmax(cpx_ctx *ctx, uint32_t max, uint32_t p_max, cpx_phi **out);
/* Phi::get_d_lpf  include/phi.cuh:23: the object's odd-only LPF array (index
 * i <-> k = 2i + 1 < max, lpf(1) = UINT32_MAX; len = its entries). */
int cpx_phi_lpf(cpx_phi *p, uint32_t **d_lpf, uint64_t *len);
/* Phi::get_d_primeList / get_h_primeList / get_primeListLength
 * include/phi.cuh:24-26: the primes from 41 (the 13th prime: the reference's
 * phi() takes the first twelve from a table of its own, src/phi.cu:11,44-45)
 * to p_max, on the device and on the host (both owned by the object). */
int cpx_phi_prime_list(cpx_phi *p, uint32_t **d_list, const uint32_t **h_list, uint32_t *len);
/* Phi::makeModCache  include/phi.cuh:32: out[k-1] = phi(p_1...p_k, k) for
 * k = 1 .. pi_max-1 (pi_max <= 16), counted on the device where the
 * primorial is below max and checked against prod (p_i - 1). */
int cpx_phi_mod_cache(cpx_phi *p, uint32_t pi_max, uint64_t *out);
int cpx_phi_destroy(cpx_phi *p);
int cpx_phi_eval(cpx_phi *p, uint32_t x, uint32_t a, uint64_t *out);          /* Phi::phi */
/* Phi::generateRange(num, a): device u32 array, entry i = #{odd k <= 2i+1 : lpf(k) > p_a},
 * freed with cpx_free. */
int cpx_phi_range(cpx_phi *p, uint32_t num, uint32_t a, uint32_t **d_out);
/* gen_d_lpf / gen_d_mu  include/sieve/lpf_mu.cuh:14-21: odd-only arrays for
 * n = bottom + 2i + 1 < top (bottom even); lpf(1) = UINT32_MAX; 32-bit
 * entries for top < 2^32. */
int cpx_gen_d_lpf(cpx_ctx *ctx, uint64_t bottom, uint64_t top, uint32_t **d_out);
int cpx_gen_d_mu(cpx_ctx *ctx, uint64_t bottom, uint64_t top, int8_t **d_out);
/* gen_d_lpf<uint64_t>  include/sieve/lpf_mu.cuh:14 (src/sieve/lpf_mu.cu:57-87):
 * 64-bit entries for windows anywhere below 2^64 - 2^34 (a prime n keeps its
 * whole value); cpx_gen_d_mu takes the same 64-bit windows. */
int cpx_gen_d_lpf64(cpx_ctx *ctx, uint64_t bottom, uint64_t top, uint64_t **d_out);
int cpx_free(cpx_ctx *ctx, void *d_ptr);
/* Device array to host memory (the host halves of gen_h_lpf / gen_h_mu,
 * include/sieve/lpf_mu.cuh:17,23, copy with it). */
int cpx_copy_to_host(cpx_ctx *ctx, void *dst, const void *d_src, size_t bytes);

/* ---- batched queries ----------------------------------------------------- */
typedef struct {
  uint32_t use_graph;      /* 1: replay a captured hipGraph of the batch launches */
  uint32_t tune;           /* 1: let the autotuner pick the segment shape first */
  uint32_t p2_table;       /* P2 of x < 2^64 by lookups in the resident pi table, extended to
                            * x/(y+1): 0 when at least 8 queries qualify (the table grows),
                            * 1 only where the resident table already reaches, 2 never
                            * (P2 is then counted in the sieve's P2 segments) */
} cpx_batch_opts;

typedef struct {
  double ms_total;         /* wall time of the call */
  double ms_plan;          /* host planning */
  double ms_tables;        /* table (re)builds */
  double ms_device;        /* device time of the batch (hipEvents) */
  double ms_segments;      /* device time of the persistent segment kernel */
  uint64_t tiles, hard_segments, p2_segments;
  uint32_t seg_threads, seg_kwords;
  uint32_t passes;         /* passes over segment ranges (1 unless a single query's hard rows
                              exceed cpx_opts.pass_bytes) */
  uint32_t reserved;
} cpx_batch_stats;

/* Many queries in one launch sequence over shared resident tables: every
 * hard-leaf and P2 segment of every query is a tile of one persistent kernel. */
int cpx_pi_batch(cpx_ctx *ctx, const cpx_u128 *xs, size_t n, cpx_u128 *out,
                 const cpx_batch_opts *opts, cpx_batch_stats *stats);
/* Same, also returning each query's term breakdown (terms may be NULL). */
int cpx_pi_batch_terms(cpx_ctx *ctx, const cpx_u128 *xs, size_t n, cpx_u128 *out,
                       cpx_terms *terms, const cpx_batch_opts *opts, cpx_batch_stats *stats);
/* cpx_pi_batch without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The batch's plan and
 * output arrays go to a static 128 MiB device array of the library (each GPU
 * the library runs on holds one until the process ends) instead of the
 * context's arena; nothing is allocated on the device or in pinned host
 * memory (host planning uses the process heap).  The context's resident
 * tables must already cover the batch -- an allocating call (cpx_pi_batch)
 * at the largest x builds them -- else CPX_EAGAIN; a batch whose state
 * exceeds 128 MiB (about one x near 8e16, or many smaller ones) gives
 * CPX_ENOMEM.  Waits for its results.  Calls on one GPU share its array and
 * are serialized by the library; the array keeps the last call's contents
 * (every call clears what must start at zero). */
int cpx_pi_batch_noalloc(cpx_ctx *ctx, const cpx_u128 *xs, size_t n, cpx_u128 *out);
/* Pipelined batches (no reference counterpart: the reference solves one x per
 * run).  cpx_pi_batch_submit plans the batch on the calling thread, queues its
 * upload and launches and returns a ticket without waiting, so the next batch
 * is planned and uploaded while this one runs; cpx_pi_batch_collect waits for
 * a ticket's batch and writes its n results to out (stats as cpx_pi_batch,
 * ms_total from submit to collect).  A context holds two uncollected batches
 * at most: a further submit (or a cpx_pi_batch / cpx_pi call while both are
 * in flight) returns CPX_ESTATE.  Tickets may be collected in any order; an
 * unknown or collected ticket gives CPX_ENOTFOUND.  opts->tune is honoured
 * only when no batch is in flight. */
int cpx_pi_batch_submit(cpx_ctx *ctx, const cpx_u128 *xs, size_t n, const cpx_batch_opts *opts,
                        uint64_t *ticket);
int cpx_pi_batch_collect(cpx_ctx *ctx, uint64_t ticket, cpx_u128 *out, cpx_batch_stats *stats);

/* ---- autotuner ----------------------------------------------------------- */
/* Time a calibration batch under each candidate and keep the fastest for this
 * context, in three passes: the segment-kernel shape (threads x words), a
 * factor on the built-in alpha(x) model (unless opts.alpha fixes alpha), and
 * the number of wheel primes c (6..8); every candidate must return the same
 * pi(x) for every query.  Also run by cpx_pi_batch when opts->tune is set. */
int cpx_autotune(cpx_ctx *ctx, const cpx_u128 *xs, size_t n, uint32_t *threads, uint32_t *kwords);
/* The context's current tuning (any pointer may be NULL). */
int cpx_get_tuning(cpx_ctx *ctx, uint32_t *threads, uint32_t *kwords, double *alpha_scale,
                   uint32_t *wheel_c);
/* Set a tuning (a supported shape, alpha factor in (0, 64), c in 1..8), e.g.
 * one cpx_autotune picked earlier, without re-timing.  No reference
 * counterpart (the reference fixes alpha and c per build). */
int cpx_set_tuning(cpx_ctx *ctx, uint32_t threads, uint32_t kwords, double alpha_scale,
                   uint32_t wheel_c);

/* ---- serving (continuous batching) ---------------------------------------
 * A server owns one executor context (the caller's tuning and tables are
 * snapshotted at creation) and one streaming launch at a time: a persistent
 * kernel that keeps taking work while the server publishes new cohorts of
 * admitted requests into it.  Requests that arrive while others run join the
 * running launch as soon as a cohort slot and the HBM budget allow; their
 * tiles run beside the earlier cohorts' tail.  In CPX_SERVER_BATCHES mode the
 * server instead runs each admitted FIFO prefix as a batch launch of its own.
 * The reference has no serving path (src/main.cpp:23-88 runs one term per
 * process). */
typedef struct {
  uint64_t hbm_budget;     /* device bytes for the cohort arena (0: 64 GiB): a cohort is admitted
                              when its planned state fits the space left in the running launch;
                              a request whose state alone exceeds the budget is rejected with
                              CPX_ENOMEM */
  uint32_t max_batch;      /* largest cohort (default 256) */
  double max_wait_ms;      /* stream mode: a pending request waits at most this long for the
                              device backlog to drain (so that it joins a larger cohort) before
                              it is published anyway (default 2); batch mode starts the next
                              batch as soon as the previous one ends */
  uint32_t slots;          /* cohorts in flight at once, 1..8 (default 4) */
  cpx_u128 max_x;          /* build the resident tables for x up to this at creation (0: grow on
                              demand, at a launch boundary) */
  uint32_t mode;           /* CPX_SERVER_STREAM (default): cohorts join one running streaming
                              launch; CPX_SERVER_BATCHES: each admitted FIFO prefix (at most
                              max_batch requests, admitted by their planned bytes) runs as a batch
                              launch of its own while later arrivals queue for the next one */
} cpx_server_opts;

#define CPX_SERVER_STREAM 0u
#define CPX_SERVER_BATCHES 1u

typedef struct {
  uint64_t ticket;
  cpx_u128 x;
  cpx_u128 pi;
  int status;              /* CPX_OK or a negative code */
  double t_arrive_ms, t_admit_ms, t_done_ms;   /* server clock */
  uint64_t launch;         /* the streaming launch that served it (from 1) */
  uint64_t cohort;         /* its cohort (from 1) */
  double dev_first_ms, dev_last_ms;   /* device clock: its cohort's first main tile began, last ended */
} cpx_result;

typedef struct {
  uint64_t submitted, completed, rejected, cancelled, batches;   /* batches: cohorts published */
  uint64_t max_inflight_bytes;   /* largest sum of the in-flight cohorts' arena needs */
  double p50_ms, p99_ms, mean_ms, qps;
  uint64_t launches;       /* streaming launches */
  uint64_t joined;         /* cohorts published into a launch that was already running another */
  uint64_t arena_bytes;    /* device bytes reserved by the cohort slots */
} cpx_server_stats;

int cpx_server_create(cpx_ctx *ctx, const cpx_server_opts *o, cpx_server **out);
int cpx_server_submit(cpx_server *s, cpx_u128 x, uint64_t *ticket);
int cpx_server_cancel(cpx_server *s, uint64_t ticket);
/* Up to max finished results; *n set to the number written. */
int cpx_server_poll(cpx_server *s, cpx_result *out, size_t max, size_t *n);
/* Block until every submitted request has finished. */
int cpx_server_drain(cpx_server *s);
int cpx_server_stats_get(cpx_server *s, cpx_server_stats *st);
int cpx_server_destroy(cpx_server *s);
/* Load generator: submit xs[i] at t_ms[i] milliseconds after the call (from
 * the calling thread); returns once the last request is queued. */
int cpx_server_replay(cpx_server *s, const cpx_u128 *xs, const double *t_ms, size_t n,
                      uint64_t *tickets);
/* Device bytes of one request's state planned alone as a server cohort, at
 * exactly the capacities its plan needs (the server admits a cohort by the
 * same measure against a slot's share of its budget; 0 on error).  Grows the
 * context's tables to cover x. */
uint64_t cpx_request_bytes(cpx_ctx *ctx, cpx_u128 x);

/* ---- host-side self checks (no GPU needed) -------------------------------- */
/* Checks the library's 128/64 division, small-quotient division and integer
 * roots against native __int128 on n pseudo-random cases; returns failures. */
int cpx_selftest_arith(uint64_t seed, uint64_t n);

/* ---- device self check (no reference counterpart) ------------------------ */
/* Runs the segment kernel's 64-lane register scan and sum on nwaves pseudo-
 * random wavefronts and compares each lane with a serial scan; *bad = number
 * of wavefronts that differ. */
int cpx_selftest_wave(cpx_ctx *ctx, uint64_t seed, uint32_t nwaves, uint32_t *bad);

/* ---- measured peaks (no reference counterpart) ---------------------------- */
/* LDS read bandwidth (bytes/s, conflict-free 16-byte reads on every CU) and
 * 32-bit integer VALU rate (lane operations/s) of the context's device, timed
 * with HIP events: the denominators of the roofline fractions. */
int cpx_probe_peaks(cpx_ctx *ctx, double *lds_bytes_per_s, double *int32_ops_per_s);

/* ---- test hook (no counterpart in the reference) ------------------------
 * The n-th device allocation the library makes from now on (tables, batch
 * arenas) fails as an exhausted device would (hipErrorOutOfMemory), so that
 * a test can check the clean CPX_ENOMEM and that nothing leaks; 0 disarms.
 * Process-wide; for tests only. */
int cpx_debug_fail_alloc(int n);

#ifdef __cplusplus
}
#endif
#endif /* CUDAPIX_H */
