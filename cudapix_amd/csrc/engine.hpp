// engine.hpp -- host side of the MI355X pi(x) engine: context, resident
// tables, batch planner and launch sequence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <mutex>
#include "cpx_internal.hpp"
#include "cpx_stream.hpp"
#include "../../include/cudapix.h"

// Grow-only device buffer.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  size_t bytes() const { return cap * sizeof(T); }
};

struct HostQuery {
  u128 x;
  uint64_t y, z, zt, sqrtx, cbrtx, qrtx, sqrtz;
  uint32_t c;
  bool small;               // answered from the pi table (x < CPX_SMALL_X)
  bool p2tab;               // P2 by pi-table lookups of x/p (the table covers x/(y+1))
};

// Which terms a plan evaluates.
enum : uint32_t {
  TERM_ORD = 1, TERM_TRIV = 2, TERM_EASY = 4, TERM_HARD = 8, TERM_P2 = 16,
  TERM_ALL = 31,
  TERM_EASY_A = 32,   // restrict V to x^(1/4) < p <= sqrt(x/y)   (V_a, src/V.cu:58-87)
  TERM_EASY_B = 64    // restrict V to sqrt(x/y) < p <= cbrt(x)   (V_b, src/V.cu:89-124)
};


// Fixed-capacity batch arena: array offsets depend only on capacities, so
// the pointers baked into a captured graph stay valid until a capacity grows.
struct Caps {
  uint64_t q, stages, tiles, groups, ord, triv, easy, fin, p2d, tm, rows, p2rows, parts, mm, mrows;
  uint64_t base;   // per-stage survivor totals of a run in passes (0 otherwise)
};

struct Layout {
  size_t counters, queries, stages, tiles, groups, ord, triv, easy, fin, p2d, rowinfo, plan_bytes;
  size_t hard_T, hard_M, hard_bact, hard_acc, hard_base, p2_cnt, p2_np, p2_loc, parts, results, out_bytes;
};

// Batch slots per context: one batch can be planned and uploaded while the
// previous one runs on the device.
#define CPX_SLOTS 2

// One batch's device arena, pinned staging, captured launch graphs and
// events, and the record of the batch it holds until collected.
struct BatchSlot {
  void *d_plan = nullptr;           // plan arena (device)
  void *h_plan = nullptr;           // its pinned staging copy
  void *d_out = nullptr;            // counters, partial sums and results (device)
  void *h_res = nullptr;            // pinned results
  Caps caps{};
  Layout L{};
  uint64_t arena_gen = 0;           // bumps when the arena is reallocated
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};   // pre-terms, segments, post
  uint64_t graph_gen = ~0ull, graph_arena_gen = ~0ull;     // table and arena generations captured
  SegShape graph_shape{0, 0};
  // batch start, batch end, segment kernel start and end, upload done
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  // the batch in flight
  bool busy = false;
  uint64_t ticket = 0;
  size_t nq = 0;
  double t_submit = 0, ms_tables = 0, ms_plan = 0;
  uint64_t tiles = 0, hard_segments = 0, p2_segments = 0;
  uint32_t passes = 1;
};

struct cpx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;     // tables and batch kernels
  hipStream_t up_stream = nullptr;  // plan uploads (overlap the previous batch's kernels)
  cpx_opts opts{};
  std::string err;
  int cu_count = 0;
  std::recursive_mutex mu;

  // ---- resident tables
  DBuf<PiEntry> small;              // pi table of [0, 2^19)
  DBuf<uint32_t> small_primes;
  uint32_t n_small_primes = 0;
  DBuf<PiEntry> pitab;
  uint64_t pitab_nent = 0, pitab_limit = 0;     // pi(n) valid for n <= pitab_limit
  DBuf<uint32_t> primes;
  DBuf<double> prinv;               // 1/p of primes[i], rounded to nearest
  uint64_t prime_limit = 0, nprimes = 0;
  std::vector<uint32_t> h_primes;
  DBuf<int32_t> mlpf;
  uint64_t mlpf_limit = 0;                      // odd m <= mlpf_limit
  DBuf<PiEntry> wheel;
  DBuf<WheelInfo> winfo;
  std::vector<WheelInfo> h_winfo;
  uint32_t wheel_c = 0;
  DBuf<uint32_t> scan_tot;
  DBuf<uint64_t> scan_base;
  uint64_t table_bytes = 0;

  // ---- batch slots
  BatchSlot slot[CPX_SLOTS];
  uint64_t next_ticket = 1;
  uint64_t generation = 0;                      // bumps when a table is reallocated

  // segment-kernel shape: 512 threads x 32 words (2^19-bit segments, two
  // blocks per CU) measured fastest for single queries to 1e16 on MI355X;
  // cpx_autotune may replace it for a batch
  SegShape shape{512, 32};
  bool shape_tuned = false;
  double alpha_scale = 1.0;                     // factor on the alpha model (autotuned)
};

// engine.hip
int ctx_fail(cpx_ctx *c, int code, const std::string &msg);
int hip_fail(cpx_ctx *c, hipError_t e, const char *what);
uint64_t choose_y_host(const cpx_ctx *c, u128 x);
int make_host_query(cpx_ctx *c, u128 x, uint64_t y, uint32_t cc, HostQuery *hq);
int ensure_tables(cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y);
// p2mode: cpx_batch_opts.p2_table.  submit_queries plans the queries into a
// free slot and queues their upload, launches and result copy without
// waiting (CPX_ESTATE when every slot holds an uncollected batch);
// collect_queries waits for a slot's batch and frees the slot; run_queries
// does both.
int submit_queries(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode,
                   bool use_graph, uint32_t *slot);
int collect_queries(cpx_ctx *c, uint32_t slot, std::vector<ResultDev> &res, cpx_batch_stats *st);
int run_queries(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode,
                std::vector<ResultDev> &res, cpx_batch_stats *st, bool use_graph);
// A batch with its plan and output arrays in a device region the caller
// provides (region_bytes long), nothing allocated on the device or in pinned
// host memory: the resident tables must already cover it (CPX_EAGAIN
// otherwise), and its state must fit the region (CPX_ENOMEM otherwise).
// Runs on the context's stream and waits for it.
int run_queries_at(cpx_ctx *c, std::vector<HostQuery> &hq, void *d_region, uint64_t region_bytes,
                   std::vector<ResultDev> &res);
// Frees every slot's buffers, graphs and events (context teardown).
void free_slots(cpx_ctx *c);
uint64_t slot_bytes(const cpx_ctx *c);
uint64_t host_pi(const cpx_ctx *c, uint64_t v);   // v <= prime_limit
uint64_t plan_bytes_estimate(const cpx_ctx *c, const HostQuery &hq);
void free_tables(cpx_ctx *c);

// ---- streaming cohorts (the server's continuous-batching launch)
struct CohortInfo {
  enum { OK = 0, TABLES = 1, CAPS = 2, TOO_BIG = 3 };
  int why = OK;                  // why stream_plan returned CPX_EAGAIN / CPX_ENOMEM
  uint32_t nq = 0, n_main = 0, n_fin = 0;
  uint32_t nranges = 0;          // used prefixes of the plan arena (same offsets in the pinned image)
  uint64_t off[12] = {}, len[12] = {};
  uint64_t copy_bytes = 0;
  uint64_t need_bytes = 0;       // plan + out bytes at exactly the capacities the plan needs
  uint64_t hard_segments = 0;
  uint64_t q_base = 0, tile_base = 0, fin_base = 0;   // its first query, tile, carry descriptor
  Caps next{};                   // the arena cursor after its entries
};
bool tables_cover(const cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y);
// p2_lookups: size the pi table for P2 by lookups of every eligible query.
void stream_table_needs(const cpx_ctx *c, std::vector<HostQuery> &hq, uint64_t *R, uint64_t *P, uint64_t *Y,
                        bool p2_lookups = false);
// Plans hq as one streaming cohort into the arena S at the cursor (its
// entries start there in every array, rounded up to CPX_STREAM_ALIGN; the
// pinned image S.h_plan mirrors the plan region): ordinary and trivial as
// tiles, stage table from the host, P2 by lookups where the resident table
// reaches.  CPX_EAGAIN: the tables do not cover the cohort (why = TABLES) or
// it does not fit behind the cursor and grow is false (CAPS); CPX_ENOMEM:
// its state exceeds the budget (TOO_BIG).  With grow (only while no launch
// uses the arena, cursor at zero) the arena is reallocated to hold mult
// cohorts like this one.
int stream_plan(cpx_ctx *c, std::vector<HostQuery> &hq, BatchSlot &S, const Caps &cursor, bool grow,
                uint64_t budget, CohortInfo *ci, double mult = 1.0);
int stream_grow_tables(cpx_ctx *c, std::vector<HostQuery> &hq, bool p2_lookups = false);
// Device bytes of hq's state as one streaming cohort at exactly the
// capacities its plan needs (grows the tables to cover it first).
int stream_need_bytes(cpx_ctx *c, std::vector<HostQuery> &hq, uint64_t *bytes);

BatchDev stream_batchdev(cpx_ctx *c, BatchSlot &S, ResultDev *d_results);

// Below this x the answer is one pi-table lookup; at and above it sqrt(x) >= 128,
// so the P2 range (which starts at sqrt(x) rounded down to a multiple of 64)
// lies above every wheel prime (<= 19).
#define CPX_SMALL_X 16384u
