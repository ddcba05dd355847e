// engine.hpp -- host side of the MI355X pi(x) engine: context, resident
// tables, batch planner and launch sequence.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <mutex>
#include "cpx_internal.hpp"
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


struct cpx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  cpx_opts opts{};
  std::string err;
  int cu_count = 0;
  std::recursive_mutex mu;

  // ---- resident tables
  DBuf<PiEntry> small;              // pi table of [0, 2^18)
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

  // ---- batch buffers (plan arena on the device, pinned staging on the host)
  void *d_plan = nullptr;  size_t d_plan_cap = 0;
  void *h_plan = nullptr;  size_t h_plan_cap = 0;
  void *d_out = nullptr;   size_t d_out_cap = 0;
  void *h_res = nullptr;   size_t h_res_cap = 0;
  uint64_t generation = 0;                      // bumps on any reallocation

  // segment-kernel shape: 512 threads x 32 words (2^19-bit segments, two
  // blocks per CU) measured fastest for single queries to 1e16 on MI355X;
  // cpx_autotune may replace it for a batch
  SegShape shape{512, 32};
  bool shape_tuned = false;
  double alpha_scale = 1.0;                     // factor on the alpha model (autotuned)

  // captured graph of the batch launch sequence
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};   // pre-terms, segments, post
  uint64_t graph_gen = ~0ull;
  SegShape graph_shape{0, 0};

  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

// engine.hip
int ctx_fail(cpx_ctx *c, int code, const std::string &msg);
int hip_fail(cpx_ctx *c, hipError_t e, const char *what);
uint64_t choose_y_host(const cpx_ctx *c, u128 x);
int make_host_query(cpx_ctx *c, u128 x, uint64_t y, uint32_t cc, HostQuery *hq);
int ensure_tables(cpx_ctx *c, uint64_t R, uint64_t P, uint64_t Y);
// p2mode: cpx_batch_opts.p2_table
int run_queries(cpx_ctx *c, const std::vector<HostQuery> &hq, uint32_t mask, uint32_t p2mode,
                std::vector<ResultDev> &res, cpx_batch_stats *st, bool use_graph);
uint64_t host_pi(const cpx_ctx *c, uint64_t v);   // v <= prime_limit
uint64_t plan_bytes_estimate(const cpx_ctx *c, const HostQuery &hq);
void free_tables(cpx_ctx *c);

// Below this x the answer is one pi-table lookup; at and above it sqrt(x) >= 128,
// so the P2 range (which starts at sqrt(x) rounded down to a multiple of 64)
// lies above every wheel prime (<= 19).
#define CPX_SMALL_X 16384u
