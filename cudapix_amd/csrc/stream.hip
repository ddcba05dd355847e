// stream.hip -- the streaming (continuous-batching) persistent kernel.
//
// One launch serves every cohort the host publishes while it runs (layout
// and protocol in cpx_stream.hpp): main tiles are the batch kernel's tiles
// plus the ordinary and trivial descriptors, then carry units and combine
// units finish a cohort and write its results to pinned host memory.
// Queries that arrive while others run thus join the same launch, and their
// tiles interleave on the device with the earlier cohorts' tail.
//
// On this device an agent-scope fence writes back and invalidates the whole
// L2, which the segment tiles' table gathers depend on, so the kernel never
// issues one per unit: data moves through fresh addresses, written-through
// stores and device-coherent atomics, and order comes from waiting for a
// wave's outstanding memory operations (order_release / order_acquire), as
// rocPRIM's look-back scan does on this architecture.
// (The reference has no serving path; its terms run one host loop each,
// src/Deleglise-Rivat/deleglise-rivat.cpp:50-80.)
#include "segtiles.hpp"
#include "cpx_terms.hpp"
#include "cpx_stream.hpp"

enum : uint32_t { SW_NONE = 0, SW_COPY, SW_COMB, SW_FIN, SW_MAIN, SW_EXIT };

struct StreamWork {
  uint32_t kind, slot;
  uint64_t idx;
  uint32_t n;              // carry or combine units [idx, idx + n)
};

// Every memory operation of this wave has completed before the next one
// (vmcnt 0), without cache maintenance.
static __device__ __forceinline__ void order_release()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(/*vmcnt*/ 0 | (/*expcnt*/ 0x7 << 4) | (/*lgkmcnt*/ 0xf << 8));
}
static __device__ __forceinline__ void order_acquire()
{
  __builtin_amdgcn_s_waitcnt(/*vmcnt*/ 0 | (/*expcnt*/ 0x7 << 4) | (/*lgkmcnt*/ 0xf << 8));
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The published state and the plan image (pinned host memory): system-scope loads.
static __device__ __forceinline__ uint64_t ld_sys(const uint64_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
static __device__ __forceinline__ uint32_t ld_sys32(const uint32_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// Reports to pinned host memory: system-scope stores.
static __device__ __forceinline__ void st_sys(uint64_t *p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// Control state (device memory): device-scope atomics.
static __device__ __forceinline__ uint64_t ld_dev(const uint64_t *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void st_dev(uint64_t *p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ void st_dev32(uint32_t *p, uint32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ bool cas_dev(uint64_t *p, uint64_t expect, uint64_t want)
{
  return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ __forceinline__ uint32_t add_dev32(uint32_t *p, uint32_t v)
{
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// An error stops every block (device word) and is reported (a plain store to
// pinned host memory: no read-modify-write over the bus).
static __device__ __forceinline__ void set_error(const StreamArgs &a, uint32_t code)
{
  st_dev(&a.ctl->error, code);
  __hip_atomic_store(&a.mb->error, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The slot whose copy (which = 0) or main (1) range holds v.
static __device__ __forceinline__ bool slot_of(const StreamPub *pub, uint32_t nslots, int which, uint64_t v,
                                               uint32_t *slot, uint64_t *local)
{
  for (uint32_t j = 0; j < nslots; j++) {
    const uint64_t r = ld_sys(which ? &pub->rec[j].main_range : &pub->rec[j].copy_range);
    const uint64_t s = r >> CPX_STREAM_NBITS, n = r & ((1ull << CPX_STREAM_NBITS) - 1);
    if (v >= s && v < s + n) { *slot = j; *local = v - s; return true; }
  }
  return false;
}

// Claim a ready carry or combine unit of slot j: head = (seq << 32) | next,
// next = ~0 until the phase is ready.  The count (the slot's record) is read
// after the head; a completed cohort leaves its heads "not ready" before the
// host may reuse the record, so a head read before a record's reuse no longer
// matches at the compare-and-swap.  A claim at or past the count (read again
// after the claim) is no work.
// A claim takes a run of up to (count / (2 * grid), at most 64) units: a
// cohort's carry has thousands of units of a few microseconds each, and one
// compare-and-swap (plus its accounting) per unit made the claims, not the
// units, the cost of the phase.
static __device__ __forceinline__ bool claim_phase(uint64_t *head, const uint64_t *count, uint64_t *idx,
                                                   uint32_t *n)
{
  for (;;) {
    const uint64_t v = ld_dev(head);
    const uint32_t i = (uint32_t)v;
    if (i == 0xffffffffu) return false;
    order_acquire();
    const uint64_t cnt = ld_sys(count);
    if (i >= cnt) return false;
    uint64_t k = cnt / (2ull * gridDim.x);
    k = k < 1 ? 1 : (k > 64 ? 64 : k);
    if (k > cnt - i) k = cnt - i;
    if (cas_dev(head, v, v + k)) {
      order_acquire();
      if (i >= ld_sys(count)) return false;
      *idx = i;
      *n = (uint32_t)k;
      return true;
    }
  }
}

// Claim indices [idx, idx + n) below the published end from a FIFO head,
// n <= kmax (1 for copy units).
static __device__ __forceinline__ bool claim_fifo(uint64_t *head, uint64_t end, uint64_t *idx,
                                                  uint64_t kmax = 1, uint64_t *n = nullptr)
{
  for (;;) {
    const uint64_t h = ld_dev(head);
    if (h >= end) return false;
    const uint64_t k = min(kmax, end - h);
    if (cas_dev(head, h, h + k)) {
      *idx = h;
      if (n) *n = k;
      return true;
    }
  }
}

// Main tiles a block claims at once: every block contends for the one head
// (each claim a compare-and-swap at device scope), so while many tiles are
// left a claim takes a run of up to 16 consecutive tiles (1 near the end,
// where the longest-first order leaves the short tiles to balance the grid).
static __device__ __forceinline__ uint64_t main_claim_run(uint64_t left)
{
  const uint64_t k = left / (4ull * gridDim.x);
  return k < 1 ? 1 : (k > 16 ? 16 : k);
}

// Thread 0: the block's next unit of work, in priority order.  A claimed run
// of main tiles [v, v + n) comes back as its first tile, with *run_next and
// *run_end set to the rest (the caller takes those before looking again).
static __device__ __attribute__((noinline)) void stream_find(const StreamArgs &a, StreamWork &w,
                                                             uint64_t *run_next, uint64_t *run_end)
{
  StreamCtl *ctl = a.ctl;
  const StreamPub *pub = a.pub;
  uint64_t v;
  uint32_t j;
  w.kind = SW_NONE;
  if (ld_dev(&ctl->error)) { w.kind = SW_EXIT; return; }
  // the main end first: the host publishes a cohort's copy units before its
  // main tiles, so every cohort below `me` has its copy units below `ce`;
  // a block that finds no copy unit to claim below `ce` and takes a main
  // tile below `me` waits only for copies other blocks are already making
  const uint64_t me = ld_sys(&pub->main_end);
  order_acquire();
  const uint64_t ce = ld_sys(&pub->copy_end);
  order_acquire();
  if (claim_fifo(&ctl->copy_head, ce, &v)) {
    uint64_t loc;
    if (slot_of(pub, a.nslots, 0, v, &j, &loc)) { w = {SW_COPY, j, loc}; return; }
    set_error(a, 4u);
    w.kind = SW_EXIT;
    return;
  }
  // combine, then carry units of cohorts whose earlier phase is done (units
  // are numbered within the cohort; its entries start at the bases)
  uint32_t k;
  for (j = 0; j < a.nslots; j++)
    if (claim_phase(&ctl->comb_head[j], &pub->rec[j].n_comb, &v, &k)) {
      w = {SW_COMB, j, ld_sys(&pub->rec[j].q_base) + v, k};
      return;
    }
  for (j = 0; j < a.nslots; j++)
    if (claim_phase(&ctl->fin_head[j], &pub->rec[j].n_fin, &v, &k)) {
      w = {SW_FIN, j, ld_sys(&pub->rec[j].fin_base) + v, k};
      return;
    }
  uint64_t n = 1;
  const uint64_t h0 = ld_dev(&ctl->main_head);
  if (claim_fifo(&ctl->main_head, me, &v, main_claim_run(me > h0 ? me - h0 : 0), &n)) {
    st_sys(&a.mb->main_claimed, v + n);
    *run_next = v + 1;
    *run_end = v + n;
    uint64_t loc;
    if (slot_of(pub, a.nslots, 1, v, &j, &loc)) {
      w = {SW_MAIN, j, ld_sys(&pub->rec[j].tile_base) + loc};
      return;
    }
    set_error(a, 4u);
    w.kind = SW_EXIT;
    return;
  }
  // nothing to do: exit once the host closed the launch and every cohort it
  // published is complete
  if (ld_sys(&pub->close)) {
    order_acquire();
    const uint64_t np = ld_sys(&pub->cohorts);
    if (ld_dev(&ctl->completed) >= np) w.kind = SW_EXIT;
  }
}

// Copy unit u of slot j's cohort: a CPX_STREAM_COPY_BYTES piece of the
// concatenation of the cohort's entries in the plan arrays, from the pinned
// image (system-scope loads) into the arena at the same offsets, stored
// written-through.
template <int T>
static __device__ __attribute__((noinline)) void stream_copy(const StreamArgs &a, uint32_t j, uint64_t u)
{
  const StreamPub *pub = a.pub;
  const StreamRec &r = pub->rec[j];
  const uint64_t total = ld_sys(&r.copy_bytes);
  const char *src = (const char *)ld_sys((const uint64_t *)&pub->image);
  char *dst = (char *)ld_sys((const uint64_t *)&pub->arena);
  const uint32_t nr = (uint32_t)ld_sys(&r.nranges);
  const uint64_t b0 = u * CPX_STREAM_COPY_BYTES;
  const uint64_t b1 = min(total, b0 + CPX_STREAM_COPY_BYTES);
  const uint32_t t = threadIdx.x;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < nr && acc < b1; k++) {
    const uint64_t off = ld_sys(&r.off[k]), len = ld_sys(&r.len[k]);
    const uint64_t lo = max(b0, acc), hi = min(b1, acc + len);
    if (lo < hi) {
      // bytes [e0, e1) (multiples of 4): 4-byte moves up to an 8-byte
      // boundary and after the last one, 8-byte moves between.  The loads
      // stay system-scope (what makes the host's writes visible to a running
      // launch); eight of them per lane are issued before their stores, so
      // a lane has eight PCIe round trips in flight instead of one.
      const uint64_t e0 = off + (lo - acc), e1 = off + (hi - acc);
      const uint64_t a8 = min(e1, (e0 + 7) & ~7ull), z8 = max(a8, e1 & ~7ull);
      for (uint64_t o = e0 + 4ull * t; o < a8; o += 4ull * T)
        st_out((uint32_t *)(dst + o), ld_sys32((const uint32_t *)(src + o)));
      const uint64_t *s8 = (const uint64_t *)(src + a8);
      uint64_t *d8 = (uint64_t *)(dst + a8);
      const uint64_t nw = (z8 - a8) >> 3;
      constexpr uint32_t U = 8;
      uint64_t i = t;
      for (; i + (U - 1) * (uint64_t)T < nw; i += U * (uint64_t)T) {
        uint64_t v[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) v[u] = ld_sys(s8 + i + u * (uint64_t)T);
#pragma unroll
        for (uint32_t u = 0; u < U; u++) st_out(d8 + i + u * (uint64_t)T, v[u]);
      }
      for (; i < nw; i += T) st_out(d8 + i, ld_sys(s8 + i));
      for (uint64_t o = z8 + 4ull * t; o < e1; o += 4ull * T)
        st_out((uint32_t *)(dst + o), ld_sys32((const uint32_t *)(src + o)));
    }
    acc += len;
  }
}

// Carry, combine and term units: rarely run, kept out of line so that their
// registers do not add to the segment tiles' live ranges.
template <int T>
static __device__ __attribute__((noinline)) void stream_fin(const BatchDev &bd, uint64_t idx, uint32_t *lds)
{
  hard_fin_desc<T>(bd, bd.fin_desc[idx], *reinterpret_cast<FinShared<T> *>(lds));
}
template <int T>
static __device__ __attribute__((noinline)) void stream_comb(const BatchDev &bd, uint64_t idx, uint32_t *lds)
{
  combine_query<T>(bd, (uint32_t)idx, *reinterpret_cast<CombShared<T> *>(lds), true);
}
template <int T>
static __device__ __attribute__((noinline)) void stream_term(const BatchDev &bd, const TileDev &tile, uint32_t *lds)
{
  const bool ord = tile.group == CPX_TILE_ORD;
  const BlockDesc d = ord ? bd.ord_desc[tile.seg] : bd.triv_desc[tile.seg];
  const i128 v = ord ? ordinary_part<T>(bd, d, threadIdx.x) : trivial_part<T>(bd, d, threadIdx.x);
  const i128 s = block_sum_i128(v, reinterpret_cast<i128 *>(lds));
  if (threadIdx.x == 0) st_out(&bd.parts[bd.queries[d.query].part_off + d.part], s);
}

// One main tile: the batch kernel's segment, easy-leaf and P2-lookup tiles,
// and the ordinary and trivial descriptors.  Inlined: out of line, the
// tile code would see the LDS arrays through generic pointers (flat
// accesses that wait on global loads too) and read the arena record
// through memory instead of kernel arguments.
template <int T, int K>
static __device__ __forceinline__ void stream_main(const BatchDev &bd, const BatchDev *bd_mem, uint64_t ti,
                                                   uint32_t *lds, SegShared<T, K> &sh)
{
  const TileDev tile = bd.tiles[ti];
  if (tile.group == CPX_TILE_NOP) {
    return;
  } else if (tile.group == CPX_TILE_EASY) {
    easy_tile<T, K>(bd, (uint32_t)tile.seg, lds, sh);
  } else if (tile.group == CPX_TILE_P2) {
    p2tab_tile<T, K>(bd, (uint32_t)tile.seg, lds);
  } else if (tile.group == CPX_TILE_ORD || tile.group == CPX_TILE_TRIV) {
    stream_term<T>(*bd_mem, tile, lds);
  } else {
    hard_segment<T, K>(bd, tile, lds, sh);
  }
}

// Thread 0's copy of the record of the cohort whose main tiles it runs.
struct StreamCohort {
  uint64_t start = 0, count = 0;   // its main tiles [start, start + count)
  uint64_t tile_base = 0, seq = 0, n_main = 0;
  uint32_t slot = 0, ncopy = 0;
  __device__ bool holds(uint64_t v) const { return count && v >= start && v < start + count; }
  // the record of the slot whose main range holds v (system-scope loads)
  __device__ bool load(const StreamArgs &a, uint64_t v)
  {
    uint64_t loc;
    if (!slot_of(a.pub, a.nslots, 1, v, &slot, &loc)) {
      count = 0;
      return false;
    }
    const StreamRec &r = a.pub->rec[slot];
    const uint64_t mr = ld_sys(&r.main_range), cr = ld_sys(&r.copy_range);
    const uint64_t mask = (1ull << CPX_STREAM_NBITS) - 1;
    start = mr >> CPX_STREAM_NBITS;
    count = mr & mask;
    tile_base = ld_sys(&r.tile_base);
    seq = ld_sys(&r.seq);
    n_main = ld_sys(&r.n_main);
    ncopy = (uint32_t)(cr & mask);
    return holds(v);
  }
};

// Thread 0, after every wave's stores of the unit completed: counts the unit
// done and opens its cohort's next phase when it was the last.
static __device__ __attribute__((noinline)) void stream_account(const StreamArgs &a, const StreamWork &w,
                                                                uint64_t now, uint64_t seq_main,
                                                                uint64_t n_main)
{
  StreamCtl *ctl = a.ctl;
  StreamMailbox *mb = a.mb;
  const uint32_t j = w.slot;
  const StreamRec &r = a.pub->rec[j];
  if (w.kind == SW_COPY) {
    add_dev32(&ctl->copied[j], 1);
    return;
  }
  if (w.kind == SW_MAIN) {
    // (seq and n_main: the block's cached copy of the cohort's record)
    const uint64_t seq = seq_main;
    __hip_atomic_fetch_max(&ctl->t_last[j], now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t n = add_dev32(&ctl->main_done[j], 1) + 1;
    if (n == n_main) {
      order_release();
      if (ld_sys(&r.n_fin)) st_dev(&ctl->fin_head[j], seq << 32);
      else st_dev(&ctl->comb_head[j], seq << 32);
    }
    return;
  }
  const uint64_t seq = ld_sys(&r.seq);
  if (w.kind == SW_FIN) {
    const uint32_t n = add_dev32(&ctl->fin_done[j], w.n) + w.n;
    if (n == ld_sys(&r.n_fin)) {
      order_release();
      st_dev(&ctl->comb_head[j], seq << 32);
    }
  } else if (w.kind == SW_COMB) {
    const uint32_t n = add_dev32(&ctl->comb_done[j], w.n) + w.n;
    if (n == ld_sys(&r.n_comb)) {
      // the cohort is complete: report, reset the slot, then signal
      order_acquire();
      st_sys(&mb->t_first[j], ld_dev(&ctl->t_first[j]));
      st_sys(&mb->t_last[j], ld_dev(&ctl->t_last[j]));
      st_dev(&ctl->fin_head[j], ~0ull);
      st_dev(&ctl->comb_head[j], ~0ull);
      st_dev32(&ctl->copied[j], 0);
      st_dev32(&ctl->main_done[j], 0);
      st_dev32(&ctl->fin_done[j], 0);
      st_dev32(&ctl->comb_done[j], 0);
      st_dev(&ctl->t_first[j], 0);
      st_dev(&ctl->t_last[j], 0);
      __hip_atomic_fetch_add(&ctl->completed, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      order_release();
      st_sys(&mb->done_seq[j], seq);
    }
  }
}

template <int T, int K>
__global__ __launch_bounds__(T, 4) void k_stream(const StreamArgs a)
{
  extern __shared__ uint32_t lds[];
  __shared__ SegShared<T, K> sh;
  __shared__ StreamWork w;
  uint64_t idle_since = wall_clock64();
  uint32_t backoff = 1;
  // thread 0: the rest of the block's claimed run of main tiles; all of them
  // lie below the main end read when the run was claimed, so their cohorts'
  // copy units had all been claimed by then
  uint64_t run_next = 0, run_end = 0;
  // thread 0: the record of the cohort whose main tiles it takes (read once
  // per cohort and run, not per tile: a record does not change while the
  // cohort has tiles this block claimed, and main indices only grow, so a
  // range cached from a completed cohort matches no later index)
  StreamCohort co{};
  for (;;) {
    if (threadIdx.x == 0) {
      if (run_next < run_end) {
        if (!co.holds(run_next) && !co.load(a, run_next)) {
          set_error(a, 4u);
          w.kind = SW_EXIT;
        } else {
          w = {SW_MAIN, co.slot, co.tile_base + (run_next - co.start)};
          run_next++;
        }
      } else {
        stream_find(a, w, &run_next, &run_end);
        if (w.kind == SW_MAIN && !co.holds(run_next - 1) && !co.load(a, run_next - 1)) {
          set_error(a, 4u);
          w.kind = SW_EXIT;
        }
      }
      if (w.kind == SW_NONE && wall_clock64() - idle_since > a.idle_ticks) {
        // the host went quiet without closing: leave, flagging any cohort left behind
        if (ld_dev(&a.ctl->completed) < ld_sys(&a.pub->cohorts)) set_error(a, 2u);
        w.kind = SW_EXIT;
      }
      order_acquire();
    }
    __syncthreads();
    const StreamWork wk = w;
    if (wk.kind == SW_EXIT) break;
    if (wk.kind == SW_NONE) {
      // back off (1 .. 32 sleeps of ~64 * 127 clocks) before looking again
      for (uint32_t i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(127);
      backoff = backoff < 32 ? backoff * 2 : 32;
      __syncthreads();
      continue;
    }
    backoff = 1;
    const uint32_t j = __builtin_amdgcn_readfirstlane(wk.slot);
    if (wk.kind == SW_COPY) {
      stream_copy<T>(a, j, wk.idx);
    } else if (wk.kind == SW_MAIN) {
      if (threadIdx.x == 0) {
        // the cohort's copy units were claimed before this tile; wait for them
        const uint32_t ncopy = co.ncopy;
        const uint64_t t0 = wall_clock64();
        w.kind = SW_MAIN;
        while (__hip_atomic_load(&a.ctl->copied[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ncopy) {
          if (wall_clock64() - t0 > a.wait_ticks) {
            set_error(a, 1u);
            w.kind = SW_EXIT;
            break;
          }
          __builtin_amdgcn_s_sleep(8);
        }
        order_acquire();
        uint64_t zero = 0;
        __hip_atomic_compare_exchange_strong(&a.ctl->t_first[j], &zero, wall_clock64(), __ATOMIC_RELAXED,
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      if (w.kind == SW_EXIT) break;
      stream_main<T, K>(a.bd, a.bd_mem, wk.idx, lds, sh);
    } else if (wk.kind == SW_FIN) {
      for (uint32_t u = 0; u < wk.n; u++) {
        if (u) __syncthreads();
        stream_fin<T>(*a.bd_mem, wk.idx + u, lds);
      }
    } else {
      for (uint32_t u = 0; u < wk.n; u++) {
        if (u) __syncthreads();
        stream_comb<T>(*a.bd_mem, wk.idx + u, lds);
      }
    }
    // every wave's stores of the unit have completed before it is counted
    order_release();
    __syncthreads();
    if (threadIdx.x == 0) stream_account(a, {wk.kind, j, wk.idx, wk.n}, wall_clock64(), co.seq, co.n_main);
    idle_since = wall_clock64();
    __syncthreads();
  }
}

#define CPX_STREAM_SHAPES(X) X(512, 32) X(256, 64) X(256, 32)

bool stream_shape_supported(SegShape s)
{
#define X(TT, KK) if (s.threads == TT && s.kwords == KK) return true;
  CPX_STREAM_SHAPES(X)
#undef X
  return false;
}

int stream_blocks_per_cu(SegShape shape)
{
  size_t lds = seg_lds_bytes(shape);
  int nb = 0;
#define X(TT, KK)                                                                            \
  if (shape.threads == TT && shape.kwords == KK) {                                           \
    (void)hipFuncSetAttribute((const void *)k_stream<TT, KK>,                                \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);         \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_stream<TT, KK>, TT, lds) !=     \
        hipSuccess)                                                                          \
      nb = 0;                                                                                \
  }
  CPX_STREAM_SHAPES(X)
#undef X
  return nb;
}

hipError_t launch_stream(const StreamArgs &a, SegShape shape, uint32_t grid, hipStream_t s)
{
  size_t lds = seg_lds_bytes(shape);
#define X(TT, KK)                                                                            \
  if (shape.threads == TT && shape.kwords == KK) {                                           \
    hipLaunchKernelGGL((k_stream<TT, KK>), dim3(grid), dim3(TT), lds, s, a);                 \
    return hipGetLastError();                                                                \
  }
  CPX_STREAM_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}
