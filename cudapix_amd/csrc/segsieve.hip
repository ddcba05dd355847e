// segsieve.hip -- the persistent segment kernel: hard-leaf phi sieve and the
// P2 segmented prime count, one LDS-resident bitmap segment per tile.
//
// Layout: a segment holds 32*K*T odd numbers lo+1, lo+3, ... as bits
// (1 = unsieved) in padded-linear LDS; thread t owns the contiguous bits
// [t*32K, (t+1)*32K) and keeps their survivor count in a register.
//
// Hard leaves (replaces S2glob::S2ctl + S2hardHost, src/sieve/S2_hard_device.cu:264-355,
// src/sieve/S2_hard_host.cu:115-191; DR split p <= x^(1/4), src/Deleglise-Rivat/S3.cu:8):
//   The segment starts as the wheel of the first c primes.  For each stage b
//   (p = p_b, c < b <= B, while p^2 <= x/lo):
//     L  leaves of p: the whole block strides the segment's m range
//        (x/(p*hi), x/(p*lo)] -- odd m with mu(m) != 0 and lpf(m) > p read
//        coalesced from the mu*lpf table when p^2 <= y, primes m from the
//        prime list when p^2 > y.  phi_local(n) for n = x/(m p) is
//          wave prefix (registers) + in-wave prefix of the owning thread
//          (LDS) + popcount of the owner's words up to n;
//        sum mu * phi_local and sum mu accumulate per thread.
//     -- barrier --
//     X  cross off the odd multiples of p in the thread's own bits (private
//        words, no atomics), update its count, wave-scan the counts and
//        publish in-wave prefixes and wave totals for the next stage.
//     -- barrier --
//   Per segment it writes T_b (survivors before stage b), sum_b(mu) and the
//   local sum; hard_finalize adds the carry sum_b prefix_b(seg) * sum_b(mu)
//   (the decomposition the reference does with scanVectorized/addMultiply,
//   src/sieve/S2_hard_device.cu:357-375).
//
// P2 (replaces the windowed upper_bound count of src/P2.cu:76-133):
//   segments of [sqrt(x) rounded down to 64, x/(y+1)] are fully sieved (wheel,
//   then the small primes shared by the block from offsets computed once per
//   segment, then one large prime per thread), then the block strides the P2
//   primes p (pi-table bits) whose x/p lands in the segment and counts primes
//   <= x/p the same way.  Per segment: prime count, number of p, sum of local
//   counts; the carry is applied in the combine kernel.
#include "cpx_device.hpp"

#define P2_MAX_COOP 512    // LDS slots for shared-prime offsets
#define HARD_SCH 256       // stages per chunk of LDS stage counters

template <int T, int K>
struct SegShared {
  uint32_t ipre[T];              // in-wave exclusive prefix of each thread's count
  uint32_t wtot[T / CPX_WAVE];   // wave totals
  int32_t wsum[T / CPX_WAVE];    // per-wave sums (P2)
  unsigned long long red_u[T / CPX_WAVE];
  long long red_s[T / CPX_WAVE];
  union {
    uint32_t offs[P2_MAX_COOP];  // P2: first multiple of each block-shared prime
    struct {
      uint32_t dec[HARD_SCH];    // hard: survivors removed by each stage of the chunk
      int32_t msum[HARD_SCH];    // hard: sum of mu of each stage's leaves (this replica)
    } st;
  } u;
  uint32_t tile;
};

// Publish the block scan of the per-thread counts: in-wave exclusive prefix
// per thread and one total per wave.  Caller syncs afterwards.
template <int T, int K>
__device__ __forceinline__ void publish_counts(uint32_t cnt, SegShared<T, K> &sh)
{
  uint32_t incl = wave_incl_scan(cnt);
  sh.ipre[threadIdx.x] = incl - cnt;
  if (lane_id() == CPX_WAVE - 1) sh.wtot[wave_id()] = incl;
}

// Number of survivors among the first `bit` bits of the segment
// (0 <= bit <= 32*K*T), given the wave prefixes wp[] held in registers.
template <int T, int K>
__device__ __forceinline__ uint32_t count_below(const uint32_t *lds, const SegShared<T, K> &sh,
                                                const uint32_t (&wp)[T / CPX_WAVE], uint32_t bit)
{
  constexpr uint32_t BPT = 32 * K;
  uint32_t owner = bit / BPT;
  uint32_t jb = bit - owner * BPT;
  if (owner >= (uint32_t)T) { owner = T - 1; jb = BPT; }
  uint32_t base = 0;
#pragma unroll
  for (uint32_t w = 0; w < (uint32_t)(T / CPX_WAVE); w++) base = (owner / CPX_WAVE == w) ? wp[w] : base;
  uint32_t c = base + sh.ipre[owner];
  const uint32_t wix = jb >> 5;
  const uint32_t a0 = owner * (K + 1);
  for (uint32_t j = 0; j < wix; j++) c += cpx_popc32(lds[a0 + j]);
  if (jb & 31u) c += cpx_popc32(lds[a0 + wix] & cpx_lowmask32(jb & 31u));
  return c;
}

template <int T, int K>
__device__ __forceinline__ uint32_t read_wave_prefix(const SegShared<T, K> &sh, uint32_t (&wp)[T / CPX_WAVE])
{
  uint32_t run = 0;
#pragma unroll
  for (uint32_t w = 0; w < (uint32_t)(T / CPX_WAVE); w++) { wp[w] = run; run += sh.wtot[w]; }
  return run;
}

// Write the counters of stages [c0, c1) of the current chunk: T_b (replica 0
// only), sum(mu) of this replica; then clear the chunk.  All threads call.
template <int T, int K>
__device__ __forceinline__ void hard_flush(SegShared<T, K> &sh, uint32_t c0, uint32_t c1,
                                           uint32_t &running, uint32_t *outT, int32_t *outM,
                                           bool write_T)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t k = c0; k < c1; k++) {
      if (write_T) outT[k] = running;
      running -= sh.u.st.dec[k - c0];
      outM[k] = sh.u.st.msum[k - c0];
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < HARD_SCH; k += T) { sh.u.st.dec[k] = 0; sh.u.st.msum[k] = 0; }
  __syncthreads();
}

template <int T, int K>
__device__ __forceinline__ void hard_segment(const BatchDev &bd, const QueryDev &q,
                                             const TileDev &tile, uint32_t *lds, SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t seg = tile.seg;
  const uint64_t lo = seg * 2 * SB;
  const uint64_t hi = lo + 2 * SB;
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const uint64_t y = q.y;
  const WheelInfo wi = bd.winfo[q.c];
  const uint64_t row = q.hard_row0 + seg;
  const HardRowInfo ri = bd.rowinfo[row];
  const uint32_t rep = tile.rep, nrep = ri.nrep;
  const uint64_t mrow = (uint64_t)ri.mrow + rep;
  uint32_t *outT = bd.hard_T + q.tm_off + seg * q.nstage;
  int32_t *outM = bd.hard_M + q.m_off + mrow * q.nstage;
  const StageDev *st = bd.stages + q.stage_off;
  const int32_t *__restrict__ mlpf = bd.mlpf;

  seg_fill_wheel(lds, K, T, lo, bd.wheel + wi.off, wi.period);
  for (uint32_t k = t; k < HARD_SCH; k += T) { sh.u.st.dec[k] = 0; sh.u.st.msum[k] = 0; }
  __syncthreads();
  uint32_t cnt = 0;
#pragma unroll 8
  for (uint32_t j = 0; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  {
    uint32_t ws = wave_sum(cnt);
    if (lane == 0) sh.wtot[wv] = ws;
  }
  __syncthreads();
  uint32_t running = 0;                                 // survivors before the next stage
  for (uint32_t w = 0; w < NW; w++) running += sh.wtot[w];
  __syncthreads();

  // stage b is active while p_b^2 <= x / lo (all leaves of p_b have n < x/p_b^2)
  uint64_t xlo = ~0ull;
  if (lo) {
    u128 qq = x / lo;
    xlo = (qq >> 64) ? ~0ull : (uint64_t)qq;
  }
  const double inv_lo = lo ? 1.0 / (double)lo : 0.0, inv_hi = 1.0 / (double)hi;

  int64_t acc = 0;
  uint32_t s = 0, chunk0 = 0;
  StageDev cur = st[0];
  for (; s < q.nstage; s++) {
    const uint32_t p = cur.p;
    const uint64_t xp = cur.xp;
    if ((uint64_t)p * p > xlo) break;
    const StageDev nxt = (s + 1 < q.nstage) ? st[s + 1] : cur;

    // m range of this segment's leaves of p (uniform over the block)
    const bool sparse = (uint64_t)p * p > y;
    uint64_t mhi = y;
    if (lo != 0 && (xp >> 50) < lo) {
      uint64_t v = cpx_div_recip(xp, lo, inv_lo);
      if (v < mhi) mhi = v;
    }
    uint64_t mlo = ~0ull;
    if ((xp >> 50) < hi) mlo = cpx_div_recip(xp, hi, inv_hi);
    if (mlo <= mhi) {
      if (mlo < cur.ylo) mlo = cur.ylo;
      if (sparse && mlo < p) mlo = p;
    }

    if (mhi > mlo) {
      // ---- leaves of p (state: crossed by p_1 .. p_{b-1}); needs every
      //      thread's words and counts, so the block meets here
      publish_counts<T, K>(cnt, sh);
      __syncthreads();
      uint32_t wp[NW];
      (void)read_wave_prefix<T, K>(sh, wp);
      int32_t smu = 0;
      const uint64_t step = (uint64_t)T * nrep;
      if (!sparse) {
        const uint64_t mtop = mhi - ((mhi & 1) ^ 1);            // largest odd <= mhi
        const uint64_t ncand = mtop > mlo ? (mtop - mlo + 1) / 2 : 0;
#pragma unroll 4
        for (uint64_t i = (uint64_t)rep * T + t; i < ncand; i += step) {
          const uint64_t m = mtop - 2 * i;
          const int32_t v = mlpf[m >> 1];
          const uint32_t av = (uint32_t)(v < 0 ? -v : v);
          if (av <= p) continue;                               // mu(m) = 0 or lpf(m) <= p
          const uint64_t n = cpx_div_smallq(xp, m);           // lo <= n < hi
          const uint32_t c = count_below<T, K>(lds, sh, wp, (uint32_t)((n - lo + 1) >> 1));
          if (v > 0) { acc += c; smu++; } else { acc -= c; smu--; }
        }
      } else {
        const uint64_t ih = pi_lookup(bd.pitab, mhi), il = pi_lookup(bd.pitab, mlo);
#pragma unroll 4
        for (uint64_t k = il + (uint64_t)rep * T + t; k < ih; k += step) {
          const uint64_t m = bd.primes[k];                     // prime, mu = -1
          const uint64_t n = cpx_div_smallq(xp, m);
          acc -= count_below<T, K>(lds, sh, wp, (uint32_t)((n - lo + 1) >> 1));
          smu--;
        }
      }
      if (smu) atomicAdd(&sh.u.st.msum[s - chunk0], smu);
      __syncthreads();                                          // readers done before cross-off
    }

    // ---- cross off the odd multiples of p in this thread's own bits
    {
      // first odd multiple of p at or above lo + 1, as a bit offset (< p)
      const uint64_t start = lo + 1;
      const uint64_t rr = start - cpx_div_recip(start, p, cur.inv_p) * p;
      uint64_t mm = start + (rr ? p - rr : 0);
      if (!(mm & 1)) mm += p;
      const uint32_t oseg = (uint32_t)((mm - start) >> 1);
      const uint32_t tb = cpx_fastmod32(t * BPT, p, cur.magic);
      uint32_t j = oseg >= tb ? oseg - tb : oseg + p - tb;
      const uint32_t a0 = t * (K + 1);
      uint32_t d = 0;
      for (; j < BPT; j += p) {
        const uint32_t addr = a0 + (j >> 5);
        const uint32_t bit = 1u << (j & 31u);
        const uint32_t w = lds[addr];
        if (w & bit) { lds[addr] = w & ~bit; d++; }
      }
      if (d) { atomicAdd(&sh.u.st.dec[s - chunk0], d); cnt -= d; }
    }
    if (s + 1 - chunk0 == HARD_SCH) {
      hard_flush<T, K>(sh, chunk0, s + 1, running, outT, outM, rep == 0);
      chunk0 = s + 1;
    }
    cur = nxt;
  }
  hard_flush<T, K>(sh, chunk0, s, running, outT, outM, rep == 0);
  long long a = wave_sum((long long)acc);
  if (lane == 0) sh.red_s[wv] = a;
  __syncthreads();
  if (t == 0) {
    long long tt = 0;
    for (uint32_t w = 0; w < NW; w++) tt += sh.red_s[w];
    bd.hard_acc[q.mrow_base + mrow] = tt;
    if (rep == 0) bd.hard_bact[row] = s;
  }
}

template <int T, int K>
__device__ __forceinline__ void p2_segment(const BatchDev &bd, const QueryDev &q, uint64_t seg,
                                           uint32_t *lds, SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t s0 = q.p2_e + seg * 2 * SB;
  const uint64_t s1 = s0 + 2 * SB;
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const WheelInfo wi = bd.winfo[q.c];
  const uint32_t *__restrict__ sp = bd.primes;
  const uint32_t np_all = (uint32_t)bd.nprimes;

  // wheel + offsets of the block-shared primes (computed once, in parallel)
  seg_fill_wheel(lds, K, T, s0, bd.wheel + wi.off, wi.period);
  const uint32_t i0 = q.c;
  uint32_t ncoop = 0;                                    // shared primes: p < BPT
  {
    // primes[i0 + k] < BPT for k < ncoop; bounded by the LDS slots
    uint32_t lo_i = i0, hi_i = min(np_all, i0 + (uint32_t)P2_MAX_COOP);
    while (lo_i < hi_i) {                                // first index with p >= BPT or p^2 >= s1
      uint32_t mid = (lo_i + hi_i) / 2;
      uint64_t pm = sp[mid];
      if (pm < BPT && pm * pm < s1) lo_i = mid + 1; else hi_i = mid;
    }
    ncoop = lo_i - i0;
  }
  for (uint32_t k = t; k < ncoop; k += T) {
    const uint32_t p = sp[i0 + k];
    sh.u.offs[k] = (uint32_t)first_odd_multiple_bit(s0, p, (uint64_t)p * p);
  }
  __syncthreads();
  for (uint32_t k = 0; k < ncoop; k++) {
    const uint32_t p = sp[i0 + k];
    for (uint64_t j = sh.u.offs[k] + (uint64_t)t * p; j < SB; j += (uint64_t)T * p)
      atomicAnd(&lds[seg_addr((uint32_t)(j >> 5), K, T)], ~(1u << (j & 31)));
  }
  for (uint32_t k = i0 + ncoop + t; k < np_all; k += T) {   // one large prime per thread
    const uint32_t p = sp[k];
    if ((uint64_t)p * p >= s1) break;
    for (uint64_t j = first_odd_multiple_bit(s0, p, (uint64_t)p * p); j < SB; j += p)
      atomicAnd(&lds[seg_addr((uint32_t)(j >> 5), K, T)], ~(1u << (j & 31)));
  }
  __syncthreads();
  uint32_t cnt = 0;
#pragma unroll 8
  for (uint32_t j = 0; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  publish_counts<T, K>(cnt, sh);
  __syncthreads();
  uint32_t wp[NW];
  const uint32_t tot = read_wave_prefix<T, K>(sh, wp);

  // P2 primes p in (max(y, x/s1), min(sqrt x, x/s0)]: block strides the pi-table entries
  unsigned long long loc = 0;
  uint32_t np = 0;
  uint64_t pmax = cpx_div128_to64(x, s0);
  if (pmax > q.sqrtx) pmax = q.sqrtx;
  uint64_t pmin = cpx_div128_to64(x, s1);
  if (pmin < q.y) pmin = q.y;
  if (pmax > pmin) {
    const uint64_t e_lo = (pmin + 1) >> 6, e_hi = pmax >> 6;
    for (uint64_t e = e_lo + t; e <= e_hi; e += T) {
      const PiEntry pe = bd.pitab[e];
      uint32_t b = pe.bits;
      const uint64_t base = 64 * e;
      // keep odd integers base + 2i + 1 in (pmin, pmax]
      if (pmax < base + 63) b &= cpx_lowmask32((uint32_t)((pmax - base + 1) >> 1));
      if (pmin > base) {
        const uint64_t kk = (pmin - base + 1) >> 1;     // bits for integers <= pmin
        b = kk >= 32 ? 0u : (b & ~cpx_lowmask32((uint32_t)kk));
      }
      while (b) {
        const uint32_t i = __builtin_ctz(b);
        b &= b - 1;
        const uint64_t p = base + 2 * i + 1;
        const uint64_t v = cpx_div128_to64(x, p);       // s0 <= v < s1
        loc += (unsigned long long)count_below<T, K>(lds, sh, wp, (uint32_t)((v - s0 + 1) >> 1));
        np++;
      }
    }
  }
  loc = wave_sum(loc);
  const uint32_t npw = wave_sum(np);
  if (lane == 0) { sh.red_u[wv] = loc; sh.wsum[wv] = (int32_t)npw; }
  __syncthreads();
  if (t == 0) {
    unsigned long long L = 0;
    uint32_t N = 0;
    for (uint32_t w = 0; w < NW; w++) { L += sh.red_u[w]; N += (uint32_t)sh.wsum[w]; }
    const uint64_t row = q.p2_row0 + seg;
    bd.p2_cnt[row] = tot;
    bd.p2_np[row] = N;
    bd.p2_loc[row] = L;
  }
}

template <int T, int K>
__global__ __launch_bounds__(T) void k_segments(BatchDev bd)
{
  extern __shared__ uint32_t lds[];
  __shared__ SegShared<T, K> sh;
  const uint32_t ntiles = *bd.ntiles;
  for (;;) {
    if (threadIdx.x == 0) sh.tile = atomicAdd(bd.tile_head, 1u);
    __syncthreads();
    const uint32_t ti = sh.tile;
    if (ti >= ntiles) break;
    const TileDev tile = bd.tiles[ti];
    const QueryDev q = bd.queries[tile.query];
    if (tile.kind == TILE_HARD) hard_segment<T, K>(bd, q, tile, lds, sh);
    else p2_segment<T, K>(bd, q, tile.seg, lds, sh);
    __syncthreads();
  }
}

size_t seg_lds_bytes(SegShape shape) { return (size_t)(shape.kwords + 1) * shape.threads * 4; }

// One-wavefront workgroups (T = 64) make every block barrier a single-wave
// barrier and let ~15 independent segment workers share a CU.
#define CPX_SEG_SHAPES(X) \
  X(64, 32) X(64, 64) X(64, 128) X(128, 32) X(128, 64) X(256, 32) X(256, 64) X(512, 32)

bool seg_shape_supported(SegShape s)
{
#define X(TT, KK) if (s.threads == TT && s.kwords == KK) return true;
  CPX_SEG_SHAPES(X)
#undef X
  return false;
}

int seg_blocks_per_cu(SegShape shape)
{
  size_t lds = seg_lds_bytes(shape);
  int nb = 0;
#define X(TT, KK)                                                                            \
  if (shape.threads == TT && shape.kwords == KK) {                                           \
    (void)hipFuncSetAttribute((const void *)k_segments<TT, KK>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);         \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_segments<TT, KK>, TT, lds) !=   \
        hipSuccess)                                                                          \
      nb = 0;                                                                                \
  }
  CPX_SEG_SHAPES(X)
#undef X
  return nb;
}

hipError_t launch_segments(BatchDev b, SegShape shape, uint32_t grid, hipStream_t s)
{
  size_t lds = seg_lds_bytes(shape);
#define X(TT, KK)                                                                            \
  if (shape.threads == TT && shape.kwords == KK) {                                           \
    /* the >64 KiB LDS attribute is set by seg_blocks_per_cu before any launch */           \
    hipLaunchKernelGGL((k_segments<TT, KK>), dim3(grid), dim3(TT), lds, s, b);               \
    return hipGetLastError();                                                                \
  }
  CPX_SEG_SHAPES(X)
#undef X
  return hipErrorInvalidValue;
}
