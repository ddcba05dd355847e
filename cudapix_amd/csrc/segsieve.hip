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
    struct {                     // hard: stage table of the current chunk
      uint32_t dec[HARD_SCH];    // survivors removed by each stage
      int32_t msum[HARD_SCH];    // sum of mu of each stage's leaves (this replica)
      uint64_t magic[HARD_SCH];  // ceil(2^64 / p)
      uint32_t p[HARD_SCH];
      uint32_t oseg[HARD_SCH];   // first odd multiple of p in the segment (bit offset)
      uint32_t cbase[HARD_SCH];  // leaf candidates: mtop/2 (dense) or first prime index
      uint32_t ccnt[HARD_SCH];   // number of candidates | 2^31 if sparse (p^2 > y)
      uint32_t lmask[HARD_SCH / 32];   // stages with candidates
      uint32_t cend;             // first inactive stage of the chunk (or ~0)
    } st;
  } u;
  uint32_t tile;
};

// Each thread publishes, packed in one LDS word, its in-wave exclusive prefix
// of the survivor counts (low PRE_BITS bits: at most 63 * 32K) and the count
// of its first K/2 words (above), so that a count below any bit scans at most
// K/2 of the owner's words.
template <int K>
struct CountPack {
  static constexpr uint32_t PRE_BITS = K <= 32 ? 16 : (K <= 64 ? 17 : 18);
  static constexpr uint32_t PRE_MASK = (1u << PRE_BITS) - 1;
};

// Publish the block scan of the per-thread counts: in-wave exclusive prefix
// and half count per thread (packed), one total per wave.  Caller syncs afterwards.
template <int T, int K>
__device__ __forceinline__ void publish_counts(uint32_t cnt, uint32_t cnt_lo, SegShared<T, K> &sh)
{
  uint32_t incl = wave_incl_scan(cnt);
  sh.ipre[threadIdx.x] = (incl - cnt) | (cnt_lo << CountPack<K>::PRE_BITS);
  if (lane_id() == CPX_WAVE - 1) sh.wtot[wave_id()] = incl;
}

// Survivor count of the thread's first K/2 words.
template <int K>
__device__ __forceinline__ uint32_t own_count_lo(const uint32_t *lds, uint32_t t)
{
  uint32_t c = 0;
#pragma unroll
  for (uint32_t j = 0; j < (uint32_t)K / 2; j++) c += cpx_popc32(lds[own_addr(t, j, K)]);
  return c;
}

// Number of survivors among the first `bit` bits of the segment
// (0 <= bit <= 32*K*T), given the wave prefixes wp[] held in registers:
// wave prefix + in-wave prefix of the owning thread + its words below the
// bit, counted from the nearer half boundary (at most K/2 words, read four
// at a time so that each group costs one LDS round trip).
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
  const uint32_t e = sh.ipre[owner];
  uint32_t c = base + (e & CountPack<K>::PRE_MASK);
  uint32_t w0 = 0;                                   // first word to scan
  if (jb >= BPT / 2) { c += e >> CountPack<K>::PRE_BITS; w0 = K / 2; }
  const uint32_t wix = jb >> 5;                      // words [w0, wix) are whole
  const uint32_t *ow = lds + owner * (K + 1);
  for (uint32_t g = w0; g < wix; g += 4) {
    const uint32_t x0 = ow[g], x1 = ow[g + 1], x2 = ow[g + 2], x3 = ow[g + 3];
    c += cpx_popc32(x0);
    c += g + 1 < wix ? cpx_popc32(x1) : 0u;
    c += g + 2 < wix ? cpx_popc32(x2) : 0u;
    c += g + 3 < wix ? cpx_popc32(x3) : 0u;
  }
  if (jb & 31u) c += cpx_popc32(ow[wix] & cpx_lowmask32(jb & 31u));
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
// only), sum(mu) of this replica; then clear the chunk.  All threads call;
// `running` is kept by wave 0.
template <int T, int K>
__device__ __forceinline__ void hard_flush(SegShared<T, K> &sh, uint32_t c0, uint32_t c1,
                                           uint32_t &running, uint32_t *outT, int32_t *outM,
                                           bool write_T)
{
  __syncthreads();
  if (wave_id() == 0) {                 // wave 0 scans the removals 64 stages at a time
    const uint32_t lane = lane_id();
    for (uint32_t k0 = c0; k0 < c1; k0 += CPX_WAVE) {
      const uint32_t k = k0 + lane;
      const uint32_t d = k < c1 ? sh.u.st.dec[k - c0] : 0u;
      const uint32_t incl = wave_incl_scan(d);
      if (k < c1) {
        if (write_T) outT[k] = running - (incl - d);
        outM[k] = sh.u.st.msum[k - c0];
      }
      running -= __shfl(incl, CPX_WAVE - 1);
    }
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < HARD_SCH; k += T) { sh.u.st.dec[k] = 0; sh.u.st.msum[k] = 0; }
  if (threadIdx.x < HARD_SCH / 32) sh.u.st.lmask[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.u.st.cend = ~0u;
  __syncthreads();
}

// Leaf candidate i of chunk stage k: dense stages read mu*lpf of m = mtop - 2i
// (index mtop/2 - i), sparse stages read the i-th prime above mlo.
__device__ __forceinline__ int32_t hard_fetch(const BatchDev &bd, uint32_t cbase, bool sparse,
                                              uint32_t i)
{
  return sparse ? (int32_t)bd.primes[cbase + i] : bd.mlpf[cbase - i];
}

// Stage table of one chunk: for its stages k < clen (stage s0 + k) the
// threads write p, the Lemire constant, the first odd multiple of p in the
// segment (bit offset), the leaf candidate range and a leafy bit; the first
// inactive stage (p^2 > x/lo) lowers st.cend.  Caller syncs before and after.
template <int T, int K>
__device__ __forceinline__ void hard_chunk_table(const BatchDev &bd, const QueryDev &q,
                                                 const StageDev *st, uint32_t s0, uint32_t clen,
                                                 uint64_t lo, uint64_t hi, uint64_t xlo,
                                                 double inv_lo, double inv_hi,
                                                 SegShared<T, K> &sh)
{
  const uint64_t y = q.y;
  for (uint32_t k = threadIdx.x; k < clen; k += T) {
    const StageDev sd = st[s0 + k];
    const uint64_t p = sd.p, xp = sd.xp;
    if (p * p > xlo) {                       // inactive: so are all later stages
      atomicMin(&sh.u.st.cend, k);
      continue;
    }
    sh.u.st.p[k] = (uint32_t)p;
    sh.u.st.magic[k] = sd.magic;
    // first odd multiple of p at or above lo + 1, as a bit offset (< p)
    const uint64_t start = lo + 1;
    const uint64_t rr = start - cpx_div_recip(start, p, sd.inv_p) * p;
    uint64_t mm = start + (rr ? p - rr : 0);
    if (!(mm & 1)) mm += p;
    sh.u.st.oseg[k] = (uint32_t)((mm - start) >> 1);
    // m range of this segment's leaves of p: x/(p hi) < m <= x/(p lo), m > y/p
    const bool sparse = p * p > y;
    uint64_t mhi = y;
    if (lo != 0 && (xp >> 50) < lo) {
      const uint64_t v = cpx_div_recip(xp, lo, inv_lo);
      if (v < mhi) mhi = v;
    }
    uint64_t mlo = ~0ull;
    if ((xp >> 50) < hi) mlo = cpx_div_recip(xp, hi, inv_hi);
    uint32_t ncand = 0, cbase = 0;
    if (mlo < mhi) {
      if (mlo < sd.ylo) mlo = sd.ylo;
      if (sparse && mlo < p) mlo = p;
      if (mhi > mlo) {
        if (!sparse) {
          const uint64_t mtop = mhi - ((mhi & 1) ^ 1);            // largest odd <= mhi
          ncand = mtop > mlo ? (uint32_t)((mtop - mlo + 1) / 2) : 0;
          cbase = (uint32_t)(mtop >> 1);
        } else {
          const uint64_t il = pi_lookup(bd.pitab, mlo);
          ncand = (uint32_t)(pi_lookup(bd.pitab, mhi) - il);
          cbase = (uint32_t)il;
        }
      }
    }
    sh.u.st.cbase[k] = cbase;
    sh.u.st.ccnt[k] = ncand | (sparse ? 0x80000000u : 0u);
    if (ncand) atomicOr(&sh.u.st.lmask[k >> 5], 1u << (k & 31));
  }
}

// First leafy stage of the chunk at or after k (HARD_SCH if none).
template <int T, int K>
__device__ __forceinline__ uint32_t hard_next_leafy(const SegShared<T, K> &sh, uint32_t k)
{
  for (uint32_t w = k >> 5; w < HARD_SCH / 32; w++) {
    uint32_t m = sh.u.st.lmask[w];
    if (w == (k >> 5)) m &= ~cpx_lowmask32(k & 31);
    if (m) return w * 32 + (uint32_t)__builtin_ctz(m);
  }
  return HARD_SCH;
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
  const WheelInfo wi = bd.winfo[q.c];
  const uint64_t row = q.hard_row0 + seg;
  const HardRowInfo ri = bd.rowinfo[row];
  const uint32_t rep = tile.rep, nrep = ri.nrep;
  const uint64_t mrow = (uint64_t)ri.mrow + rep;
  // counters of the active stages only: the plan sized this row for ri.act
  // stages, and the stage loop stops there at the latest
  const uint32_t nst = min(q.nstage, ri.act);
  uint32_t *outT = bd.hard_T + q.tm_off + ri.toff;
  int32_t *outM = bd.hard_M + q.m_off + ri.moff + (uint64_t)rep * ri.act;
  const StageDev *st = bd.stages + q.stage_off;
  const uint32_t i0 = rep * T + t, step = T * nrep;     // this thread's leaf candidates

  // stage b is active while p_b^2 <= x / lo (all leaves of p_b have n < x/p_b^2)
  uint64_t xlo = ~0ull;
  if (lo) {
    u128 qq = x / lo;
    xlo = (qq >> 64) ? ~0ull : (uint64_t)qq;
  }
  const double inv_lo = lo ? 1.0 / (double)lo : 0.0, inv_hi = 1.0 / (double)hi;

  seg_fill_wheel(lds, K, T, lo, bd.wheel + wi.off, wi.period);
  for (uint32_t k = t; k < HARD_SCH; k += T) { sh.u.st.dec[k] = 0; sh.u.st.msum[k] = 0; }
  if (t < HARD_SCH / 32) sh.u.st.lmask[t] = 0;
  if (t == 0) sh.u.st.cend = ~0u;
  __syncthreads();
  uint32_t cnt_lo = own_count_lo<K>(lds, t), cnt = cnt_lo;   // survivors: first half, all
#pragma unroll 8
  for (uint32_t j = K / 2; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  {
    uint32_t ws = wave_sum(cnt);
    if (lane == 0) sh.wsum[wv] = (int32_t)ws;           // wtot is rewritten by the stages
  }
  uint32_t s0 = 0, clen = min((uint32_t)HARD_SCH, nst);
  hard_chunk_table<T, K>(bd, q, st, s0, clen, lo, hi, xlo, inv_lo, inv_hi, sh);
  __syncthreads();
  uint32_t running = 0;                                 // survivors before the next stage
  for (uint32_t w = 0; w < NW; w++) running += (uint32_t)sh.wsum[w];

  int64_t acc = 0;
  uint32_t nact = 0;                                    // active stages of the segment
  for (;;) {
    const uint32_t cend = min(clen, sh.u.st.cend);
    // one-ahead prefetch of this thread's first candidate of the next leafy stage
    // (and of x/p for that stage)
    uint32_t pf_k = hard_next_leafy<T, K>(sh, 0);
    int32_t pf = 0;
    uint64_t pf_xp = 0;
    if (pf_k < cend) {
      const uint32_t cc = sh.u.st.ccnt[pf_k];
      pf_xp = st[s0 + pf_k].xp;
      if (i0 < (cc & 0x7fffffffu)) pf = hard_fetch(bd, sh.u.st.cbase[pf_k], cc >> 31, i0);
    }
    uint32_t lw = 0;
    for (uint32_t k = 0; k < cend; k++) {
      if ((k & 31) == 0) lw = sh.u.st.lmask[k >> 5];
      const uint32_t p = sh.u.st.p[k];
      if ((lw >> (k & 31)) & 1u) {
        // ---- leaves of p (state: crossed by p_1 .. p_{b-1}); needs every
        //      thread's words and counts, so the block meets here
        publish_counts<T, K>(cnt, cnt_lo, sh);
        __syncthreads();
        uint32_t wp[NW];
        (void)read_wave_prefix<T, K>(sh, wp);
        const uint32_t cc = sh.u.st.ccnt[k], cb = sh.u.st.cbase[k];
        const bool sparse = cc >> 31;
        const uint32_t ncand = cc & 0x7fffffffu;
        const uint64_t xp = pf_xp;                            // x / p, fetched ahead
        const double xpf = (double)xp;
        int32_t smu = 0;
        int32_t v = pf;
        const uint32_t nk = hard_next_leafy<T, K>(sh, k + 1);
        for (uint32_t i = i0; i < ncand; i += step) {
          if (i != i0) v = hard_fetch(bd, cb, sparse, i);
          uint64_t m;
          if (sparse) {
            m = (uint32_t)v;                                  // prime, mu = -1
            v = -1;
          } else {
            const uint32_t av = (uint32_t)(v < 0 ? -v : v);
            if (av <= p) continue;                            // mu(m) = 0 or lpf(m) <= p
            m = 2 * (uint64_t)(cb - i) + 1;
          }
          const uint64_t n = cpx_div_leaf(xp, xpf, (uint32_t)m);   // lo <= n < hi (< 2^50)
          const uint32_t c = count_below<T, K>(lds, sh, wp, (uint32_t)((n - lo + 1) >> 1));
          if (v > 0) { acc += c; smu++; } else { acc -= c; smu--; }
        }
        pf_k = nk;
        if (nk < cend) {
          const uint32_t c2 = sh.u.st.ccnt[nk];
          pf_xp = st[s0 + nk].xp;
          if (i0 < (c2 & 0x7fffffffu)) pf = hard_fetch(bd, sh.u.st.cbase[nk], c2 >> 31, i0);
        }
        {
          const int32_t wm = (int32_t)wave_sum((uint32_t)smu);   // one LDS add per wave
          if (wm && lane == 0) atomicAdd(&sh.u.st.msum[k], wm);
        }
        __syncthreads();                                      // readers done before cross-off
      }
      // ---- cross off the odd multiples of p in this thread's own bits
      {
        const uint32_t oseg = sh.u.st.oseg[k];
        const uint32_t tb = cpx_fastmod32(t * BPT, p, sh.u.st.magic[k]);
        uint32_t j = oseg >= tb ? oseg - tb : oseg + p - tb;
        const uint32_t a0 = t * (K + 1);
        uint32_t d = 0, dlo = 0;
        for (; j < BPT; j += p) {
          const uint32_t addr = a0 + (j >> 5);
          const uint32_t bit = 1u << (j & 31u);
          const uint32_t w = lds[addr];
          if (w & bit) { lds[addr] = w & ~bit; d++; dlo += j < BPT / 2; }
        }
        cnt -= d;
        cnt_lo -= dlo;
        const uint32_t wd = wave_sum(d);                       // one LDS add per wave
        if (wd && lane == 0) atomicAdd(&sh.u.st.dec[k], wd);
      }
    }
    nact = s0 + cend;
    hard_flush<T, K>(sh, s0, nact, running, outT, outM, rep == 0);
    if (cend < clen || nact >= nst) break;
    s0 = nact;
    clen = min((uint32_t)HARD_SCH, nst - s0);
    hard_chunk_table<T, K>(bd, q, st, s0, clen, lo, hi, xlo, inv_lo, inv_hi, sh);
    __syncthreads();
  }
  long long a = wave_sum((long long)acc);
  if (lane == 0) sh.red_s[wv] = a;
  __syncthreads();
  if (t == 0) {
    long long tt = 0;
    for (uint32_t w = 0; w < NW; w++) tt += sh.red_s[w];
    bd.hard_acc[q.mrow_base + mrow] = tt;
    if (rep == 0) bd.hard_bact[row] = nact;
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
  const uint32_t cnt_lo = own_count_lo<K>(lds, t);
  uint32_t cnt = cnt_lo;
#pragma unroll 8
  for (uint32_t j = K / 2; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  publish_counts<T, K>(cnt, cnt_lo, sh);
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

// Self check of the register wavefront scan and sum against a serial scan
// through LDS: one 64-lane wavefront per block, n blocks over in[64 * n].
__global__ __launch_bounds__(CPX_WAVE) void k_selftest_wave(const uint32_t *__restrict__ in,
                                                            uint32_t n, uint32_t *bad)
{
  __shared__ uint32_t v[CPX_WAVE];
  const uint32_t l = threadIdx.x;
  if (blockIdx.x >= n) return;
  const uint32_t a = in[(uint64_t)blockIdx.x * CPX_WAVE + l];
  v[l] = a;
  __syncthreads();
  uint32_t ref = 0, tot = 0;
  for (uint32_t k = 0; k < CPX_WAVE; k++) {
    if (k <= l) ref += v[k];
    tot += v[k];
  }
  const uint32_t incl = wave_incl_scan(a);
  const uint32_t sum = wave_sum(a);
  if (incl != ref || sum != tot) atomicAdd(bad, 1u);
  // the leaf divide against the exact 64-bit divide: m in [2, 2^32), a with
  // a / m < 2^50 (splitmix64 of the input word and the lane)
  uint64_t h = ((uint64_t)a << 32) ^ ((uint64_t)blockIdx.x * CPX_WAVE + l) ^ 0x9e3779b97f4a7c15ull;
  h = (h ^ (h >> 30)) * 0xbf58476d1ce4e5b9ull;
  h = (h ^ (h >> 27)) * 0x94d049bb133111ebull;
  h ^= h >> 31;
  uint32_t m = (uint32_t)(h >> (32 + (a & 31)));
  if (m < 2) m = 2;
  uint64_t g = h * 0x2545f4914f6cdd1dull + l;
  if (m < (1u << 14)) g %= ((uint64_t)m << 50);
  if ((a & 7) == 0) g = ((uint64_t)m << 32) * ((uint64_t)1 << (a >> 27)) - 1 - (a & 1);  // near multiples
  if (g / m < (1ull << 50) && cpx_div_leaf(g, (double)g, m) != g / m) atomicAdd(bad, 1u);
}

hipError_t launch_selftest_wave(const uint32_t *in, uint32_t n, uint32_t *bad, hipStream_t s)
{
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_selftest_wave, dim3(n), dim3(CPX_WAVE), 0, s, in, n, bad);
  return hipGetLastError();
}
