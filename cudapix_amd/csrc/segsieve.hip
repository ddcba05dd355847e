// segsieve.hip -- the persistent segment kernel: hard-leaf phi sieve of one
// LDS-resident bitmap segment per tile, with the P2 prime count of the same
// segment as an epilogue.
//
// Layout: a segment of [lo, lo + 2*SB) holds its SB = 32*K*T odd numbers
// lo+1, lo+3, ... as bits (1 = unsieved) in padded-linear LDS; thread t owns
// the contiguous bits [t*32K, (t+1)*32K) and keeps their survivor count (and
// that of its first K/2 words) in registers; wave w owns 64 threads' bits.
//
// Hard leaves (replaces S2glob::S2ctl + S2hardHost, src/sieve/S2_hard_device.cu:264-355,
// src/sieve/S2_hard_host.cu:115-191; DR split p <= x^(1/4), src/Deleglise-Rivat/S3.cu:8):
//   The segment starts as the wheel of the first c primes.  Stages b (p = p_b,
//   c < b <= B, while p^2 <= x/lo) run in chunks of HARD_SCH: the block first
//   tabulates the chunk (p, x/p, first multiple, leaf candidate range of the
//   segment and its split among the waves), then every wave walks the chunk's
//   stages on its own, with no block barrier:
//     L  leaves of p whose n = x/(p m) falls in the wave's bits: m odd with
//        mu(m) != 0 and lpf(m) > p from the mu*lpf table (p^2 <= y) or m prime
//        (p^2 > y); the wave publishes its in-wave prefix counts and counts
//        phi_wave(n) = in-wave prefix of the owning lane + popcounts of its
//        words below n; sum mu * phi_wave per thread, sum mu per wave;
//     X  cross off the odd multiples of p in the thread's own bits, recording
//        each wave's removals.
//   At the end of a chunk the block adds the survivors of the earlier waves
//   to every leaf's count (from the per-stage, per-wave removals and sums of
//   mu), and writes T_b (survivors before stage b) and sum_b(mu); the carry
//   kernel adds sum_b prefix_b(seg) * sum_b(mu) over the segments (the
//   decomposition the reference does with scanVectorized/addMultiply,
//   src/sieve/S2_hard_device.cu:357-375).
//
// P2: the segments holding x/p for the P2 primes p in (y, sqrt x] get, on
// replica 0 of their tile, the rest of a prime sieve (primes above the last
// stage up to sqrt(hi)) and the P2 counts; see p2_epilogue.
#include "cpx_device.hpp"

#define HARD_SCH 128       // stages per chunk of the hard stage table

template <int T, int K>
struct SegShared {
  uint32_t ipre[T];              // in-wave exclusive prefix of each thread's count
  uint32_t wtot[T / CPX_WAVE];   // wave totals
  int32_t wsum[T / CPX_WAVE];    // per-wave sums (P2)
  unsigned long long red_u[T / CPX_WAVE];
  long long red_s[T / CPX_WAVE];
  union {
    struct {                     // hard: stage table and per-wave counters of the chunk
      uint64_t xp[HARD_SCH];     // floor(x / p)
      uint32_t p[HARD_SCH];
      float rinv[HARD_SCH];      // 1 / p
      uint32_t oseg[HARD_SCH];   // first odd multiple of p in the segment (bit offset)
      uint32_t cbase[HARD_SCH];  // leaf candidates: mtop/2 (dense) or first prime index
      uint32_t ccnt[HARD_SCH];   // candidates | HARD_WIDE | HARD_SPARSE
      int32_t wm[HARD_SCH][T / CPX_WAVE];    // sum of mu of each wave's leaves per stage
      uint16_t dw[HARD_SCH][T / CPX_WAVE];   // survivors each wave removes per stage
      uint16_t wb[HARD_SCH][T / CPX_WAVE];   // first candidate of each wave (non-wide stages)
      uint32_t c0[T / CPX_WAVE]; // each wave's survivors at the chunk start
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

// Leaf candidate i of chunk stage k: dense stages read mu*lpf of m = mtop - 2i
// (index mtop/2 - i), sparse stages read the i-th prime above mlo.
__device__ __forceinline__ int32_t hard_fetch(const BatchDev &bd, uint32_t cbase, bool sparse,
                                              uint32_t i)
{
  return sparse ? (int32_t)bd.primes[cbase + i] : bd.mlpf[cbase - i];
}

// ccnt word of the stage table: number of candidates (bits 0-29), "wide"
// (bit 30: 2^16 candidates or more, so the per-wave boundaries are not
// tabulated) and "sparse" (bit 31: p^2 > y, candidates are primes).
#define HARD_NC_MASK 0x3fffffffu
#define HARD_WIDE 0x40000000u
#define HARD_SPARSE 0x80000000u

// Stage table of one chunk: for its stages k < clen (stage s0 + k) the
// threads write p, 1/p, x/p, the first odd multiple of p in the segment (bit
// offset), the leaf candidate range, a leafy bit, and for every wave w > 0
// the boundary b_w of its candidates (wave w owns the leaves with n = x/(p m)
// in [lo_w, lo_w + 4*32K*64)): for dense stages (m = mtop - 2i, n grows with
// i) b_w counts the candidates with n < lo_w and is the wave's first index;
// for sparse stages (m = the i-th prime, n falls with i) it counts those with
// n >= lo_w and is the wave's end.  Clears the per-wave sums of mu; the first
// inactive stage (p^2 > x/lo) lowers st.cend.  Caller syncs before and after.
template <int T, int K>
__device__ __forceinline__ void hard_chunk_table(const BatchDev &bd, const QueryDev &q,
                                                 const StageDev *st, uint32_t s0, uint32_t clen,
                                                 uint64_t lo, uint64_t hi, uint64_t xlo,
                                                 double inv_lo, double inv_hi,
                                                 SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint64_t WB = (uint64_t)CPX_WAVE * 32 * K;
  const uint64_t y = q.y;
  for (uint32_t k = threadIdx.x; k < clen * NW; k += T) (&sh.u.st.wm[0][0])[k] = 0;
  for (uint32_t k = threadIdx.x; k < clen; k += T) {
    const StageDev sd = st[s0 + k];
    const uint64_t p = sd.p, xp = sd.xp;
    if (p * p > xlo) {                       // inactive: so are all later stages
      atomicMin(&sh.u.st.cend, k);
      continue;
    }
    sh.u.st.p[k] = (uint32_t)p;
    sh.u.st.rinv[k] = 1.0f / (float)p;
    sh.u.st.xp[k] = xp;
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
    sh.u.st.ccnt[k] = ncand | (sparse ? HARD_SPARSE : 0u) | (ncand > 0xffffu ? HARD_WIDE : 0u);
    if (ncand) atomicOr(&sh.u.st.lmask[k >> 5], 1u << (k & 31));
  }
  if (NW == 1) return;
  __syncthreads();
  // n >= lo_w <=> m <= M = floor(x/p / lo_w)
  const uint32_t cend = min(clen, sh.u.st.cend);
  for (uint32_t idx = threadIdx.x; idx < cend * NW; idx += T) {
    const uint32_t k = idx / NW, w = idx % NW;
    const uint32_t cc = sh.u.st.ccnt[k];
    if (w == 0 || !(cc & HARD_NC_MASK) || (cc & HARD_WIDE)) continue;
    const uint32_t ncand = cc & HARD_NC_MASK, cb = sh.u.st.cbase[k];
    const uint64_t xp = sh.u.st.xp[k];
    const uint64_t lo_w = lo + 2 * (uint64_t)w * WB;          // > 0
    const uint64_t M = cpx_div_recip(xp, lo_w, 1.0 / (double)lo_w);
    uint64_t b;
    if (!(cc & HARD_SPARSE)) {
      const uint64_t mtop = 2 * (uint64_t)cb + 1;            // m = mtop - 2i
      b = mtop > M ? (mtop - M + 1) / 2 : 0;
    } else {
      const uint64_t pm = pi_lookup(bd.pitab, M < y ? M : y);
      b = pm > cb ? pm - cb : 0;
    }
    sh.u.st.wb[k][w] = (uint16_t)(b < ncand ? b : ncand);
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

// This wave's candidate index range [ia, ib) of chunk stage k.  Stages with
// 2^16 candidates or more derive it here: n = x/(p m) in [lo_w, hi_w) <=>
// floor(xp / hi_w) < m <= floor(xp / lo_w).
template <int T, int K>
__device__ __forceinline__ void hard_wave_range(const BatchDev &bd, const QueryDev &q,
                                                const SegShared<T, K> &sh, uint32_t k, uint32_t wv,
                                                uint64_t lo_w, uint64_t hi_w, double inv_lo_w,
                                                double inv_hi_w, uint32_t &ia, uint32_t &ib)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  const uint32_t cc = sh.u.st.ccnt[k], ncand = cc & HARD_NC_MASK;
  if (!(cc & HARD_WIDE)) {
    if (!(cc & HARD_SPARSE)) {               // n grows with the index
      ia = wv ? sh.u.st.wb[k][wv] : 0u;
      ib = wv + 1 < NW ? sh.u.st.wb[k][wv + 1] : ncand;
    } else {                                 // n falls with the index
      ia = wv + 1 < NW ? sh.u.st.wb[k][wv + 1] : 0u;
      ib = wv ? sh.u.st.wb[k][wv] : ncand;
    }
    return;
  }
  const uint32_t cb = sh.u.st.cbase[k];
  const uint64_t xp = sh.u.st.xp[k];
  const uint64_t Mlo = lo_w ? cpx_div_recip(xp, lo_w, inv_lo_w) : ~0ull;
  const uint64_t Mhi = cpx_div_recip(xp, hi_w, inv_hi_w);
  if (!(cc & HARD_SPARSE)) {                                 // m = mtop - 2i
    const uint64_t mtop = 2 * (uint64_t)cb + 1;
    const uint64_t a = mtop > Mlo ? (mtop - Mlo + 1) / 2 : 0;
    const uint64_t b = mtop > Mhi ? (mtop - Mhi + 1) / 2 : 0;
    ia = (uint32_t)min(a, (uint64_t)ncand);
    ib = (uint32_t)min(b, (uint64_t)ncand);
  } else {                                                   // primes[cb + i]
    const uint64_t y = q.y;
    const uint64_t pl = pi_lookup(bd.pitab, Mhi < y ? Mhi : y);
    const uint64_t ph = pi_lookup(bd.pitab, Mlo < y ? Mlo : y);
    ia = (uint32_t)(pl > cb ? min(pl - cb, (uint64_t)ncand) : 0);
    ib = (uint32_t)(ph > cb ? min(ph - cb, (uint64_t)ncand) : 0);
  }
}

// Survivors among the first `gw` bits (0 <= gw <= 64 * 32K) of this wave's
// part of the segment: in-wave prefix of the owning lane + its words below,
// from the packed prefix/half-count words the wave published.
template <int T, int K>
__device__ __forceinline__ uint32_t count_in_wave(const uint32_t *lds, const SegShared<T, K> &sh,
                                                  uint32_t wv, uint32_t gw)
{
  constexpr uint32_t BPT = 32 * K;
  uint32_t ol = gw / BPT;
  uint32_t jb = gw - ol * BPT;
  if (ol >= (uint32_t)CPX_WAVE) { ol = CPX_WAVE - 1; jb = BPT; }
  const uint32_t owner = wv * CPX_WAVE + ol;
  const uint32_t e = sh.ipre[owner];
  uint32_t c = e & CountPack<K>::PRE_MASK;
  uint32_t w0 = 0;
  if (jb >= BPT / 2) { c += e >> CountPack<K>::PRE_BITS; w0 = K / 2; }
  const uint32_t wix = jb >> 5;
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

// End of a chunk of stages [c0, c1) (all threads call; `running` is kept by
// wave 0): writes T_b = survivors of the segment before stage b (replica 0)
// and this replica's sum of mu per stage, and adds to acc the part of every
// leaf's phi that the in-wave count left out -- the survivors of the earlier
// waves at that stage:
//   sum_k sum_w wm[k][w] * (A_w - sum_{j<k} sum_{w'<w} dw[j][w'])
// with A_w the earlier waves' survivors at the chunk start.  Then resets the
// chunk's leafy mask and end marker.
template <int T, int K>
__device__ __forceinline__ void hard_flush(SegShared<T, K> &sh, uint32_t c0, uint32_t c1,
                                           uint32_t &running, uint32_t *outT, int32_t *outM,
                                           bool write_T, int64_t &acc)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t n = c1 - c0;
  __syncthreads();
  if (NW > 1) {
    uint32_t A = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) A += w < wv ? sh.u.st.c0[w] : 0u;
    uint32_t carry = 0;
    int64_t corr = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += CPX_WAVE) {
      const uint32_t k = k0 + lane;
      uint32_t G = 0;
      int32_t m = 0;
      if (k < n) {
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) G += w < wv ? (uint32_t)sh.u.st.dw[k][w] : 0u;
        m = sh.u.st.wm[k][wv];
      }
      const uint32_t incl = wave_incl_scan(G);
      corr += (int64_t)m * ((int64_t)A - (int64_t)(carry + incl - G));
      carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, CPX_WAVE - 1);
    }
    acc += corr;
  }
  if (wv == 0) {                       // T_b and the sum of mu per stage
    for (uint32_t k0 = 0; k0 < n; k0 += CPX_WAVE) {
      const uint32_t k = k0 + lane;
      uint32_t d = 0;
      int32_t ms = 0;
      if (k < n) {
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) { d += sh.u.st.dw[k][w]; ms += sh.u.st.wm[k][w]; }
      }
      const uint32_t incl = wave_incl_scan(d);
      if (k < n) {
        if (write_T) outT[c0 + k] = running - (incl - d);
        outM[c0 + k] = ms;
      }
      running -= (uint32_t)__builtin_amdgcn_readlane((int)incl, CPX_WAVE - 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < HARD_SCH / 32) sh.u.st.lmask[threadIdx.x] = 0;
  if (threadIdx.x == 0) sh.u.st.cend = ~0u;
  __syncthreads();
}

// P2 leaves of a segment (replaces the windowed upper_bound count of
// src/P2.cu:76-133), run by replica 0 of a hard tile after its stages: the
// bitmap already holds [lo, hi) sieved by the wheel and the stage primes
// p_(c+1) .. p_kfirst; the primes from index kfirst with p^2 < hi are
// crossed here (one per thread, from p^2), which leaves exactly the primes
// above p2_e.  Then for each P2 prime p in (max(y, x/hi), min(sqrt x, x/lo)]
// (pi-table bits, block-strided) the count of primes in (p2_e, x/p] held in
// this segment; per segment: primes in (p2_e, hi), number of p, sum of the
// counts.  The carry is applied in the combine kernel.
template <int T, int K>
__device__ __forceinline__ void p2_epilogue(const BatchDev &bd, const QueryDev &q, uint64_t seg,
                                            uint64_t lo, uint32_t kfirst, uint32_t *lds,
                                            SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t hi = lo + 2 * SB;
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const uint32_t *__restrict__ sp = bd.primes;
  const uint32_t np_all = (uint32_t)bd.nprimes;

  __syncthreads();
  for (uint32_t k = kfirst + t; k < np_all; k += T) {
    const uint32_t p = sp[k];
    if ((uint64_t)p * p >= hi) break;
    const uint64_t o = first_odd_multiple_bit_r(lo, p, (uint64_t)p * p);
    if (o >= SB) continue;
    for (uint32_t j = (uint32_t)o; j < (uint32_t)SB; j += p) {   // SB + p < 2^32
      const uint32_t g = j >> 5;
      atomicAnd(&lds[g + g / K], ~(1u << (j & 31)));
    }
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
  // survivors at or below p2_e (1, the stage primes and what is below
  // sqrt x are not counted): only the first segment of the range has any
  const uint32_t ce = q.p2_e >= lo ? count_below<T, K>(lds, sh, wp, (uint32_t)min((q.p2_e - lo + 1) >> 1, SB)) : 0u;

  unsigned long long loc = 0;
  uint32_t np = 0;
  uint64_t pmax = q.sqrtx;
  if (lo) {
    const uint64_t v = cpx_div128_to64(x, lo);
    if (v < pmax) pmax = v;
  }
  uint64_t pmin = cpx_div128_to64(x, hi);
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
        const uint64_t v = cpx_div128_to64(x, p);       // lo <= v < hi
        loc += (unsigned long long)(count_below<T, K>(lds, sh, wp, (uint32_t)((v - lo + 1) >> 1)) - ce);
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
    const uint64_t row = q.p2_row0 + (seg - q.p2_seg0);
    bd.p2_cnt[row] = tot - ce;
    bd.p2_np[row] = N;
    bd.p2_loc[row] = L;
  }
}

template <int T, int K>
__device__ __forceinline__ void hard_segment(const BatchDev &bd, const QueryDev &q,
                                             const TileDev &tile, uint32_t *lds, SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  constexpr uint32_t WB = CPX_WAVE * BPT;                // bits per wave
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
  // this wave's integers [lo_w, hi_w); this lane's first bit, for t*32K mod p
  const uint64_t lo_w = lo + 2 * (uint64_t)wv * WB, hi_w = lo_w + 2 * (uint64_t)WB;
  const double inv_lo_w = lo_w ? 1.0 / (double)lo_w : 0.0, inv_hi_w = 1.0 / (double)hi_w;
  const uint32_t b0 = t * BPT;                          // < 2^20
  const float fb0 = (float)b0;
  const uint32_t cstep = CPX_WAVE * nrep;               // candidate stride of a lane

  // stage b is active while p_b^2 <= x / lo (all leaves of p_b have n < x/p_b^2)
  uint64_t xlo = ~0ull;
  if (lo) {
    u128 qq = x / lo;
    xlo = (qq >> 64) ? ~0ull : (uint64_t)qq;
  }
  const double inv_lo = lo ? 1.0 / (double)lo : 0.0, inv_hi = 1.0 / (double)hi;

  seg_fill_wheel(lds, K, T, lo, bd.wheel + wi.off, wi.period);
  if (t < HARD_SCH / 32) sh.u.st.lmask[t] = 0;
  if (t == 0) sh.u.st.cend = ~0u;
  __syncthreads();
  uint32_t cnt_lo = own_count_lo<K>(lds, t), cnt = cnt_lo;   // survivors: first half, all
#pragma unroll 8
  for (uint32_t j = K / 2; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  {
    uint32_t ws = wave_sum(cnt);
    if (lane == 0) sh.wsum[wv] = (int32_t)ws;
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
    {
      const uint32_t ws = wave_sum(cnt);                // this wave's survivors at the chunk start
      if (lane == 0) sh.u.st.c0[wv] = ws;
    }
    // one-ahead prefetch of this lane's first candidate of the next leafy stage
    uint32_t pf_k = hard_next_leafy<T, K>(sh, 0);
    int32_t pf = 0;
    if (pf_k < cend) {
      uint32_t ia, ib;
      hard_wave_range<T, K>(bd, q, sh, pf_k, wv, lo_w, hi_w, inv_lo_w, inv_hi_w, ia, ib);
      const uint32_t cc = sh.u.st.ccnt[pf_k];
      if (ia + rep * CPX_WAVE + lane < ib)
        pf = hard_fetch(bd, sh.u.st.cbase[pf_k], cc & HARD_SPARSE, ia + rep * CPX_WAVE + lane);
    }
    uint32_t lw = 0;
    for (uint32_t k = 0; k < cend; k++) {
      if ((k & 31) == 0) lw = sh.u.st.lmask[k >> 5];
      const uint32_t p = sh.u.st.p[k];
      if ((lw >> (k & 31)) & 1u) {
        // ---- leaves of p whose n falls in this wave's bits (state: crossed
        //      by p_1 .. p_{b-1}); the wave publishes its own prefix, no
        //      other wave is involved
        {
          const uint32_t incl = wave_incl_scan(cnt);
          sh.ipre[t] = (incl - cnt) | (cnt_lo << CountPack<K>::PRE_BITS);
        }
        const uint32_t cc = sh.u.st.ccnt[k], cb = sh.u.st.cbase[k];
        const bool sparse = cc & HARD_SPARSE;
        const uint64_t xp = sh.u.st.xp[k];
        const double xpf = (double)xp;
        uint32_t ia, ib;
        hard_wave_range<T, K>(bd, q, sh, k, wv, lo_w, hi_w, inv_lo_w, inv_hi_w, ia, ib);
        int32_t smu = 0;
        if (ib > ia) {
          int32_t v = pf;
          for (uint32_t i = ia + rep * CPX_WAVE + lane; i < ib; i += cstep) {
            if (i != ia + rep * CPX_WAVE + lane) v = hard_fetch(bd, cb, sparse, i);
            uint64_t m;
            int32_t sg = -1;
            if (sparse) {
              m = (uint32_t)v;                              // prime, mu = -1
            } else {
              const uint32_t av = (uint32_t)(v < 0 ? -v : v);
              if (av <= p) continue;                        // mu(m) = 0 or lpf(m) <= p
              m = 2 * (uint64_t)(cb - i) + 1;
              sg = v > 0 ? 1 : -1;
            }
            const uint64_t n = cpx_div_leaf(xp, xpf, (uint32_t)m);   // lo_w <= n < hi_w
            const uint32_t c = count_in_wave<T, K>(lds, sh, wv, (uint32_t)((n - lo_w + 1) >> 1));
            if (sg > 0) { acc += c; smu++; } else { acc -= c; smu--; }
          }
          // sum of mu over the wave (one candidate per lane at most: two ballots)
          int32_t wm;
          if (ib - ia <= cstep)
            wm = (int32_t)__popcll(__ballot(smu > 0)) - (int32_t)__popcll(__ballot(smu < 0));
          else
            wm = (int32_t)wave_sum((uint32_t)smu);
          if (lane == 0) sh.u.st.wm[k][wv] = wm;
        }
        const uint32_t nk = hard_next_leafy<T, K>(sh, k + 1);
        if (nk < cend) {
          uint32_t na, nb;
          hard_wave_range<T, K>(bd, q, sh, nk, wv, lo_w, hi_w, inv_lo_w, inv_hi_w, na, nb);
          if (na + rep * CPX_WAVE + lane < nb)
            pf = hard_fetch(bd, sh.u.st.cbase[nk], sh.u.st.ccnt[nk] & HARD_SPARSE,
                            na + rep * CPX_WAVE + lane);
        }
      }
      // ---- cross off the odd multiples of p in this thread's own bits,
      //      from (t * 32K) mod p by the f32 reciprocal (t * 32K < 2^20, so
      //      the truncated estimate is the quotient or one off either way)
      {
        const uint32_t oseg = sh.u.st.oseg[k];
        uint32_t tb = b0;
        if (p <= b0) {                                    // p < 2^20 here
          const uint32_t qe = (uint32_t)(fb0 * sh.u.st.rinv[k]);
          int32_t r = (int32_t)(b0 - __umul24(qe, p));
          if (r < 0) r += (int32_t)p;
          else if (r >= (int32_t)p) r -= (int32_t)p;
          tb = (uint32_t)r;
        }
        uint32_t j = oseg >= tb ? oseg - tb : oseg + p - tb;
        const uint32_t a0 = t * (K + 1);
        uint32_t wd;
        if (p > BPT) {                                    // at most one multiple per thread
          uint32_t hit = 0;
          if (j < BPT) {
            const uint32_t addr = a0 + (j >> 5);
            const uint32_t bit = 1u << (j & 31u);
            const uint32_t w = lds[addr];
            if (w & bit) { lds[addr] = w & ~bit; hit = 1; }
          }
          cnt -= hit;
          cnt_lo -= hit & (j < BPT / 2 ? 1u : 0u);
          wd = (uint32_t)__popcll(__ballot(hit));
        } else {
          uint32_t d = 0, dlo = 0;
          for (; j < BPT; j += p) {
            const uint32_t addr = a0 + (j >> 5);
            const uint32_t bit = 1u << (j & 31u);
            const uint32_t w = lds[addr];
            if (w & bit) { lds[addr] = w & ~bit; d++; dlo += j < BPT / 2; }
          }
          cnt -= d;
          cnt_lo -= dlo;
          wd = wave_sum(d);
        }
        if (lane == 0) sh.u.st.dw[k][wv] = (uint16_t)wd;   // this wave's removals
      }
    }
    nact = s0 + cend;
    hard_flush<T, K>(sh, s0, nact, running, outT, outM, rep == 0, acc);
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
  if (rep == 0 && seg >= q.p2_seg0 && seg < q.p2_seg0 + q.p2_nseg)
    p2_epilogue<T, K>(bd, q, seg, lo, q.c + nact, lds, sh);
}

template <int T, int K>
// (at least 4 waves per SIMD: two 512-thread blocks per CU need <= 128 VGPRs)
__global__ __launch_bounds__(T, 4) void k_segments(BatchDev bd)
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
    hard_segment<T, K>(bd, q, tile, lds, sh);
    __syncthreads();
  }
}

size_t seg_lds_bytes(SegShape shape) { return (size_t)(shape.kwords + 1) * shape.threads * 4; }

// One-wavefront workgroups (T = 64) make every block barrier a single-wave
// barrier and let ~15 independent segment workers share a CU.
#define CPX_SEG_SHAPES(X) \
  X(64, 32) X(64, 64) X(64, 128) X(128, 32) X(128, 64) X(256, 32) X(256, 64) X(256, 128) \
  X(512, 32) X(512, 64)

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
