// segtiles.hpp -- the tiles of the persistent segment kernel (segsieve.hip)
// and of the streaming kernel (stream.hip): hard-leaf phi sieve of one
// LDS-resident bitmap segment per tile, with the P2 prime count of the same
// segment as an epilogue; easy-leaf and P2-lookup tiles.
//
// Layout: a segment of [lo, lo + 2*SB) holds its SB = 32*K*T odd numbers
// lo+1, lo+3, ... as bits (1 = unsieved) in padded-linear LDS; thread t owns
// the contiguous bits [t*32K, (t+1)*32K) and keeps their survivor count (and
// that of its first K/2 words) in registers; wave w owns 64 threads' bits.
//
// A tile is one segment of a group of up to CPX_GROUP queries (similar x,
// same wheel): the sieve state after stage b does not depend on x, so the
// tile sieves the segment once and evaluates every member's leaves against
// it.
//
// Hard leaves (replaces S2glob::S2ctl + S2hardHost, src/sieve/S2_hard_device.cu:264-355,
// src/sieve/S2_hard_host.cu:115-191; DR split p <= x^(1/4), src/Deleglise-Rivat/S3.cu:8):
//   The segment starts as the wheel of the first c primes.  Stages b (p = p_b,
//   c < b <= B, while p^2 <= x/lo for some member) run in chunks of HARD_SCH
//   (stage, member) items: the block first tabulates the chunk (p, first
//   multiple; per item x/p, the leaf candidate range of the segment and its
//   split among the waves), then every wave walks the chunk's stages on its
//   own, with no block barrier:
//     L  leaves of p whose n = x/(p m) falls in the wave's bits, all members'
//        in one list: m odd with mu(m) != 0 and lpf(m) > p from the mu*lpf
//        table (p^2 <= y) or m prime (p^2 > y); the wave publishes its
//        in-wave prefix counts and counts phi_wave(n) = in-wave prefix of the
//        owning lane + popcounts of its words below n; sum mu * phi_wave per
//        thread and member, sum mu per wave and member;
//     X  cross off the odd multiples of p in the thread's own bits, recording
//        each wave's removals.
//   At the end of a chunk the block adds the survivors of the earlier waves
//   to every leaf's count (from the per-stage, per-wave removals and sums of
//   mu), and writes T_b (survivors before stage b) and sum_b(mu); the carry
//   kernel adds sum_b prefix_b(seg) * sum_b(mu) over the segments (the
//   decomposition the reference does with scanVectorized/addMultiply,
//   src/sieve/S2_hard_device.cu:357-375).
//
// P2: the segments holding x/p for a member's P2 primes p in (y, sqrt x]
// get, on replica 0 of their tile, the rest of a prime sieve (primes above
// the last stage up to sqrt(hi)) and each such member's P2 counts; see
// p2_epilogue.
#pragma once
#include "cpx_device.hpp"

#define HARD_SCH 128       // items (stage x member slot) per chunk
#define HARD_SST 64        // stages per chunk at most
#ifndef P2_WAVE_P
#define P2_WAVE_P 1024     // P2 sieve: primes below this are crossed by a whole wave
#endif
static_assert(CPX_GROUP == 1 || CPX_GROUP == 2 || CPX_GROUP == 4 || CPX_GROUP == 8,
              "member slots are powers of two, so a leafy-mask word holds whole stages");

// A group member's share of the current segment (zero when its hard range
// ends before the segment).
struct GroupMember {
  uint64_t y;
  uint64_t stage_off;   // its stage table
  uint64_t toff;        // hard_T index of its stage 0 in this segment's row
  uint64_t moff;        // hard_M index of its stage 0 for this replica
  uint64_t acc_slot;    // hard_acc index for this replica
  uint64_t row;         // its hard row (hard_bact)
  uint32_t nst;         // its active stages in the segment
  uint32_t on;          // the segment is in its hard range
  uint32_t p2;          // ... and in its P2 range
  uint32_t pad;
};

template <int T, int K>
struct SegShared {
  uint32_t ipre[T];              // in-wave exclusive prefix of each thread's count
  uint32_t wtot[T / CPX_WAVE];   // wave totals
  int32_t wsum[T / CPX_WAVE];    // per-wave sums (P2)
  unsigned long long red_u[T / CPX_WAVE];
  long long red_s[CPX_GROUP][T / CPX_WAVE];
  GroupMember gm[CPX_GROUP];
  // each wave's leaf list of its current stage: member g's entries start at
  // list offset lpre[w][g], with candidate index lia[w][g]
  __attribute__((aligned(16))) uint32_t lpre[T / CPX_WAVE][CPX_GROUP];
  uint32_t lia[T / CPX_WAVE][CPX_GROUP];
  union {
    struct {                     // hard: chunk table and per-wave counters
      // per item i = k * gs + g (stage k of the chunk, member g; gs member
      // slots, a power of two)
      uint64_t xp[HARD_SCH];     // floor(x_g / p_k), low 64 bits
      uint16_t xph[HARD_SCH];    // ... and bits 64-79 (nonzero only for x above 2^68)
      uint32_t cbase[HARD_SCH];  // leaf candidates: mtop/2 (dense) or first prime index
      uint32_t ccnt[HARD_SCH];   // candidates | HARD_WIDE | HARD_SPARSE
      int32_t wm[HARD_SCH][T / CPX_WAVE];    // sum of mu of each wave's leaves
      uint16_t wb[HARD_SCH][T / CPX_WAVE];   // first candidate of each wave (non-wide items)
      uint32_t lmask[HARD_SCH / 32];         // items with candidates
      // per stage k
      uint32_t p[HARD_SST];
      float rinv[HARD_SST];      // 1 / p
      uint32_t oseg[HARD_SST];   // first odd multiple of p in the segment (bit offset)
      uint16_t dw[HARD_SST][T / CPX_WAVE];   // survivors each wave removes
      uint32_t c0[T / CPX_WAVE]; // each wave's survivors at the chunk start
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
// bit, counted from the nearer half boundary.  The K/2 words of that half,
// the packed word and the partial word are loaded before any is used (one
// LDS round trip); words at or above the bit are masked out, and at the end
// of a thread's bits the partial word is its padding word K, masked to nothing.
template <int T, int K>
__device__ __forceinline__ uint32_t count_below(const uint32_t *lds, const SegShared<T, K> &sh,
                                                const uint32_t (&wp)[T / CPX_WAVE], uint32_t bit)
{
  constexpr uint32_t BPT = 32 * K, H = K / 2;
  uint32_t owner = bit / BPT;
  uint32_t jb = bit - owner * BPT;
  if (owner >= (uint32_t)T) { owner = T - 1; jb = BPT; }
  uint32_t base = 0;
#pragma unroll
  for (uint32_t w = 0; w < (uint32_t)(T / CPX_WAVE); w++) base = (owner / CPX_WAVE == w) ? wp[w] : base;
  const bool upper = jb >= BPT / 2;
  const uint32_t *ow = lds + owner * (K + 1) + (upper ? H : 0u);
  const uint32_t nf = (jb >> 5) - (upper ? H : 0u);  // whole words below the bit: 0 .. H
  uint32_t wd[H];
#pragma unroll
  for (uint32_t r = 0; r < H; r++) wd[r] = ow[r];
  const uint32_t part = ow[nf];
  const uint32_t e = sh.ipre[owner];
  uint32_t c = base + (e & CountPack<K>::PRE_MASK) + (upper ? e >> CountPack<K>::PRE_BITS : 0u);
#pragma unroll
  for (uint32_t r = 0; r < H; r++) c += cpx_popc32(r < nf ? wd[r] : 0u);
  return c + cpx_popc32(part & cpx_lowmask32(jb & 31u));
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

// Table of one chunk of stages [s0, s0 + clen) (all members share the stage
// primes p_k = primes[c + s0 + k]).  Per stage: p, 1/p and the first odd
// multiple of p in the segment (bit offset).  Per item (stage k, member g)
// active for g (s0 + k below its active stages): x_g/p, the leaf candidate
// range, a leafy bit, and for every wave w > 0 the boundary b_w of its
// candidates (wave w owns the leaves with n = x/(p m) in
// [lo_w, lo_w + 4*32K*64)): for dense items (m = mtop - 2i, n grows with i)
// b_w counts the candidates with n < lo_w and is the wave's first index; for
// sparse items (m = the i-th prime, n falls with i) it counts those with
// n >= lo_w and is the wave's end.  Clears the per-wave sums of mu.  Caller
// syncs before and after.
template <int T, int K, int GS>
__device__ __forceinline__ void hard_chunk_table(const BatchDev &bd, uint32_t c, uint32_t s0,
                                                 uint32_t clen, uint64_t lo, uint64_t hi,
                                                 double inv_lo, double inv_hi, SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint64_t WB = (uint64_t)CPX_WAVE * 32 * K;
  const uint32_t nit = clen * GS;
  for (uint32_t k = threadIdx.x; k < nit * NW; k += T) (&sh.u.st.wm[0][0])[k] = 0;
  for (uint32_t k = threadIdx.x; k < clen; k += T) {
    const uint64_t p = bd.primes[c + s0 + k];
    sh.u.st.p[k] = (uint32_t)p;
    sh.u.st.rinv[k] = 1.0f / (float)p;
    // first odd multiple of p at or above lo + 1, as a bit offset (< p)
    const uint64_t start = lo + 1;
    const uint64_t rr = start - cpx_div_recip(start, p, 1.0 / (double)p) * p;
    uint64_t mm = start + (rr ? p - rr : 0);
    if (!(mm & 1)) mm += p;
    sh.u.st.oseg[k] = (uint32_t)((mm - start) >> 1);
  }
  for (uint32_t it = threadIdx.x; it < nit; it += T) {
    const uint32_t k = it / GS, g = it % GS;
    const GroupMember &m = sh.gm[g];
    uint32_t ncand = 0, cbase = 0;
    bool sparse = false;
    uint64_t xp = 0;
    uint32_t xph = 0;
    if (s0 + k < m.nst) {
      const StageDev sd = bd.stages[m.stage_off + s0 + k];
      const uint64_t p = sd.p, y = m.y;
      xp = sd.xp;
      xph = sd.xph;
      // m range of this segment's leaves of p: x/(p hi) < m <= x/(p lo), m > y/p
      sparse = p * p > y;
      uint64_t mhi = y;
      if (lo != 0 && cpx_quot_below_2_50(xp, xph, lo)) {
        const uint64_t v = xph ? cpx_div_recip_w(xp, xph, lo, inv_lo) : cpx_div_recip(xp, lo, inv_lo);
        if (v < mhi) mhi = v;
      }
      uint64_t mlo = ~0ull;
      if (cpx_quot_below_2_50(xp, xph, hi))
        mlo = xph ? cpx_div_recip_w(xp, xph, hi, inv_hi) : cpx_div_recip(xp, hi, inv_hi);
      if (sparse && mhi > sd.xp3) mhi = sd.xp3;        // q above x/p^3: pi-table lookups
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
    }
    sh.u.st.xp[it] = xp;
    sh.u.st.xph[it] = (uint16_t)xph;
    sh.u.st.cbase[it] = cbase;
    sh.u.st.ccnt[it] = ncand | (sparse ? HARD_SPARSE : 0u) | (ncand > 0xffffu ? HARD_WIDE : 0u);
    if (ncand) atomicOr(&sh.u.st.lmask[it >> 5], 1u << (it & 31));
  }
  if (NW == 1) return;
  __syncthreads();
  // n >= lo_w <=> m <= M = floor(x/p / lo_w)
  for (uint32_t idx = threadIdx.x; idx < nit * NW; idx += T) {
    const uint32_t it = idx / NW, w = idx % NW;
    const uint32_t cc = sh.u.st.ccnt[it];
    if (w == 0 || !(cc & HARD_NC_MASK) || (cc & HARD_WIDE)) continue;
    const uint32_t ncand = cc & HARD_NC_MASK, cb = sh.u.st.cbase[it];
    const uint64_t xp = sh.u.st.xp[it];
    const uint32_t xph = sh.u.st.xph[it];
    const uint64_t lo_w = lo + 2 * (uint64_t)w * WB;          // > 0
    // (a quotient of 2^50 or more exceeds every m <= y: no n below lo_w)
    const uint64_t M = !cpx_quot_below_2_50(xp, xph, lo_w) ? ~0ull
                       : xph ? cpx_div_recip_w(xp, xph, lo_w, 1.0 / (double)lo_w)
                             : cpx_div_recip(xp, lo_w, 1.0 / (double)lo_w);
    uint64_t b;
    if (!(cc & HARD_SPARSE)) {
      const uint64_t mtop = 2 * (uint64_t)cb + 1;            // m = mtop - 2i
      b = mtop > M ? (mtop - M + 1) / 2 : 0;
    } else {
      const uint64_t y = sh.gm[it % GS].y;
      const uint64_t pm = pi_lookup(bd.pitab, M < y ? M : y);
      b = pm > cb ? pm - cb : 0;
    }
    sh.u.st.wb[it][w] = (uint16_t)(b < ncand ? b : ncand);
  }
}

// Clear bits j, j + p, ... below lim of a thread's own words, adding to d
// the bits that were still set; returns the first multiple at or above lim.
// Four and-with-returns are issued before their results are used, so each
// lane keeps four LDS round trips in flight (the LDS applies one lane's
// operations in order, so multiples that share a word still see each
// other).  ubfe and the shift take the bit offset modulo 32.
__device__ __forceinline__ uint32_t cross_own_run(uint32_t *own, uint32_t j, uint32_t lim, uint32_t p,
                                                  uint32_t &d)
{
  for (; j + 3 * p < lim; j += 4 * p) {
    const uint32_t j1 = j + p, j2 = j1 + p, j3 = j2 + p;
    const uint32_t o0 = atomicAnd(&own[j >> 5], ~(1u << (j & 31u)));
    const uint32_t o1 = atomicAnd(&own[j1 >> 5], ~(1u << (j1 & 31u)));
    const uint32_t o2 = atomicAnd(&own[j2 >> 5], ~(1u << (j2 & 31u)));
    const uint32_t o3 = atomicAnd(&own[j3 >> 5], ~(1u << (j3 & 31u)));
    d += __builtin_amdgcn_ubfe(o0, j, 1u) + __builtin_amdgcn_ubfe(o1, j1, 1u) +
         __builtin_amdgcn_ubfe(o2, j2, 1u) + __builtin_amdgcn_ubfe(o3, j3, 1u);
  }
  for (; j < lim; j += p) d += __builtin_amdgcn_ubfe(atomicAnd(&own[j >> 5], ~(1u << (j & 31u))), j, 1u);
  return j;
}

// First leafy item of the chunk at or after item k (HARD_SCH if none).
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

// This wave's candidate index range [ia, ib) of chunk item k (of a member
// with this y).  Items with 2^16 candidates or more derive it here:
// n = x/(p m) in [lo_w, hi_w) <=> floor(xp / hi_w) < m <= floor(xp / lo_w).
template <int T, int K>
__device__ __forceinline__ void hard_wave_range(const BatchDev &bd, uint64_t y,
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
  const uint32_t xph = sh.u.st.xph[k];
  // (quotients of 2^50 or more exceed every m <= y)
  const uint64_t Mlo = !(lo_w && cpx_quot_below_2_50(xp, xph, lo_w)) ? ~0ull
                       : xph ? cpx_div_recip_w(xp, xph, lo_w, inv_lo_w) : cpx_div_recip(xp, lo_w, inv_lo_w);
  const uint64_t Mhi = !cpx_quot_below_2_50(xp, xph, hi_w) ? ~0ull
                       : xph ? cpx_div_recip_w(xp, xph, hi_w, inv_hi_w) : cpx_div_recip(xp, hi_w, inv_hi_w);
  if (!(cc & HARD_SPARSE)) {                                 // m = mtop - 2i
    const uint64_t mtop = 2 * (uint64_t)cb + 1;
    const uint64_t a = mtop > Mlo ? (mtop - Mlo + 1) / 2 : 0;
    const uint64_t b = mtop > Mhi ? (mtop - Mhi + 1) / 2 : 0;
    ia = (uint32_t)min(a, (uint64_t)ncand);
    ib = (uint32_t)min(b, (uint64_t)ncand);
  } else {                                                   // primes[cb + i]
    const uint64_t pl = pi_lookup(bd.pitab, Mhi < y ? Mhi : y);
    const uint64_t ph = pi_lookup(bd.pitab, Mlo < y ? Mlo : y);
    ia = (uint32_t)(pl > cb ? min(pl - cb, (uint64_t)ncand) : 0);
    ib = (uint32_t)(ph > cb ? min(ph - cb, (uint64_t)ncand) : 0);
  }
}

// Survivors among the first `gw` bits (0 <= gw <= 64 * 32K) of this wave's
// part of the segment: in-wave prefix of the owning lane + its words below,
// from the packed prefix/half-count words the wave published.  All K/2 words
// of the owner's half, the packed word and the partial word are loaded before
// any is used, so the count costs one LDS round trip; the words at or above
// the bit are masked out.  (For gw at the end of a lane's bits the partial
// word is the owner's padding word K, masked to nothing.)
template <int T, int K>
__device__ __forceinline__ uint32_t count_in_wave(const uint32_t *lds, const SegShared<T, K> &sh,
                                                  uint32_t wv, uint32_t gw)
{
  constexpr uint32_t BPT = 32 * K, H = K / 2;
  uint32_t ol = gw / BPT;
  uint32_t jb = gw - ol * BPT;
  if (ol >= (uint32_t)CPX_WAVE) { ol = CPX_WAVE - 1; jb = BPT; }
  const uint32_t owner = wv * CPX_WAVE + ol;
  const bool upper = jb >= BPT / 2;
  const uint32_t *ow = lds + owner * (K + 1) + (upper ? H : 0u);
  const uint32_t nf = (jb >> 5) - (upper ? H : 0u);  // whole words below the bit: 0 .. H
  uint32_t wd[H];
#pragma unroll
  for (uint32_t r = 0; r < H; r++) wd[r] = ow[r];
  const uint32_t part = ow[nf];
  const uint32_t e = sh.ipre[owner];
  uint32_t c = (e & CountPack<K>::PRE_MASK) + (upper ? e >> CountPack<K>::PRE_BITS : 0u);
#pragma unroll
  for (uint32_t r = 0; r < H; r++) c += cpx_popc32(r < nf ? wd[r] : 0u);
  return c + cpx_popc32(part & cpx_lowmask32(jb & 31u));
}

// This wave's leaves of chunk stage k as one list over the stage's leafy
// members (lbits): lane g < GS finds member g's candidate range [a, b) and
// takes list offsets [my_lo, my_hi) = [pre_g, pre_g + b - a); the offsets
// and starts go to the wave's list slots.  Returns the list length
// (wave-uniform).  All lanes call.
template <int T, int K, int GS>
__device__ __forceinline__ uint32_t hard_lane_list(const BatchDev &bd, SegShared<T, K> &sh, uint32_t k,
                                                   uint32_t lbits, uint32_t wv, uint32_t lane,
                                                   uint64_t lo_w, uint64_t hi_w, double inv_lo_w,
                                                   double inv_hi_w, uint32_t &my_lo, uint32_t &my_hi,
                                                   uint32_t &ia1)
{
  if (GS == 1) {                       // one member: every lane derives the range, kept in ia1
    uint32_t a = 0, b = 0;
    if (lbits & 1u) hard_wave_range<T, K>(bd, sh.gm[0].y, sh, k, wv, lo_w, hi_w, inv_lo_w, inv_hi_w, a, b);
    ia1 = a;
    my_lo = 0;
    my_hi = b > a ? b - a : 0u;
    return my_hi;
  }
  uint32_t a = 0, n = 0;
  if (lane < GS && ((lbits >> lane) & 1u)) {
    uint32_t b;
    hard_wave_range<T, K>(bd, sh.gm[lane].y, sh, k * GS + lane, wv, lo_w, hi_w, inv_lo_w, inv_hi_w, a, b);
    n = b > a ? b - a : 0u;
  }
  const uint32_t incl = wave_incl_scan(n);
  my_lo = incl - n;
  my_hi = incl;
  if (lane < GS) {
    sh.lpre[wv][lane] = my_lo;
    sh.lia[wv][lane] = a;
  }
  return (uint32_t)__builtin_amdgcn_readlane((int)incl, GS - 1);
}

// Member g and candidate index i of entry f of this wave's list (ia1: the
// candidate start of a one-member list).
template <int T, int K, int GS>
__device__ __forceinline__ uint32_t hard_lane_entry(const SegShared<T, K> &sh, uint32_t wv, uint32_t f,
                                                    uint32_t ia1, uint32_t &i)
{
  if (GS == 1) {
    i = ia1 + f;
    return 0;
  }
  uint32_t g = 0;
  if (GS > 1) {
    uint32_t pre[GS];
#pragma unroll
    for (uint32_t h = 0; h < GS; h++) pre[h] = sh.lpre[wv][h];
#pragma unroll
    for (uint32_t h = 1; h < GS; h++) g += f >= pre[h] ? 1u : 0u;
  }
  i = sh.lia[wv][g] + (f - sh.lpre[wv][g]);
  return g;
}

// Add one window of list entries [fb, fb + 64) (lane l holds entry fb + l:
// val = mu * phi_wave, sg = mu, zero past the list) to the sums of member
// lanes: lane g < GS gains the sums over its entries [my_lo, my_hi).  One
// wave scan of val * 256 + sg (|val| <= 2^16, so a window's sum fits 31
// bits and its low byte is the sum of sg), read at each member's ends.
__device__ __forceinline__ void hard_lane_accumulate(int32_t val, int32_t sg, uint32_t fb, uint32_t my_lo,
                                                     uint32_t my_hi, int64_t &acc, int32_t &wmv)
{
  const uint32_t S = wave_incl_scan((uint32_t)(val * 256 + sg));
  // list offsets are below 2^30: an item has at most y/2 candidates, and
  // groups of more than one member have y < 2^28
  const int32_t lo = min(max((int32_t)my_lo - (int32_t)fb, 0), CPX_WAVE);
  const int32_t hi = min(max((int32_t)my_hi - (int32_t)fb, 0), CPX_WAVE);
  const uint32_t shi = (uint32_t)__builtin_amdgcn_ds_bpermute((hi > 0 ? hi - 1 : 0) * 4, (int)S);
  const uint32_t slo = (uint32_t)__builtin_amdgcn_ds_bpermute((lo > 0 ? lo - 1 : 0) * 4, (int)S);
  const int32_t d = (int32_t)((hi > 0 ? shi : 0u) - (lo > 0 ? slo : 0u));
  const int32_t sgs = (int32_t)((uint32_t)d << 24) >> 24;
  acc += (d - sgs) >> 8;
  wmv += sgs;
}

// Leafy-mask bits of chunk stage k (one per member slot).
template <int T, int K, int GS>
__device__ __forceinline__ uint32_t hard_stage_bits(const SegShared<T, K> &sh, uint32_t k)
{
  const uint32_t it0 = k * GS;
  return (sh.u.st.lmask[it0 >> 5] >> (it0 & 31)) & (GS == 32 ? ~0u : (1u << GS) - 1u);
}

// End of a chunk of stages [s0, s0 + n) (all threads call and keep
// `running`): writes, for every member g and its active stages b, T_b =
// survivors of the segment before stage b (replica 0) and this replica's sum
// of mu, and adds to acc[g] the part of every leaf's phi that the in-wave
// count left out -- the survivors of the earlier waves at that stage:
//   sum_k sum_w wm[k,g][w] * (A_w - sum_{j<k} sum_{w'<w} dw[j][w'])
// with A_w the earlier waves' survivors at the chunk start.  Then resets the
// chunk's leafy mask.
template <int T, int K, int GS>
__device__ __forceinline__ void hard_flush(const BatchDev &bd, SegShared<T, K> &sh, uint32_t s0,
                                           uint32_t n, uint32_t &running, bool write_T, int64_t &acc)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  const uint32_t lane = lane_id(), wv = wave_id();
  __syncthreads();
  if (NW > 1) {
    uint32_t A = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) A += w < wv ? sh.u.st.c0[w] : 0u;
    if (GS == 1) {                     // sums per lane: lane k takes stage k
      uint32_t carry = 0;
      for (uint32_t k0 = 0; k0 < n; k0 += CPX_WAVE) {
        const uint32_t k = k0 + lane;
        uint32_t Gk = 0;
        if (k < n) {
#pragma unroll
          for (uint32_t w = 0; w < NW; w++) Gk += w < wv ? (uint32_t)sh.u.st.dw[k][w] : 0u;
        }
        const uint32_t incl = wave_incl_scan(Gk);
        if (k < n) acc += (int64_t)sh.u.st.wm[k][wv] * ((int64_t)A - (int64_t)(carry + incl - Gk));
        carry += (uint32_t)__builtin_amdgcn_readlane((int)incl, CPX_WAVE - 1);
      }
    } else {                           // sums in member lanes: the earlier waves' removals
      // before each stage, R_k, go to this wave's part of the boundary table
      // (no wave reads it after the first barrier; the next chunk rewrites it)
      uint32_t *Rk = reinterpret_cast<uint32_t *>(&sh.u.st.wb[0][0]) + wv * CPX_WAVE;
      const uint32_t k = lane;         // n <= 64 stages
      uint32_t Gk = 0;
      if (k < n) {
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) Gk += w < wv ? (uint32_t)sh.u.st.dw[k][w] : 0u;
      }
      const uint32_t incl = wave_incl_scan(Gk);
      if (k < n) Rk[k] = incl - Gk;
      if (lane < GS) {
        for (uint32_t j = 0; j < n; j++)
          acc += (int64_t)sh.u.st.wm[j * GS + lane][wv] * ((int64_t)A - (int64_t)Rk[j]);
      }
    }
  }
  // T_b and the sums of mu per stage: wave w writes those of members
  // w, w + NW, ...; every wave keeps `running`
  for (uint32_t k0 = 0; k0 < n; k0 += CPX_WAVE) {
    const uint32_t k = k0 + lane;
    uint32_t d = 0;
    if (k < n) {
#pragma unroll
      for (uint32_t w = 0; w < NW; w++) d += sh.u.st.dw[k][w];
    }
    const uint32_t incl = wave_incl_scan(d);
    if (k < n) {
      const uint32_t Tk = running - (incl - d);
      for (uint32_t g = wv; g < GS; g += NW) {
        const GroupMember &m = sh.gm[g];
        if (s0 + k >= m.nst) continue;
        int32_t ms = 0;
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) ms += sh.u.st.wm[k * GS + g][w];
        if (write_T) st_out(&bd.hard_T[m.toff + s0 + k], (uint32_t)Tk);
        st_out(&bd.hard_M[m.moff + s0 + k], (int32_t)ms);
      }
    }
    running -= (uint32_t)__builtin_amdgcn_readlane((int)incl, CPX_WAVE - 1);
  }
  // no wave reads the leafy mask after the first barrier
  if (threadIdx.x < HARD_SCH / 32) sh.u.st.lmask[threadIdx.x] = 0;
  __syncthreads();
}

// P2 leaves of a segment (replaces the windowed upper_bound count of
// src/P2.cu:76-133), run by replica 0 of a hard tile after its stages: the
// bitmap already holds [lo, hi) sieved by the wheel and the stage primes
// p_(c+1) .. p_kfirst; the primes from index kfirst with p^2 < hi are
// crossed here (from p^2), which leaves exactly the primes above the largest
// p2_e of the members.  Then, for each member whose P2 range holds the
// segment and each of its P2 primes p in (max(y, x/hi), min(sqrt x, x/lo)]
// (pi-table bits, block-strided), the count of primes in (p2_e, x/p] held in
// this segment; per member and segment: primes in (p2_e, hi), number of p,
// sum of the counts.  The carry is applied in the combine kernel.
template <int T, int K>
__device__ __forceinline__ void p2_epilogue(const BatchDev &bd, const GroupDev &grp, uint64_t seg,
                                            uint64_t lo, uint32_t kfirst, uint32_t *lds,
                                            SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t hi = lo + 2 * SB;
  const uint32_t *__restrict__ sp = bd.primes;
  const uint32_t np_all = (uint32_t)bd.nprimes;

  __syncthreads();
  // primes below P2_WAVE_P hit the segment more than SB / P2_WAVE_P times:
  // one prime per wave, its lanes taking every 64th multiple, so no lane
  // walks a long chain alone
  for (uint32_t k = kfirst + (uint32_t)__builtin_amdgcn_readfirstlane((int)wv); k < np_all; k += NW) {
    const uint32_t p = sp[k];
    if (p >= P2_WAVE_P || (uint64_t)p * p >= hi) break;
    const uint32_t o = (uint32_t)first_odd_multiple_bit_r(lo, p, (uint64_t)p * p);
    // o < SB + p and 64 p < 2^16: no 32-bit overflow
    for (uint32_t j = o + lane * p; j < (uint32_t)SB; j += CPX_WAVE * p) {
      const uint32_t g = j >> 5;
      atomicAnd(&lds[g + g / K], ~(1u << (j & 31)));
    }
  }
  // the rest: one prime per lane
  for (uint32_t k = kfirst + t; k < np_all; k += T) {
    const uint32_t p = sp[k];
    if ((uint64_t)p * p >= hi) break;
    if (p < P2_WAVE_P) continue;
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
  for (uint32_t g = 0; g < grp.n; g++) {
    if (!sh.gm[g].p2) continue;
    const QueryDev &q = bd.queries[grp.q[g]];
    const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
    // x / p < z < 2^50 for every P2 prime: the reciprocal divide when p < 2^32
    const bool fastdiv = q.sqrtx < (1ull << 32);
    const double xf = (double)x;
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
          const uint64_t v = fastdiv ? cpx_div_leaf(q.x_lo, xf, (uint32_t)p)
                                     : cpx_div128_to64(x, p);   // lo <= v < hi
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
      st_out(&bd.p2_cnt[row], (uint32_t)(tot - ce));
      st_out(&bd.p2_np[row], (uint32_t)N);
      st_out(&bd.p2_loc[row], (uint64_t)L);
    }
    __syncthreads();
  }
}

// The stage chunks of one tile with GS member slots (items k * GS + g):
// leaves, cross-off and chunk flushes; leaves the members' sums of
// mu * phi in sh.red_s.
template <int T, int K, int GS>
__device__ __forceinline__ void hard_chunks(const BatchDev &bd, SegShared<T, K> &sh, uint32_t *lds,
                                            uint32_t c, uint32_t nst, uint32_t rep, uint32_t cstep,
                                            uint64_t lo, uint64_t hi, uint64_t lo_w, uint32_t cnt,
                                            uint32_t cnt_lo)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  constexpr uint32_t WB = CPX_WAVE * BPT;
  constexpr uint32_t SST = HARD_SST < HARD_SCH / GS ? HARD_SST : HARD_SCH / GS;   // stages per chunk
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t hi_w = lo_w + 2 * (uint64_t)WB;
  const double inv_lo_w = lo_w ? 1.0 / (double)lo_w : 0.0, inv_hi_w = 1.0 / (double)hi_w;
  const double inv_lo = lo ? 1.0 / (double)lo : 0.0, inv_hi = 1.0 / (double)hi;
  const uint32_t b0 = t * BPT;                          // < 2^20
  const float fb0 = (float)b0;
  uint32_t s0 = 0, clen = min(SST, nst);
  if (clen) hard_chunk_table<T, K, GS>(bd, c, s0, clen, lo, hi, inv_lo, inv_hi, sh);
  __syncthreads();
  uint32_t running = 0;                                 // survivors before the next stage
  for (uint32_t w = 0; w < NW; w++) running += (uint32_t)sh.wsum[w];

  // sums of mu * phi_wave: per lane (GS = 1), or member g's in lane g
  int64_t acc = 0;
  while (clen) {
    {
      const uint32_t ws = wave_sum(cnt);                // this wave's survivors at the chunk start
      if (lane == 0) sh.u.st.c0[wv] = ws;
    }
    // one-ahead prefetch: the next leafy stage's list and this lane's first
    // candidate there
    const uint32_t f0 = rep * CPX_WAVE + lane;          // this lane's first list entry
    uint32_t nk = hard_next_leafy<T, K>(sh, 0) / GS;   // next leafy stage (>= clen: none)
    uint32_t my_lo = 0, my_hi = 0, ntot = 0, ia1 = 0;
    int32_t pf = 0;
    if (nk < clen) {
      ntot = hard_lane_list<T, K, GS>(bd, sh, nk, hard_stage_bits<T, K, GS>(sh, nk), wv, lane, lo_w, hi_w,
                                      inv_lo_w, inv_hi_w, my_lo, my_hi, ia1);
      if (f0 < ntot) {
        uint32_t i;
        const uint32_t it = nk * GS + hard_lane_entry<T, K, GS>(sh, wv, f0, ia1, i);
        pf = hard_fetch(bd, sh.u.st.cbase[it], sh.u.st.ccnt[it] & HARD_SPARSE, i);
      }
    }
    for (uint32_t k = 0; k < clen; k++) {
      const uint32_t it0 = k * GS;
      const uint32_t p = sh.u.st.p[k];
      if (k == nk) {
        // ---- leaves of p whose n falls in this wave's bits (state: crossed
        //      by p_1 .. p_{b-1}), all members' in one list; the wave
        //      publishes its own prefix, no other wave is involved
        if (ntot > rep * CPX_WAVE) {                    // (a wave with no entries reads none)
          const uint32_t incl = wave_incl_scan(cnt);
          sh.ipre[t] = (incl - cnt) | (cnt_lo << CountPack<K>::PRE_BITS);
        }
        int32_t wmv = 0;                                // sum of mu: the wave's (GS = 1) or member lane's
        int32_t v = pf;
        for (uint32_t fb = rep * CPX_WAVE; fb < ntot; fb += cstep) {   // every lane runs every window
          const uint32_t f = fb + lane;
          int32_t val = 0, sg = 0;
          if (f < ntot) {
            uint32_t i;
            const uint32_t it = it0 + hard_lane_entry<T, K, GS>(sh, wv, f, ia1, i);
            const uint32_t cc = sh.u.st.ccnt[it], cb = sh.u.st.cbase[it];
            const bool sparse = cc & HARD_SPARSE;
            if (fb != rep * CPX_WAVE) v = hard_fetch(bd, cb, sparse, i);
            uint32_t m;
            if (sparse) {
              m = (uint32_t)v;                                // prime, mu = -1
              sg = -1;
            } else {                                          // mu(m) = 0 or lpf(m) <= p: no leaf
              const uint32_t av = (uint32_t)(v < 0 ? -v : v);
              m = 2 * (cb - i) + 1;
              sg = av <= p ? 0 : (v > 0 ? 1 : -1);
            }
            if (sg) {
              const uint64_t xp = sh.u.st.xp[it];
              const double xpf = (double)xp + (double)sh.u.st.xph[it] * 0x1p64;
              const uint64_t n = cpx_div_leaf(xp, xpf, m);   // lo_w <= n < hi_w
              const int32_t cn = (int32_t)count_in_wave<T, K>(lds, sh, wv,
                                                              ((uint32_t)n - (uint32_t)lo_w + 1u) >> 1);
              val = sg > 0 ? cn : -cn;
            }
          }
          if (GS == 1) {
            acc += val;
            wmv += (int32_t)__popcll(__ballot(sg > 0)) - (int32_t)__popcll(__ballot(sg < 0));
          } else {
            hard_lane_accumulate(val, sg, fb, my_lo, my_hi, acc, wmv);
          }
        }
        if (lane < GS) sh.u.st.wm[it0 + lane][wv] = wmv;
        // next leafy stage
        nk = hard_next_leafy<T, K>(sh, it0 + GS) / GS;
        if (nk < clen) {
          ntot = hard_lane_list<T, K, GS>(bd, sh, nk, hard_stage_bits<T, K, GS>(sh, nk), wv, lane, lo_w,
                                          hi_w, inv_lo_w, inv_hi_w, my_lo, my_hi, ia1);
          if (f0 < ntot) {
            uint32_t i;
            const uint32_t it = nk * GS + hard_lane_entry<T, K, GS>(sh, wv, f0, ia1, i);
            pf = hard_fetch(bd, sh.u.st.cbase[it], sh.u.st.ccnt[it] & HARD_SPARSE, i);
          }
        }
      }
      // ---- cross off the odd multiples of p in this thread's own bits,
      //      from (t * 32K) mod p by the f32 reciprocal (t * 32K < 2^20, so
      //      the truncated estimate is the quotient or one off either way)
      {
        const uint32_t oseg = sh.u.st.oseg[k];
        uint32_t tb = b0;
        if (p <= b0) {                                    // p < 2^20 here
          // b0 mod p in f32: b0 < 2^20 and q p < 2^21 are exact, so the fma
          // gives b0 - q p exactly, with q the quotient or one off
          const float q = truncf(fb0 * sh.u.st.rinv[k]);
          int32_t r = (int32_t)fmaf(-q, (float)p, fb0);
          r += r < 0 ? (int32_t)p : 0;
          r -= r >= (int32_t)p ? (int32_t)p : 0;
          tb = (uint32_t)r;
        }
        uint32_t j = oseg >= tb ? oseg - tb : oseg + p - tb;
        uint32_t *const own = lds + t * (K + 1);
        uint32_t wd;
        // (one LDS and-with-return per multiple: the old word says whether
        // the bit was still set; the words are the thread's own)
        if (p > BPT) {                                    // at most one multiple per thread
          uint32_t hit = 0;
          if (j < BPT) {
            const uint32_t old = atomicAnd(&own[j >> 5], ~(1u << (j & 31u)));
            hit = (old >> (j & 31u)) & 1u;
          }
          cnt -= hit;
          cnt_lo -= hit & (j < BPT / 2 ? 1u : 0u);
          wd = (uint32_t)__popcll(__ballot(hit));
        } else {
          // the first half's multiples, then the second's: the removals of
          // the first half are the running sum at the switch
          uint32_t d = 0;
          j = cross_own_run(own, j, BPT / 2, p, d);
          cnt_lo -= d;
          cross_own_run(own, j, BPT, p, d);
          cnt -= d;
          wd = wave_sum(d);
        }
        if (lane == 0) sh.u.st.dw[k][wv] = (uint16_t)wd;   // this wave's removals
      }
    }
    hard_flush<T, K, GS>(bd, sh, s0, clen, running, rep == 0, acc);
    s0 += clen;
    clen = min(SST, nst - s0);
    if (clen) {
      hard_chunk_table<T, K, GS>(bd, c, s0, clen, lo, hi, inv_lo, inv_hi, sh);
      __syncthreads();
    }
  }
  if (GS == 1) {
    const long long a = wave_sum((long long)acc);
    if (lane == 0) sh.red_s[0][wv] = a;
  } else if (lane < GS) {
    sh.red_s[lane][wv] = acc;
  }
}

// One tile: segment `seg` of a group, replica `rep` of `nrep`.  The wheel
// and the stage primes sieve the segment once; every member's leaves are
// evaluated against that one bitmap.
template <int T, int K>
__device__ __forceinline__ void hard_segment(const BatchDev &bd, const TileDev &tile, uint32_t *lds,
                                             SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  constexpr uint32_t WB = CPX_WAVE * BPT;                // bits per wave
  constexpr uint32_t G = CPX_GROUP;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t seg = tile.seg;
  const uint64_t lo = seg * 2 * SB;
  const uint64_t hi = lo + 2 * SB;
  const GroupDev &grp = bd.groups[tile.group];
  const uint32_t rep = tile.rep;
  const uint32_t c = bd.queries[grp.q[0]].c;           // members share the wheel
  const WheelInfo wi = bd.winfo[c];
  // members' rows of this segment: counters of the active stages only (the
  // plan sized each row for ri.act stages)
  if (t < G) {
    GroupMember m{};
    if (t < grp.n) {
      const QueryDev &q = bd.queries[grp.q[t]];
      if (seg < q.hard_nseg) {
        const uint64_t row = q.hard_row0 + seg;
        const HardRowInfo ri = bd.rowinfo[row];
        m.y = q.y;
        m.stage_off = q.stage_off;
        m.toff = q.tm_off + ri.toff - bd.pass_toff;
        m.moff = q.m_off + ri.moff - bd.pass_moff + (uint64_t)rep * ri.act;
        m.acc_slot = q.mrow_base + ri.mrow + rep;
        m.row = row;
        m.nst = min(q.nstage, ri.act);
        m.on = 1;
        m.p2 = seg >= q.p2_seg0 && seg < q.p2_seg0 + q.p2_nseg;
      }
    }
    sh.gm[t] = m;
  }
  const uint64_t lo_w = lo + 2 * (uint64_t)wv * WB;     // this wave's integers from lo_w
  const uint32_t cstep = CPX_WAVE * (uint32_t)tile.nrep;   // candidate stride of a lane

  seg_fill_wheel(lds, K, T, lo, bd.wheel + wi.off, wi.period);
  if (t < HARD_SCH / 32) sh.u.st.lmask[t] = 0;
  __syncthreads();
  uint32_t nst = 0, p2any = 0;                          // stages of the segment: the members' most
#pragma unroll
  for (uint32_t g = 0; g < G; g++) { nst = max(nst, sh.gm[g].nst); p2any |= sh.gm[g].p2; }
  uint32_t cnt_lo = own_count_lo<K>(lds, t), cnt = cnt_lo;   // survivors: first half, all
#pragma unroll 8
  for (uint32_t j = K / 2; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[own_addr(t, j, K)]);
  {
    uint32_t ws = wave_sum(cnt);
    if (lane == 0) sh.wsum[wv] = (int32_t)ws;
  }
  switch (grp.n) {                                      // member slots: a power of two
  case 1: hard_chunks<T, K, 1>(bd, sh, lds, c, nst, rep, cstep, lo, hi, lo_w, cnt, cnt_lo); break;
#if CPX_GROUP >= 2
  case 2: hard_chunks<T, K, 2>(bd, sh, lds, c, nst, rep, cstep, lo, hi, lo_w, cnt, cnt_lo); break;
#endif
#if CPX_GROUP >= 4
  case 3: case 4: hard_chunks<T, K, 4>(bd, sh, lds, c, nst, rep, cstep, lo, hi, lo_w, cnt, cnt_lo); break;
#endif
#if CPX_GROUP >= 8
  default: hard_chunks<T, K, 8>(bd, sh, lds, c, nst, rep, cstep, lo, hi, lo_w, cnt, cnt_lo); break;
#endif
  }
  __syncthreads();
  if (t < G && sh.gm[t].on) {
    long long tt = 0;
    for (uint32_t w = 0; w < NW; w++) tt += sh.red_s[t][w];
    st_out(&bd.hard_acc[sh.gm[t].acc_slot], (int64_t)tt);
    if (rep == 0) st_out(&bd.hard_bact[sh.gm[t].row], (uint32_t)sh.gm[t].nst);
  }
  if (rep == 0 && p2any) p2_epilogue<T, K>(bd, grp, seg, lo, c + nst, lds, sh);
}

// Easy leaves (replaces S2_a / V_a / V_b, src/Deleglise-Rivat/S2.cu:72-99,
// src/V.cu:22-124) of one descriptor, as a tile of the persistent kernel:
//   V = sum_{x^(1/4) < p <= cbrt x} sum_{p < q <= min(y, x/p^2)} [pi(x/pq) - pi(p) + 2]
// over the descriptor's primes p = primes[j], j in [begin, end) (all these
// leaves have x/pq < p^2).  The block sets up T primes at a time in parallel
// (x/p, the q range, the (2 - pi(p)) terms), then strides its threads over
// each prime's q's with U = 4 independent prime -> quotient -> pi-table
// chains per thread.  Uses the first 3T words of the segment LDS.
template <int T, int K>
__device__ __forceinline__ void easy_tile(const BatchDev &bd, uint32_t di, uint32_t *lds, SegShared<T, K> &sh)
{
  constexpr uint32_t U = 4;
  constexpr uint32_t SBT = T < 256 ? T : 256, SUB = T / SBT, BIG = 8 * U * T;
  uint64_t *sxp = reinterpret_cast<uint64_t *>(lds);    // x / p of the chunk's primes
  uint32_t *sk1 = lds + 2 * T;                          // pi(min(y, x / p^2)): q indices end here
  uint32_t *sk0 = lds + 3 * T;                          // max(j + 1, pi(x / p^3)): and start here
  const uint32_t t = threadIdx.x;
  const BlockDesc d = bd.easy_desc[di];
  const QueryDev &q = bd.queries[d.query];
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  i128 sum = 0;
  for (uint64_t j0 = d.begin; j0 < d.end; j0 += T) {
    const uint32_t nj = (uint32_t)min((uint64_t)T, d.end - j0);
    if (t < nj) {
      const uint64_t j = j0 + t;
      const uint64_t p = bd.primes[j];
      const uint64_t xp = cpx_div128_to64(x, p);
      uint64_t qmax = q.y;
      if ((xp >> 50) < p) {
        const uint64_t v = cpx_div_smallq(xp, p);
        if (v < qmax) qmax = v;
      }
      const uint64_t k1 = pi_lookup(bd.pitab, qmax);
      // q > x/p^3 (x/(pq) < p^2): no bound above x^(1/4), where x/p^3 < p
      const uint64_t t3 = (xp / p) / p;
      const uint64_t k0 = t3 > p ? max(j + 1, pi_lookup(bd.pitab, t3 < q.y ? t3 : q.y)) : j + 1;
      sxp[t] = xp;
      sk1[t] = (uint32_t)k1;
      sk0[t] = (uint32_t)k0;
      if (k1 > k0) sum += (i128)(k1 - k0) * (1 - (i128)j);     // (2 - pi(p)) per q, pi(p) = j + 1
    }
    __syncthreads();
    // primes with at least BIG q's take the whole block one at a time (U * T
    // slots per step); the others go to sub-blocks of SBT threads, every SUB-th
    // prime to one sub-block (U * SBT slots per step: fewer idle lanes in a
    // prime's last step)
    for (uint32_t pass = 0; pass < 2; pass++) {
      const uint32_t nthr = pass ? SBT : T, stride = U * nthr, tl = pass ? t % SBT : t;
      for (uint32_t jj = pass ? t / SBT : 0; jj < nj; jj += pass ? SUB : 1) {
        const uint64_t k1 = sk1[jj], k0 = sk0[jj], xp = sxp[jj];
        if (k1 <= k0 || ((k1 - k0 >= BIG) != (pass == 0))) continue;
        // a chain past k1 repeats index k and is masked; x/(p q) < x/p^2 <
        // 2^50, so the reciprocal divide with the tabulated 1/q is exact
        const double xpf = (double)xp;
        uint64_t s = 0;
        for (uint64_t k = k0 + tl; k < k1; k += stride) {
          uint32_t qv[U];
          double rv[U];
          bool ok[U];
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            ok[u] = k + u * nthr < k1;
            const uint64_t kk = ok[u] ? k + u * nthr : k;
            qv[u] = bd.primes[kk];
            rv[u] = bd.prinv[kk];
          }
          PiEntry pe[U];
          uint32_t ix[U];
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            const uint64_t v = cpx_div_leaf_r(xp, xpf, qv[u], rv[u]);
            ix[u] = (uint32_t)(((v & 63) + 1) >> 1) | (v >= 2 ? 0x80u : 0u);
            pe[u] = bd.pitab[v >> 6];
          }
#pragma unroll
          for (uint32_t u = 0; u < U; u++) {
            const uint32_t c = (ix[u] >> 7) + pe[u].count + cpx_popc32(pe[u].bits & cpx_lowmask32(ix[u] & 63));
            s += ok[u] ? c : 0u;
          }
        }
        sum += (i128)s;
      }
    }
    __syncthreads();
  }
  // block sum: wave sums, then wave 0 adds them
  sum = wave_sum_i128(sum);
  i128 *red = reinterpret_cast<i128 *>(lds);
  if (lane_id() == 0) red[wave_id()] = sum;
  __syncthreads();
  if (t == 0) {
    i128 tot = 0;
    for (uint32_t w = 0; w < T / CPX_WAVE; w++) tot += red[w];
    // a descriptor of hard leaves (its part in the hard group) adds mu * phi
    const bool hard = d.part >= q.n_ord_parts + q.n_triv_parts + q.n_easy_parts;
    st_out(&bd.parts[q.part_off + d.part], hard ? -tot : tot);
  }
  (void)sh;
}

// P2 by lookups (replaces the windowed upper_bound count of src/P2.cu:76-133)
// for one descriptor, as a tile of the persistent kernel: the sum of
// pi(x/p) over P2 primes p = primes[k], k in [begin, end) (y < p <= sqrt x,
// x < 2^64), x/p by the reciprocal divide with the tabulated 1/p (x/p <=
// x/(y+1) < 2^50), pi from the resident table, which reaches x/(y+1); four
// independent chains per thread.  k_combine subtracts sum (pi(p) - 1) in
// closed form.  Uses the first T/16 words of the segment LDS.
template <int T, int K>
__device__ __forceinline__ void p2tab_tile(const BatchDev &bd, uint32_t di, uint32_t *lds)
{
  constexpr uint32_t U = 4;
  const uint32_t t = threadIdx.x;
  const BlockDesc d = bd.p2_desc[di];
  const QueryDev &q = bd.queries[d.query];
  const uint64_t x = q.x_lo;
  const double xf = (double)x;
  uint64_t s = 0;
  for (uint64_t k = d.begin + t; k < d.end; k += U * T) {
    uint64_t v[U];
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint64_t kk = k + u * T < d.end ? k + u * T : k;
      v[u] = cpx_div_leaf_r(x, xf, bd.primes[kk], bd.prinv[kk]);
    }
#pragma unroll
    for (uint32_t u = 0; u < U; u++) {
      const uint64_t c = pi_lookup(bd.pitab, v[u]);
      s += k + u * T < d.end ? c : 0u;
    }
  }
  const uint64_t ws = wave_sum(s);
  uint64_t *red = reinterpret_cast<uint64_t *>(lds);
  if (lane_id() == 0) red[wave_id()] = ws;
  __syncthreads();
  if (t == 0) {
    i128 tot = 0;
    for (uint32_t w = 0; w < T / CPX_WAVE; w++) tot += (i128)red[w];
    st_out(&bd.parts[q.part_off + d.part], (i128)tot);
  }
}

