// segsieve.hip -- the persistent segment kernel: hard-leaf phi sieve and the
// P2 segmented prime count, one LDS-resident bitmap segment per tile.
//
// Hard leaves (replaces S2glob::S2ctl + S2hardHost, src/sieve/S2_hard_device.cu:264-355,
// src/sieve/S2_hard_host.cu:115-191; DR split p <= x^(1/4), src/Deleglise-Rivat/S3.cu:8):
//   A segment holds 32*K*T odd numbers lo+1, lo+3, ... as bits (1 = unsieved).
//   It starts as the wheel of the first c primes.  For each stage b
//   (p = p_b, c < b <= B, while p^2 <= x/lo):
//     1. wavefront inclusive scan of per-thread survivor counts + one LDS
//        exchange of wave totals  -> every thread's exclusive prefix, and the
//        segment total T_b (written out for the carry);
//     2. leaves of p: thread t owns bits [t*32K, (t+1)*32K) and evaluates the
//        leaves whose n = x/(m p) falls in its bits, m descending so n
//        ascends and one running popcount over its own words gives
//        phi_local(n) = prefix + survivors <= n.  Dense stages (p^2 <= y)
//        scan odd m with mu(m) != 0 and lpf(m) > p from the mu*lpf table;
//        sparse stages (p^2 > y) walk only primes m (range from the pi table);
//     3. cross off the odd multiples of p inside the thread's own bits,
//        updating its survivor count (no LDS atomics: words are private).
//   One barrier per stage.  Per segment it writes T_b, sum_b(mu) and the local
//   sum  sum mu * phi_local; hard_finalize adds the carry
//   sum_b prefix_b(seg) * sum_b(mu)  (the decomposition the reference does with
//   scanVectorized/addMultiply, src/sieve/S2_hard_device.cu:357-375).
//
// P2 (replaces the windowed upper_bound count of src/P2.cu:76-133):
//   segments of [sqrt(x), x/(y+1)] are fully sieved (wheel + primes with
//   p^2 < segment top), then each thread counts primes <= x/p for the P2
//   primes p (from the pi-table bits) whose x/p lands in its own bits.  Per
//   segment: prime count, number of p, and sum of local counts; the carry is
//   applied in the combine kernel.
#include "cpx_device.hpp"

template <int T, int K>
struct SegShared {
  uint32_t wtot[2][T / CPX_WAVE];
  int32_t wsum[2][T / CPX_WAVE];
  unsigned long long red_u[T / CPX_WAVE];
  long long red_s[T / CPX_WAVE];
  uint32_t tile;
};

template <int T, int K>
__device__ __forceinline__ void hard_segment(const BatchDev &bd, const QueryDev &q, uint64_t seg,
                                             uint32_t *lds, SegShared<T, K> &sh)
{
  constexpr uint32_t NW = T / CPX_WAVE;
  constexpr uint32_t BPT = 32 * K;
  const uint32_t t = threadIdx.x, lane = lane_id(), wv = wave_id();
  const uint64_t SB = (uint64_t)BPT * T;
  const uint64_t lo = seg * 2 * SB;
  const uint64_t lo_t = lo + 2ull * BPT * t;
  const uint64_t hi_t = lo_t + 2ull * BPT;
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const uint64_t y = q.y;
  const WheelInfo wi = bd.winfo[q.c];

  seg_fill_wheel(lds, K, T, lo, bd.wheel + wi.off, wi.period);
  __syncthreads();
  uint32_t cnt = 0;
#pragma unroll 8
  for (uint32_t j = 0; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[j * T + t]);

  // stage b is active while p_b^2 <= x / lo (all leaves of p_b have n < x/p_b^2)
  uint64_t xlo = ~0ull;
  if (lo) {
    u128 qq = x / lo;
    xlo = (qq >> 64) ? ~0ull : (uint64_t)qq;
  }
  const uint64_t row = q.hard_row0 + seg;
  uint32_t *outT = bd.hard_T + q.tm_off + seg * q.nstage;
  int32_t *outM = bd.hard_M + q.tm_off + seg * q.nstage;
  const StageDev *st = bd.stages + q.stage_off;
  const int32_t *__restrict__ mlpf = bd.mlpf;

  int64_t acc = 0;
  uint32_t s = 0;
  StageDev cur = st[0];
  for (; s < q.nstage; s++) {
    const uint32_t p = cur.p;
    const uint64_t xp = cur.xp;
    if ((uint64_t)p * p > xlo) break;
    StageDev nxt = (s + 1 < q.nstage) ? st[s + 1] : cur;
    const uint32_t buf = s & 1u;

    // 1. scan of survivors (state: crossed by p_1 .. p_{b-1})
    uint32_t incl = wave_incl_scan(cnt);
    if (lane == CPX_WAVE - 1) sh.wtot[buf][wv] = incl;
    __syncthreads();
    uint32_t pre = incl - cnt, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) {
      uint32_t v = sh.wtot[buf][w];
      tot += v;
      pre += (w < wv) ? v : 0u;
    }
    if (t == 0) {
      outT[s] = tot;
      if (s > 0) {
        int32_t m = 0;
#pragma unroll
        for (uint32_t w = 0; w < NW; w++) m += sh.wsum[buf ^ 1u][w];
        outM[s - 1] = m;
      }
    }

    // 2. leaves of p with n in this thread's bits
    int32_t smu = 0;
    {
      const bool sparse = (uint64_t)p * p > y;
      uint64_t ylo = cpx_div_smallq(y, p);
      uint64_t mhi = y;
      if (lo_t != 0 && (xp >> 50) < lo_t) {
        uint64_t v = cpx_div_smallq(xp, lo_t);
        if (v < mhi) mhi = v;
      }
      uint64_t mlo = ~0ull;
      if ((xp >> 50) < hi_t) mlo = cpx_div_smallq(xp, hi_t);
      if (mlo < ylo) mlo = ylo;
      if (sparse && mlo < p) mlo = p;
      if (mhi > mlo) {
        uint32_t wcur = 0, run = 0;
        if (!sparse) {
          uint64_t m = mhi - ((mhi & 1) ^ 1);          // largest odd <= mhi
          for (; m > mlo; m -= 2) {
            int32_t v = mlpf[m >> 1];
            uint32_t av = (uint32_t)(v < 0 ? -v : v);
            if (av <= p) continue;                     // mu(m) = 0 or lpf(m) <= p
            uint64_t n = cpx_div_smallq(xp, m);       // lo_t <= n < hi_t
            uint32_t nb = (uint32_t)((n - lo_t + 1) >> 1);
            uint32_t wix = nb >> 5;
            while (wcur < wix) { run += cpx_popc32(lds[wcur * T + t]); wcur++; }
            uint32_t c = run + ((nb & 31u) ? cpx_popc32(lds[wix * T + t] & cpx_lowmask32(nb & 31u)) : 0u);
            int64_t phi = (int64_t)pre + c;
            if (v > 0) { acc += phi; smu++; } else { acc -= phi; smu--; }
          }
        } else {
          uint64_t ih = pi_lookup(bd.pitab, mhi), il = pi_lookup(bd.pitab, mlo);
          for (uint64_t k = ih; k > il; k--) {
            uint64_t m = bd.primes[k - 1];             // prime, mu = -1
            uint64_t n = cpx_div_smallq(xp, m);
            uint32_t nb = (uint32_t)((n - lo_t + 1) >> 1);
            uint32_t wix = nb >> 5;
            while (wcur < wix) { run += cpx_popc32(lds[wcur * T + t]); wcur++; }
            uint32_t c = run + ((nb & 31u) ? cpx_popc32(lds[wix * T + t] & cpx_lowmask32(nb & 31u)) : 0u);
            acc -= (int64_t)pre + c;
            smu--;
          }
        }
      }
    }
    smu = wave_sum(smu);
    if (lane == 0) sh.wsum[buf][wv] = smu;

    // 3. cross off odd multiples of p in this thread's own bits
    {
      uint32_t oseg = (uint32_t)first_odd_multiple_bit(lo, p, 0);
      uint32_t tb = (t * BPT) % p;
      uint32_t j = oseg >= tb ? oseg - tb : oseg + p - tb;
      for (; j < BPT; j += p) {
        uint32_t addr = (j >> 5) * T + t;
        uint32_t bit = 1u << (j & 31u);
        uint32_t w = lds[addr];
        if (w & bit) { lds[addr] = w & ~bit; cnt--; }
      }
    }
    cur = nxt;
  }
  __syncthreads();
  if (t == 0 && s > 0) {
    int32_t m = 0;
#pragma unroll
    for (uint32_t w = 0; w < NW; w++) m += sh.wsum[(s - 1) & 1u][w];
    outM[s - 1] = m;
  }
  long long a = wave_sum((long long)acc);
  if (lane == 0) sh.red_s[wv] = a;
  __syncthreads();
  if (t == 0) {
    long long tot = 0;
    for (uint32_t w = 0; w < NW; w++) tot += sh.red_s[w];
    bd.hard_acc[row] = tot;
    bd.hard_bact[row] = s;
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
  const uint64_t lo_t = s0 + 2ull * BPT * t;
  const uint64_t hi_t = lo_t + 2ull * BPT;
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const WheelInfo wi = bd.winfo[q.c];

  seg_fill_wheel(lds, K, T, s0, bd.wheel + wi.off, wi.period);
  __syncthreads();
  seg_cross_primes(lds, K, T, s0, SB, bd.primes, q.c, (uint32_t)bd.nprimes);
  __syncthreads();
  uint32_t cnt = 0;
#pragma unroll 8
  for (uint32_t j = 0; j < (uint32_t)K; j++) cnt += cpx_popc32(lds[j * T + t]);
  uint32_t incl = wave_incl_scan(cnt);
  if (lane == CPX_WAVE - 1) sh.wtot[0][wv] = incl;
  __syncthreads();
  uint32_t pre = incl - cnt, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < NW; w++) {
    uint32_t v = sh.wtot[0][w];
    tot += v;
    pre += (w < wv) ? v : 0u;
  }

  // P2 primes p in (max(y, x/hi_t), min(sqrt x, x/lo_t)], descending
  unsigned long long loc = 0;
  uint32_t np = 0;
  uint64_t pmax = cpx_div128_to64(x, lo_t);
  if (pmax > q.sqrtx) pmax = q.sqrtx;
  uint64_t pmin = cpx_div128_to64(x, hi_t);
  if (pmin < q.y) pmin = q.y;
  if (pmax > pmin) {
    uint32_t wcur = 0, run = 0;
    for (uint64_t e = pmax >> 6;; e--) {
      PiEntry pe = bd.pitab[e];
      uint32_t b = pe.bits;
      // keep odd integers 64e + 2i + 1 in (pmin, pmax]
      uint64_t base = 64 * e;
      if (pmax < base + 63) b &= cpx_lowmask32((uint32_t)((pmax - base + 1) >> 1));
      if (pmin + 1 > base + 1) {
        uint64_t kk = (pmin - base + 1) >> 1;          // bits for integers <= pmin
        if (kk >= 32) b = 0; else b &= ~cpx_lowmask32((uint32_t)kk);
      }
      while (b) {
        uint32_t i = 31u - (uint32_t)__builtin_clz(b);
        b &= ~(1u << i);
        uint64_t p = base + 2 * i + 1;
        uint64_t v = cpx_div128_to64(x, p);            // lo_t <= v < hi_t
        uint32_t nb = (uint32_t)((v - lo_t + 1) >> 1);
        uint32_t wix = nb >> 5;
        while (wcur < wix) { run += cpx_popc32(lds[wcur * T + t]); wcur++; }
        uint32_t c = run + ((nb & 31u) ? cpx_popc32(lds[wix * T + t] & cpx_lowmask32(nb & 31u)) : 0u);
        loc += (unsigned long long)pre + c;
        np++;
      }
      if (base <= pmin + 1) break;
    }
  }
  loc = wave_sum(loc);
  uint32_t npw = wave_sum(np);
  if (lane == 0) { sh.red_u[wv] = loc; sh.wsum[1][wv] = (int32_t)npw; }
  __syncthreads();
  if (t == 0) {
    unsigned long long L = 0;
    uint32_t N = 0;
    for (uint32_t w = 0; w < NW; w++) { L += sh.red_u[w]; N += (uint32_t)sh.wsum[1][w]; }
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
    if (tile.kind == TILE_HARD) hard_segment<T, K>(bd, q, tile.seg, lds, sh);
    else p2_segment<T, K>(bd, q, tile.seg, lds, sh);
    __syncthreads();
  }
}

size_t seg_lds_bytes(SegShape shape) { return (size_t)shape.kwords * shape.threads * 4; }

#define CPX_SEG_SHAPES(X) \
  X(256, 32) X(256, 64) X(256, 128) X(512, 32) X(512, 64) X(1024, 32)

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
    (void)hipFuncSetAttribute((const void *)k_segments<TT, KK>,                                  \
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);               \
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
