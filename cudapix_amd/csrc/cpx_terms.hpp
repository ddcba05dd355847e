// cpx_terms.hpp -- block-level bodies of the cheap terms, the hard-leaf
// carry and the per-query combine, shared by their grid-stride kernels
// (terms.hip) and by the streaming kernel, which runs them as tiles.  Every
// function is written for a block of NT threads (a multiple of 64) and takes
// its LDS scratch from the caller.
#pragma once
#include "cpx_device.hpp"

// ----------------------------------------------------------- ordinary
// S0 = sum over odd n <= y with lpf(n) > p_c of mu(n) * phi(x/n, c),
// phi(u, c) = floor(u / W_c) * totient(W_c) + phi(u mod W_c, c)
// (src/ordinary.cu:101-115 and :118-135 for 128-bit x).  Items: odd index i,
// n = 2i + 1.  Returns this thread's share of descriptor d.
template <uint32_t NT>
__device__ __forceinline__ i128 ordinary_part(const BatchDev &bd, const BlockDesc &d, uint32_t t)
{
  const QueryDev &q = bd.queries[d.query];
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  const WheelInfo wi = bd.winfo[q.c];
  const PiEntry *wh = bd.wheel + wi.off;
  i128 sum = 0;
  for (uint64_t i = d.begin + t; i < d.end; i += NT) {
    int32_t v = bd.mlpf[i];
    uint32_t av = (uint32_t)(v < 0 ? -v : v);
    if (av <= wi.pc) continue;
    uint64_t n = 2 * i + 1;
    uint64_t rr;
    u128 m = cpx_div128_64(x, n, &rr);
    uint64_t r;
    u128 k = cpx_div128_64(m, wi.W, &r);
    uint32_t kk = (uint32_t)((r + 1) >> 1);
    PiEntry e = wh[kk >> 5];
    uint64_t ph = e.count + cpx_popc32(e.bits & cpx_lowmask32(kk & 31u));
    i128 phi = (i128)(k * wi.totient) + (i128)ph;
    sum += v > 0 ? phi : -phi;
  }
  return sum;
}

// ------------------------------------------------------------ trivial U
// U = sum over primes sqrt(z) < p <= cbrt(x) of [pi(y) - pi(x / p^2)]
// (src/trivial.cu:31-62).  Items: prime index i (p = primes[i]).
template <uint32_t NT>
__device__ __forceinline__ i128 trivial_part(const BatchDev &bd, const BlockDesc &d, uint32_t t)
{
  const QueryDev &q = bd.queries[d.query];
  const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
  i128 sum = 0;
  for (uint64_t i = d.begin + t; i < d.end; i += NT) {
    uint64_t p = bd.primes[i];
    uint64_t qq = cpx_div128_to64(x, p * p);    // < y
    sum += (i128)q.a - (i128)pi_lookup(bd.pitab, qq);
  }
  return sum;
}

// -------------------------------------------------------- hard carry
// Hard-leaf carry of stages [s0, s0 + ns) (ns <= CPX_FIN_STAGES) of one
// query: for each of them
//   carry_s = sum over segments r (in order, while s is active in r) of
//             msum_s(r) * (T_s(0) + ... + T_s(r-1))
// with msum_s(r) the replicas' sum(mu) of the leaves of p_s in segment r and
// T_s(r) its survivors before stage s.  Activity is a prefix of the segments
// (bact is non-increasing in r), so the block scans NT segments at a time
// and stops after the first chunk that ends inactive for s0.  A thread reads
// its segment's ns consecutive T and sum(mu) entries, which share cache lines
// (the row record is read once for all of them).
template <uint32_t NT>
struct FinShared {
  i128 red[NT / CPX_WAVE];
  unsigned long long wt[CPX_FIN_STAGES][NT / CPX_WAVE];
  uint32_t stop;
};

// Writes the descriptor's carry to bd.parts; every thread of the block calls it.
// In a run in passes it scans the pass's rows only, starts each stage's
// prefix from hard_base (the survivors of the earlier passes' segments),
// writes the advanced totals back, and adds to the partial sum after the
// first pass: the stages' carries over the passes then sum to the one-launch carry.
template <uint32_t NT>
__device__ __forceinline__ void hard_fin_desc(const BatchDev &bd, const BlockDesc &d, FinShared<NT> &fs)
{
  constexpr uint32_t FS = CPX_FIN_STAGES, NW = NT / CPX_WAVE;
  const uint32_t t = threadIdx.x;
  const QueryDev &q = bd.queries[d.query];
  const uint64_t s0 = d.begin;
  const uint32_t ns = (uint32_t)(d.end - d.begin);           // 1..FS, block-uniform
  i128 carry = 0;
  // survivors of each stage in the rows before this launch's (runs in passes)
  unsigned long long base[FS];
#pragma unroll
  for (uint32_t j = 0; j < FS; j++) base[j] = bd.hard_base && j < ns ? bd.hard_base[s0 + j] : 0ull;
  const uint64_t r_end = min(q.hard_nseg, bd.pass_r1);
  // the next chunk's row records are fetched while this chunk is summed
  uint32_t ba_n = 0;
  HardRowInfo ri_n{};
  auto fetch = [&](uint64_t r) {
    ba_n = 0;
    if (r < r_end) {
      ba_n = bd.hard_bact[q.hard_row0 + r];
      if (ba_n > s0) ri_n = bd.rowinfo[q.hard_row0 + r];
    }
  };
  fetch(bd.pass_r0 + t);
  for (uint64_t r0 = bd.pass_r0; r0 < r_end; r0 += NT) {
    const uint32_t ba = ba_n;
    const HardRowInfo ri = ri_n;
    fetch(r0 + NT + t);
    unsigned long long tv[FS], incl[FS];
    long long msum[FS];
#pragma unroll
    for (uint32_t j = 0; j < FS; j++) { tv[j] = 0; msum[j] = 0; }
    bool act = false;
    if (ba > s0) {
      act = true;
      const uint32_t *T = bd.hard_T + (q.tm_off + ri.toff - bd.pass_toff) + s0;
      const int32_t *m = bd.hard_M + (q.m_off + ri.moff - bd.pass_moff) + s0;
      const uint32_t nj = min(ns, ba - (uint32_t)s0);   // this row's stages of the block
#pragma unroll
      for (uint32_t j = 0; j < FS; j++)
        if (j < nj) tv[j] = T[j];
      // replicas outermost: each step's nj loads are independent and
      // adjacent, so they are in flight together
#pragma unroll 2
      for (uint32_t k = 0; k < ri.nrep; k++) {
        const int32_t *mk = m + (uint64_t)k * ri.act;
#pragma unroll
        for (uint32_t j = 0; j < FS; j++)
          if (j < nj) msum[j] += mk[j];
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < FS; j++) {
      incl[j] = 0;
      if (j < ns) {
        incl[j] = wave_incl_scan(tv[j]);
        if (lane_id() == CPX_WAVE - 1) fs.wt[j][wave_id()] = incl[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < FS; j++) {
      if (j < ns) {
        unsigned long long pre = incl[j] - tv[j], tot = 0;
        for (uint32_t w = 0; w < NW; w++) {
          tot += fs.wt[j][w];
          if (w < wave_id()) pre += fs.wt[j][w];
        }
        carry += (i128)(base[j] + pre) * msum[j];
        base[j] += tot;
      }
    }
    // (read before the next write: the next chunk syncs after its scans)
    if (t == NT - 1) fs.stop = act ? 0u : 1u;
    __syncthreads();
    if (fs.stop) break;
  }
  i128 tt = block_sum_i128(carry, fs.red);
  if (t == 0) {
    i128 *out = &bd.parts[q.part_off + d.part];
    st_out(out, bd.pass_add ? *out + tt : tt);
    if (bd.hard_base) {
#pragma unroll
      for (uint32_t j = 0; j < FS; j++)
        if (j < ns) bd.hard_base[s0 + j] = base[j];
    }
  }
}

// ------------------------------------------------------------- combine
// One query: P2 carry scan over P2 segments, hard local sums, then
//   pi = S0 + (S1 + U) + V + S3 + pi(y) - 1 - P2
// (src/Deleglise-Rivat/deleglise-rivat.cpp:77).  With table_pi the query's
// pi(sqrt x) and pi(p2_e) are read from the pi table here instead of from
// the record (the streaming path has no setup kernel).
template <uint32_t NT>
struct CombShared {
  i128 red[NT / CPX_WAVE];
  uint64_t wtot[NT / CPX_WAVE];
};

static __device__ __forceinline__ void put_i128(uint64_t out[2], i128 v)
{
  out[0] = (uint64_t)v;
  out[1] = (uint64_t)((u128)v >> 64);
}

// A result record, written word by word with system-scope coherence: the
// streaming kernel's results live in pinned host memory that the host reads
// while the kernel still runs.
static __device__ __forceinline__ void st_result(ResultDev *dst, const ResultDev &r)
{
  const uint64_t *src = reinterpret_cast<const uint64_t *>(&r);
  uint64_t *d = reinterpret_cast<uint64_t *>(dst);
  for (uint32_t k = 0; k < sizeof(ResultDev) / 8; k++)
    __hip_atomic_store(d + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <uint32_t NT>
__device__ __forceinline__ void combine_query(const BatchDev &bd, uint32_t qi, CombShared<NT> &cs,
                                              bool table_pi)
{
  const uint32_t t = threadIdx.x;
  const QueryDev &q = bd.queries[qi];
  const uint64_t pi_e = table_pi ? pi_lookup(bd.pitab, q.p2_e) : q.pi_e;
  const uint64_t pi_sqrtx = table_pi ? pi_lookup(bd.pitab, q.sqrtx) : q.pi_sqrtx;
  // P2: sum over segments of loc + np * (pi(p2_e) + primes in earlier segments)
  i128 p2 = 0;
  uint64_t base = pi_e;
  for (uint64_t r0 = 0; r0 < q.p2_nseg; r0 += NT) {
    uint64_t r = r0 + t;
    uint64_t c = 0, np = 0, loc = 0;
    if (r < q.p2_nseg) {
      c = bd.p2_cnt[q.p2_row0 + r];
      np = bd.p2_np[q.p2_row0 + r];
      loc = bd.p2_loc[q.p2_row0 + r];
    }
    uint64_t incl = wave_incl_scan(c);
    if (lane_id() == CPX_WAVE - 1) cs.wtot[wave_id()] = incl;
    __syncthreads();
    uint64_t pre = incl - c, tot = 0;
    for (uint32_t w = 0; w < NT / CPX_WAVE; w++) {
      tot += cs.wtot[w];
      if (w < wave_id()) pre += cs.wtot[w];
    }
    p2 += (i128)loc + (i128)np * (i128)(base + pre);
    base += tot;
    __syncthreads();
  }
  // hard: local sums
  i128 hacc = 0;
  for (uint64_t r = t; r < q.hard_nmrow; r += NT) hacc += bd.hard_acc[q.mrow_base + r];
  i128 p2s = block_sum_i128(p2, cs.red);
  i128 hs = block_sum_i128(hacc, cs.red);
  if (t == 0 && q.small) {
    ResultDev res;
    i128 pi = (i128)pi_lookup(bd.pitab, q.x_lo);
    put_i128(res.pi, pi);
    put_i128(res.ordinary, 0);
    put_i128(res.trivial, 0);
    put_i128(res.easy, 0);
    put_i128(res.hard, 0);
    put_i128(res.p2, 0);
    res.pi_y = q.a;
    res.status = 1;
    st_result(&bd.results[qi], res);
  }
  // the term partial sums, by term (ordinary | trivial U | easy | hard carry
  // | P2 lookups)
  i128 ps0 = 0, psu = 0, psv = 0, psh = 0, psp = 0;
  if (!q.small) {
    const i128 *pp = bd.parts + q.part_off;
    const uint32_t n0 = q.n_ord_parts, n1 = n0 + q.n_triv_parts, n2 = n1 + q.n_easy_parts,
                   n3 = n2 + q.n_hard_parts, n4 = n3 + q.n_p2_parts;
    for (uint32_t i = t; i < n4; i += NT) {
      const i128 v = pp[i];
      if (i < n0) ps0 += v;
      else if (i < n1) psu += v;
      else if (i < n2) psv += v;
      else if (i < n3) psh += v;
      else psp += v;
    }
  }
  psp = block_sum_i128(psp, cs.red);
  ps0 = block_sum_i128(ps0, cs.red);
  psu = block_sum_i128(psu, cs.red);
  psv = block_sum_i128(psv, cs.red);
  psh = block_sum_i128(psh, cs.red);
  if (t == 0 && !q.small) {
    const i128 s0 = ps0, u = psu, v = psv, hc = psh;
    // S1 = C(pi(y) - pi(cbrt x), 2)
    i128 kk = (i128)q.a - (i128)q.pi_cbrt;
    i128 s1 = kk > 0 ? kk * (kk - 1) / 2 : 0;
    i128 triv = s1 + u;
    i128 hard = -(hs + hc);
    // P2 = sum pi(x/p) - sum_{p} (pi(p) - 1), pi(p) = a+1 .. A
    i128 A = (i128)pi_sqrtx, a = (i128)q.a;
    i128 P2 = 0;
    if (A > a) P2 = p2s + psp - (A * (A - 1) / 2 - a * (a - 1) / 2);
    i128 pi = s0 + triv + v + hard + a - 1 - P2;
    ResultDev res;
    put_i128(res.pi, pi);
    put_i128(res.ordinary, s0);
    put_i128(res.trivial, triv);
    put_i128(res.easy, v);
    put_i128(res.hard, hard);
    put_i128(res.p2, P2);
    res.pi_y = q.a;
    res.status = 1;
    st_result(&bd.results[qi], res);
  }
  __syncthreads();
}
