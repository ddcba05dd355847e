// cpx_device.hpp -- device building blocks shared by the table builders and
// the segment kernel: 64-wide wavefront scans/reductions, and the segmented
// odd-only Eratosthenes cross-off used for pi tables, prime counts and P2.
#pragma once
#include "cpx_internal.hpp"

#define CPX_WAVE 64

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (CPX_WAVE - 1); }
__device__ __forceinline__ uint32_t wave_id() { return threadIdx.x / CPX_WAVE; }

// Inclusive prefix sum of 32-bit values across the 64 lanes of a wavefront in
// registers (DPP): Hillis-Steele within each row of 16 lanes (row_shr 1, 2,
// 4, 8, zero-filled at the row start), then row_bcast:15 adds row 0's total
// to row 1 and row 2's to row 3, and row_bcast:31 adds lane 31's running
// total to rows 2 and 3 -- the GCN cross-lane scan of AMD's GPUOpen note
// "AMD GCN Assembly: Cross-Lane Operations" (2016).  The disabled rows of the
// two broadcast steps keep their value.  Every lane must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
  return v;
}

// floor(a / m) for a < 2^64 and 2 <= m < 2^32 whose quotient is below 2^50,
// given af = (double)a.  The hardware reciprocal is only about 2^-23 accurate;
// two Newton steps (error squares each step) bring r within a few ulp of 1/m,
// so af * r is within 0.375 of a/m (relative error < 1.5 * 2^-52, quotient
// < 2^50).  Adding 2^52 in the same fma rounds it to an integer in the low
// mantissa bits, within 0.875 of a/m: floor(a/m) or one above or below, and
// one exact remainder test (a - q*m lies in [-m, 2m), exact in int64 whatever
// the wrap of q*m) fixes it.  (One Newton step is not enough: near 2^50 the
// estimate is then off by up to ~16.)
__device__ __forceinline__ uint64_t cpx_div_leaf(uint64_t a, double af, uint32_t m)
{
  const double md = (double)m;
  double r = __builtin_amdgcn_rcp(md);
  r = __builtin_fma(__builtin_fma(-md, r, 1.0), r, r);
  r = __builtin_fma(__builtin_fma(-md, r, 1.0), r, r);
  const double est = __builtin_fma(af, r, 0x1p52);
  uint64_t q = (uint64_t)__builtin_bit_cast(long long, est) & ((1ull << 52) - 1);
  const int64_t rem = (int64_t)(a - q * (uint64_t)m);
  if (rem < 0) q--;
  else if (rem >= (int64_t)m) q++;
  return q;
}

// cpx_div_leaf with r = 1/m rounded to nearest given (a table of prime
// reciprocals): af * r is then within a/m * 2^-52 (< 0.25) of a/m, and the
// same rounding and remainder test apply.
__device__ __forceinline__ uint64_t cpx_div_leaf_r(uint64_t a, double af, uint32_t m, double r)
{
  const double est = __builtin_fma(af, r, 0x1p52);
  uint64_t q = (uint64_t)__builtin_bit_cast(long long, est) & ((1ull << 52) - 1);
  const int64_t rem = (int64_t)(a - q * (uint64_t)m);
  if (rem < 0) q--;
  else if (rem >= (int64_t)m) q++;
  return q;
}

// Sum of 32-bit values over the wavefront, in every lane (DPP scan, then lane 63).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
  return (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(v), CPX_WAVE - 1);
}

// Inclusive prefix sum across the 64 lanes of a wavefront (any width).
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
  const uint32_t l = lane_id();
#pragma unroll
  for (int d = 1; d < CPX_WAVE; d <<= 1) {
    T u = __shfl_up(v, d, CPX_WAVE);
    if (l >= (uint32_t)d) v += u;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int d = CPX_WAVE / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, CPX_WAVE);
  return v;
}

__device__ __forceinline__ i128 wave_sum_i128(i128 v)
{
  uint64_t lo = (uint64_t)v, hi = (uint64_t)((u128)v >> 64);
#pragma unroll
  for (int d = CPX_WAVE / 2; d > 0; d >>= 1) {
    uint64_t olo = __shfl_xor(lo, d, CPX_WAVE), ohi = __shfl_xor(hi, d, CPX_WAVE);
    u128 s = (((u128)hi << 64) | lo) + (((u128)ohi << 64) | olo);
    lo = (uint64_t)s; hi = (uint64_t)(s >> 64);
  }
  return (i128)(((u128)hi << 64) | lo);
}

// Block-wide sum of an i128 through LDS scratch (>= blockDim/64 entries).
__device__ __forceinline__ i128 block_sum_i128(i128 v, i128 *scratch)
{
  v = wave_sum_i128(v);
  if (lane_id() == 0) scratch[wave_id()] = v;
  __syncthreads();
  i128 t = 0;
  if (threadIdx.x == 0) {
    for (uint32_t w = 0; w < blockDim.x / CPX_WAVE; w++) t += scratch[w];
  }
  __syncthreads();
  return t;   // valid in thread 0
}

// Output rows that another block reads in the same launch (the streaming
// kernel's carry and combine units read what segment tiles wrote): stored
// with device-scope coherence, i.e. written through to the point every XCD's
// L2 reads from, so that no agent-scope fence (a whole-L2 write-back and
// invalidate on this device) is needed between producer and consumer.  The
// producer still waits for its stores (stream_release) before it signals.
__device__ __forceinline__ void st_out(uint32_t *p, uint32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_out(int32_t *p, int32_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_out(uint64_t *p, uint64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_out(int64_t *p, int64_t v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_out(i128 *p, i128 v)
{
  uint64_t *q = reinterpret_cast<uint64_t *>(p);
  st_out(q, (uint64_t)v);
  st_out(q + 1, (uint64_t)((u128)v >> 64));
}

// Padded-linear LDS layout of a segment bitmap: segment word g lives at
// lds[g + g / K].  Thread t owns the contiguous words [t*K, (t+1)*K); its
// j-th word sits at lds[t*(K+1) + j], so lanes walking their own words in
// step hit distinct banks (K+1 is odd), while block-wide strided cross-off
// (lanes p bits apart) stays nearly linear.  Needs K*T + T words.
__device__ __forceinline__ uint32_t seg_addr(uint32_t g, uint32_t K, uint32_t T)
{
  (void)T;
  return g + g / K;
}

__device__ __forceinline__ uint32_t own_addr(uint32_t t, uint32_t j, uint32_t K)
{
  return t * (K + 1) + j;
}

// Bit offset, inside a segment whose bit i is the integer lo + 2i + 1 (lo even),
// of the first odd multiple of the odd prime p that is >= max(lo + 1, from).
__device__ __forceinline__ uint64_t first_odd_multiple_bit(uint64_t lo, uint32_t p, uint64_t from)
{
  uint64_t start = lo + 1 > from ? lo + 1 : from;    // first candidate integer
  uint64_t r = cpx_mod64_32(start, p);
  uint64_t m = start + (r ? p - r : 0);               // first multiple >= start
  if (!(m & 1)) m += p;                               // odd multiple
  return (m - lo - 1) >> 1;
}

// first_odd_multiple_bit for max(lo + 1, from) / p < 2^50, with the
// reciprocal divide in place of the 64-bit remainder.
__device__ __forceinline__ uint64_t first_odd_multiple_bit_r(uint64_t lo, uint32_t p, uint64_t from)
{
  const uint64_t start = lo + 1 > from ? lo + 1 : from;
  const uint64_t r = start - cpx_div_leaf(start, (double)start, p) * p;
  uint64_t m = start + (r ? p - r : 0);
  if (!(m & 1)) m += p;
  return (m - lo - 1) >> 1;
}

// Cross off, in an interleaved LDS bitmap of `nbits` odd numbers starting at
// integer lo (lo even), all odd multiples >= p^2 of the odd sieving primes
// sp[i0 .. i1) with p^2 < lo + 2*nbits.  Small primes are shared by the whole
// block (each thread takes every T-th multiple); large ones are one per
// thread.  Uses LDS atomics (several primes hit one word); caller syncs.
__device__ __forceinline__ void seg_cross_primes(uint32_t *lds, uint32_t K, uint32_t T, uint64_t lo,
                                                 uint64_t nbits, const uint32_t *__restrict__ sp,
                                                 uint32_t i0, uint32_t i1)
{
  const uint64_t hi = lo + 2 * nbits;
  const uint32_t t = threadIdx.x;
  // shared primes: p < nbits / T keep every thread busy
  const uint64_t coop_lim = nbits / T;
  uint32_t i = i0;
  for (; i < i1; i++) {
    uint32_t p = sp[i];
    if ((uint64_t)p >= coop_lim) break;
    if ((uint64_t)p * p >= hi) { i = i1; break; }
    uint64_t o = first_odd_multiple_bit(lo, p, (uint64_t)p * p);
    for (uint64_t j = o + (uint64_t)t * p; j < nbits; j += (uint64_t)T * p) {
      uint32_t g = (uint32_t)(j >> 5);
      atomicAnd(&lds[seg_addr(g, K, T)], ~(1u << (j & 31)));
    }
  }
  // one prime per thread
  for (uint32_t k = i + t; k < i1; k += T) {
    uint32_t p = sp[k];
    if ((uint64_t)p * p >= hi) break;
    uint64_t o = first_odd_multiple_bit(lo, p, (uint64_t)p * p);
    for (uint64_t j = o; j < nbits; j += p) {
      uint32_t g = (uint32_t)(j >> 5);
      atomicAnd(&lds[seg_addr(g, K, T)], ~(1u << (j & 31)));
    }
  }
}

// Fill an interleaved segment bitmap from a wheel pattern: segment word g is
// pattern word (w0 + g) mod P, w0 = lo / 64.  Coalesced global reads.
__device__ __forceinline__ void seg_fill_wheel(uint32_t *lds, uint32_t K, uint32_t T, uint64_t lo,
                                               const PiEntry *__restrict__ pat, uint32_t P)
{
  const uint32_t t = threadIdx.x;
  const uint32_t W = K * T;
  // the pattern index steps by T mod P instead of a division per word, so
  // the loads of the unrolled loop go out together
  const uint32_t step = T % P;
  uint32_t idx = (uint32_t)(((lo >> 6) + t) % P);
#pragma unroll 8
  for (uint32_t g = t; g < W; g += T) {
    lds[seg_addr(g, K, T)] = pat[idx].bits;
    idx += step;
    if (idx >= P) idx -= P;
  }
}
