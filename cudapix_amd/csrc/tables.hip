// tables.hip -- device-resident tables of the pi(x) engine, built by
// hand-written sieves (replacing CUDASieve's prime generation/counting and the
// reference's LPF / Moebius sieves, src/sieve/lpf_mu.cu:25-302, and the pi
// table builder, src/pitable.cu:10-25,172-218):
//   * compact pi table (PiEntry per 64 integers) by a segmented odd-only
//     sieve of Eratosthenes in LDS, then an exclusive scan of popcounts;
//   * prime list up to a limit, compacted straight out of the pi table;
//   * mu(m)*lpf(m) for odd m (int32), by a segmented LDS sieve in which all
//     threads work on one sieving prime at a time (no atomics);
//   * wheel patterns of the first c primes with prefix counts (phi(r, c)).
#include "cpx_device.hpp"

// ------------------------------------------------------------ small sieve
// One workgroup: pi-table entries for [0, 64 * nent), nent <= 8192.
__global__ __launch_bounds__(256) void k_small_sieve(PiEntry *out, uint32_t nent)
{
  __shared__ uint32_t bm[8192];
  __shared__ uint32_t wsum[4];
  const uint32_t t = threadIdx.x, T = blockDim.x;
  for (uint32_t w = t; w < nent; w += T) bm[w] = 0xffffffffu;
  __syncthreads();
  if (t == 0) bm[0] &= ~1u;                       // 1 is not prime
  __syncthreads();
  const uint32_t nbits = nent * 32;
  for (uint32_t i = 1;; i++) {                      // bit i <-> 2i + 1
    uint32_t p = 2 * i + 1;
    if ((uint64_t)p * p >= 64ull * nent) break;
    bool prime = (bm[i >> 5] >> (i & 31)) & 1u;
    if (prime) {
      for (uint32_t j = (p * p - 1) / 2 + t * p; j < nbits; j += T * p)
        atomicAnd(&bm[j >> 5], ~(1u << (j & 31)));
    }
    __syncthreads();
  }
  // exclusive prefix of popcounts (each thread a contiguous chunk)
  uint32_t per = (nent + T - 1) / T;
  uint32_t b0 = t * per, b1 = min(nent, b0 + per);
  uint32_t c = 0;
  for (uint32_t w = b0; w < b1; w++) c += cpx_popc32(bm[w]);
  uint32_t incl = wave_incl_scan(c);
  if (lane_id() == CPX_WAVE - 1) wsum[wave_id()] = incl;
  __syncthreads();
  uint32_t pre = incl - c;
  for (uint32_t w = 0; w < wave_id(); w++) pre += wsum[w];
  for (uint32_t w = b0; w < b1; w++) {
    out[w].bits = bm[w];
    out[w].count = pre;
    pre += cpx_popc32(bm[w]);
  }
}

hipError_t launch_small_sieve(PiEntry *out, uint64_t nent, hipStream_t s)
{
  if (nent > 8192) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_sieve, dim3(1), dim3(256), 0, s, out, (uint32_t)nent);
  return hipGetLastError();
}

// -------------------------------------------------------- pi table segments
// Segment k: integers [k * 2*SB, (k+1) * 2*SB), SB = 32*K*T odd bits.
// Writes bits only; counts are filled by the scan kernels below.
template <int T, int K>
__global__ __launch_bounds__(T) void k_pitab_seg(PiEntry *tab, uint64_t nent,
                                                 const uint32_t *__restrict__ sp, uint32_t nsp)
{
  extern __shared__ uint32_t lds[];
  const uint32_t t = threadIdx.x;
  const uint64_t SB = 32ull * K * T;
  const uint64_t seg = blockIdx.x;
  const uint64_t lo = seg * 2 * SB;
  for (uint32_t g = t; g < (uint32_t)(K * T); g += T) lds[seg_addr(g, K, T)] = 0xffffffffu;
  __syncthreads();
  seg_cross_primes(lds, K, T, lo, SB, sp, 1, nsp);  // sp[0] = 2 is skipped
  __syncthreads();
  if (seg == 0 && t == 0) lds[seg_addr(0, K, T)] &= ~1u;   // 1 is not prime
  __syncthreads();
  const uint64_t e0 = seg * (SB / 32);
  for (uint32_t g = t; g < (uint32_t)(K * T); g += T) {
    uint64_t e = e0 + g;
    if (e < nent) { tab[e].bits = lds[seg_addr(g, K, T)]; tab[e].count = 0; }
  }
}

// Exclusive scan of popcount(bits) into count, in three passes.
#define SCAN_CHUNK 4096
__global__ __launch_bounds__(256) void k_scan_local(PiEntry *tab, uint64_t nent, uint32_t *tot)
{
  __shared__ uint32_t wsum[4];
  const uint32_t t = threadIdx.x;
  const uint64_t c0 = (uint64_t)blockIdx.x * SCAN_CHUNK;
  const uint32_t per = SCAN_CHUNK / 256;
  uint64_t b0 = c0 + (uint64_t)t * per;
  uint32_t c = 0;
  for (uint32_t k = 0; k < per; k++) if (b0 + k < nent) c += cpx_popc32(tab[b0 + k].bits);
  uint32_t incl = wave_incl_scan(c);
  if (lane_id() == CPX_WAVE - 1) wsum[wave_id()] = incl;
  __syncthreads();
  uint32_t pre = incl - c;
  for (uint32_t w = 0; w < wave_id(); w++) pre += wsum[w];
  for (uint32_t k = 0; k < per; k++) {
    if (b0 + k < nent) {
      uint32_t b = tab[b0 + k].bits;
      tab[b0 + k].count = pre;
      pre += cpx_popc32(b);
    }
  }
  if (t == 255) tot[blockIdx.x] = pre;
}

__global__ __launch_bounds__(256) void k_scan_tot(const uint32_t *tot, uint64_t n, uint64_t *base)
{
  // one block: chunked serial scan by 256 threads
  __shared__ uint64_t part[256];
  const uint32_t t = threadIdx.x;
  uint64_t per = (n + 255) / 256;
  uint64_t b0 = t * per, b1 = b0 + per < n ? b0 + per : n;
  uint64_t s = 0;
  for (uint64_t k = b0; k < b1; k++) s += tot[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    uint64_t r = 0;
    for (int k = 0; k < 256; k++) { uint64_t v = part[k]; part[k] = r; r += v; }
  }
  __syncthreads();
  uint64_t r = part[t];
  for (uint64_t k = b0; k < b1; k++) { base[k] = r; r += tot[k]; }
}

__global__ __launch_bounds__(256) void k_scan_add(PiEntry *tab, uint64_t nent, const uint64_t *base)
{
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nent;
       e += (uint64_t)gridDim.x * blockDim.x)
    tab[e].count += (uint32_t)base[e / SCAN_CHUNK];
}

static hipError_t scan_entries(PiEntry *tab, uint64_t nent, uint32_t *tot, uint64_t *base,
                               hipStream_t s)
{
  uint64_t nb = (nent + SCAN_CHUNK - 1) / SCAN_CHUNK;
  hipLaunchKernelGGL(k_scan_local, dim3((uint32_t)nb), dim3(256), 0, s, tab, nent, tot);
  hipLaunchKernelGGL(k_scan_tot, dim3(1), dim3(256), 0, s, tot, nb, base);
  uint32_t g = (uint32_t)((nent + 255) / 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_scan_add, dim3(g), dim3(256), 0, s, tab, nent, base);
  return hipGetLastError();
}

// seg_tot / seg_base must hold ceil(nent / SCAN_CHUNK) entries.
hipError_t launch_pitab_segments(PiEntry *tab, uint64_t nent, const uint32_t *sprimes,
                                 uint32_t nsprimes, uint32_t *seg_tot, uint64_t *seg_base,
                                 hipStream_t s)
{
  constexpr int T = 256, K = 32;                   // 2^18 odd bits per segment
  const uint64_t ent_per_seg = (uint64_t)K * T;
  uint64_t nseg = (nent + ent_per_seg - 1) / ent_per_seg;
  size_t lds = (size_t)(K + 1) * T * 4;
  hipLaunchKernelGGL((k_pitab_seg<T, K>), dim3((uint32_t)nseg), dim3(T), lds, s, tab, nent,
                     sprimes, nsprimes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return scan_entries(tab, nent, seg_tot, seg_base, s);
}

// ------------------------------------------------------------ prime list
// primes[0] = 2, then every odd prime <= limit in order, at index count+1+rank.
__global__ __launch_bounds__(256) void k_extract(const PiEntry *__restrict__ tab, uint64_t limit,
                                                 uint32_t *primes, uint64_t cap)
{
  uint64_t nent = limit / 64 + 1;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nent;
       w += (uint64_t)gridDim.x * blockDim.x) {
    PiEntry e = tab[w];
    uint32_t b = e.bits;
    uint32_t k = e.count + 1;
    while (b) {
      uint32_t i = __builtin_ctz(b);
      uint64_t n = 64 * w + 2 * i + 1;
      if (n > limit) break;
      if (k < cap) primes[k] = (uint32_t)n;
      k++;
      b &= b - 1;
    }
    if (w == 0 && cap) primes[0] = 2;
  }
}

// 1/p rounded to nearest (IEEE division) for every listed prime.
__global__ __launch_bounds__(256) void k_recip(const uint32_t *__restrict__ primes, double *rinv, uint64_t n)
{
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    rinv[i] = 1.0 / (double)primes[i];
}

hipError_t launch_recip(const uint32_t *primes, double *rinv, uint64_t n, hipStream_t s)
{
  if (!n) return hipSuccess;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_recip, dim3((uint32_t)g), dim3(256), 0, s, primes, rinv, n);
  return hipGetLastError();
}

hipError_t launch_extract_primes(const PiEntry *tab, uint64_t limit, uint32_t *primes,
                                 uint64_t cap, hipStream_t s)
{
  uint64_t nent = limit / 64 + 1;
  uint32_t g = (uint32_t)((nent + 255) / 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_extract, dim3(g), dim3(256), 0, s, tab, limit, primes, cap);
  return hipGetLastError();
}

// ---------------------------------------------------------- mu * lpf sieve
// Segment of MS odd numbers m = m0 + 2i + 1.  For each odd sieving prime p
// (all threads together, one prime per round, so every entry has a single
// writer in a round): divide the cofactor r by p exactly (multiplication by
// the inverse of p mod 2^32), record lpf on first hit, flip the parity, and
// flag m squareful when p still divides the cofactor (divisibility test by
// the same inverse: x*inv mod 2^32 <= (2^32-1)/p, Granlund-Montgomery 1994).  At the end a cofactor r > 1 is the one
// prime factor above sqrt(limit).  (Same odd-only layout and lpf(1) = "max"
// convention as src/sieve/lpf_mu.cu:25-197.)
#define MLPF_SEG 4096
__global__ __launch_bounds__(256) void k_mlpf(int32_t *mlpf, uint32_t *lpf_out, int8_t *mu_out,
                                              uint64_t ibase, uint64_t nodd,
                                              const uint32_t *__restrict__ sp, uint32_t nsp)
{
  __shared__ uint32_t r[MLPF_SEG];
  __shared__ uint32_t info[MLPF_SEG];             // lpf (bits 0-29), parity (30), squareful (31)
  const uint32_t t = threadIdx.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * MLPF_SEG;          // first output index
  const uint64_t m0 = 2 * (ibase + i0);                         // m = m0 + 2i + 1
  for (uint32_t i = t; i < MLPF_SEG; i += blockDim.x) {
    r[i] = (uint32_t)(m0 + 2 * i + 1);
    info[i] = 0;
  }
  __syncthreads();
  const uint64_t mhi = m0 + 2 * MLPF_SEG;
  for (uint32_t k = 1; k < nsp; k++) {
    uint32_t p = sp[k];
    if ((uint64_t)p * p >= mhi) break;
    uint32_t inv = p;                                            // p^-1 mod 2^32 (Newton)
    inv *= 2u - p * inv; inv *= 2u - p * inv; inv *= 2u - p * inv; inv *= 2u - p * inv;
    const uint32_t lim = 0xffffffffu / p;                        // x*inv <= lim <=> p | x
    uint64_t o = first_odd_multiple_bit(m0, p, p);
    for (uint64_t j = o + t * (uint64_t)p; j < MLPF_SEG; j += (uint64_t)blockDim.x * p) {
      uint32_t q = r[j] * inv;                                   // exact r / p
      r[j] = q;
      uint32_t f = info[j];
      if ((f & 0x3fffffffu) == 0) f |= p;
      f ^= 1u << 30;
      if (q * inv <= lim) f |= 1u << 31;                         // p^2 divides m
      info[j] = f;
    }
    __syncthreads();
  }
  for (uint32_t i = t; i < MLPF_SEG; i += blockDim.x) {
    uint64_t idx = i0 + i;
    if (idx >= nodd) break;
    uint32_t f = info[i], rc = r[i];
    uint32_t lp = f & 0x3fffffffu;
    uint32_t par = (f >> 30) & 1u;
    bool sq = f >> 31;
    if (rc > 1) { par ^= 1u; if (!lp) lp = rc; }
    int32_t mu = sq ? 0 : (par ? -1 : 1);
    uint32_t lpf = lp ? lp : 0x7fffffffu;                       // m = 1
    if (mlpf) mlpf[idx] = mu == 0 ? 0 : (mu > 0 ? (int32_t)lpf : -(int32_t)lpf);
    if (lpf_out) lpf_out[idx] = lp ? lp : 0xffffffffu;
    if (mu_out) mu_out[idx] = (int8_t)mu;
  }
}

hipError_t launch_mlpf(int32_t *mlpf, uint64_t nodd, const uint32_t *primes, uint32_t nsp,
                       hipStream_t s)
{
  uint32_t g = (uint32_t)((nodd + MLPF_SEG - 1) / MLPF_SEG);
  hipLaunchKernelGGL(k_mlpf, dim3(g), dim3(256), 0, s, mlpf, (uint32_t *)nullptr,
                     (int8_t *)nullptr, (uint64_t)0, nodd, primes, nsp);
  return hipGetLastError();
}

hipError_t launch_mu_lpf_split(uint32_t *lpf, int8_t *mu, uint64_t nodd, const uint32_t *primes,
                               uint32_t nsp, hipStream_t s)
{
  uint32_t g = (uint32_t)((nodd + MLPF_SEG - 1) / MLPF_SEG);
  hipLaunchKernelGGL(k_mlpf, dim3(g), dim3(256), 0, s, (int32_t *)nullptr, lpf, mu, (uint64_t)0,
                     nodd, primes, nsp);
  return hipGetLastError();
}

// The same sieve for odd m of 64 bits (a window anywhere below 2^64 - 2^34):
// the cofactor in 64 bits, exact division by p^-1 mod 2^64 (Newton from p,
// five steps for 3 -> 96 correct bits), the divisibility test x*inv <=
// (2^64-1)/p, and the lpf of a prime m above 2^32 kept whole.  Sieving
// primes reach sqrt(top) < 2^32.  (The 64-bit instances of
// src/sieve/lpf_mu.cu:57-87,153-197.)
__global__ __launch_bounds__(256) void k_mlpf64(uint64_t *lpf_out, int8_t *mu_out, uint64_t ibase, uint64_t nodd,
                                                const uint32_t *__restrict__ sp, uint32_t nsp)
{
  __shared__ uint64_t r[MLPF_SEG];
  __shared__ uint32_t lp[MLPF_SEG];
  __shared__ uint8_t fl[MLPF_SEG];                 // parity (bit 0), squareful (bit 1)
  const uint32_t t = threadIdx.x;
  const uint64_t i0 = (uint64_t)blockIdx.x * MLPF_SEG;
  const uint64_t m0 = 2 * (ibase + i0);
  for (uint32_t i = t; i < MLPF_SEG; i += blockDim.x) {
    r[i] = m0 + 2 * i + 1;
    lp[i] = 0;
    fl[i] = 0;
  }
  __syncthreads();
  const uint64_t mhi = m0 + 2 * MLPF_SEG;
  for (uint32_t k = 1; k < nsp; k++) {
    const uint32_t p = sp[k];
    if ((uint64_t)p * p >= mhi) break;
    uint64_t inv = p;
    for (int it = 0; it < 5; it++) inv *= 2u - p * inv;
    const uint64_t lim = ~0ull / p;
    const uint64_t o = first_odd_multiple_bit(m0, p, p);
    for (uint64_t j = o + t * (uint64_t)p; j < MLPF_SEG; j += (uint64_t)blockDim.x * p) {
      const uint64_t q = r[j] * inv;
      r[j] = q;
      if (!lp[j]) lp[j] = p;
      uint32_t f = fl[j] ^ 1u;
      if (q * inv <= lim) f |= 2u;
      fl[j] = (uint8_t)f;
    }
    __syncthreads();
  }
  for (uint32_t i = t; i < MLPF_SEG; i += blockDim.x) {
    const uint64_t idx = i0 + i;
    if (idx >= nodd) break;
    uint32_t par = fl[i] & 1u;
    const bool sq = (fl[i] >> 1) & 1u;
    uint64_t lpf = lp[i];
    if (r[i] > 1) { par ^= 1u; if (!lpf) lpf = r[i]; }
    if (lpf_out) lpf_out[idx] = lpf ? lpf : ~0ull;                // m = 1
    if (mu_out) mu_out[idx] = (int8_t)(sq ? 0 : (par ? -1 : 1));
  }
}

hipError_t launch_mu_lpf64_range(uint64_t *lpf, int8_t *mu, uint64_t ibase, uint64_t nodd,
                                 const uint32_t *primes, uint32_t nsp, hipStream_t s)
{
  uint32_t g = (uint32_t)((nodd + MLPF_SEG - 1) / MLPF_SEG);
  hipLaunchKernelGGL(k_mlpf64, dim3(g), dim3(256), 0, s, lpf, mu, ibase, nodd, primes, nsp);
  return hipGetLastError();
}

// lpf / mu of the odd n = 2(ibase + i) + 1, i < nodd (gen_d_lpf / gen_d_mu with an
// even bottom = 2 * ibase).
hipError_t launch_mu_lpf_range(uint32_t *lpf, int8_t *mu, uint64_t ibase, uint64_t nodd,
                               const uint32_t *primes, uint32_t nsp, hipStream_t s)
{
  uint32_t g = (uint32_t)((nodd + MLPF_SEG - 1) / MLPF_SEG);
  hipLaunchKernelGGL(k_mlpf, dim3(g), dim3(256), 0, s, (int32_t *)nullptr, lpf, mu, ibase, nodd,
                     primes, nsp);
  return hipGetLastError();
}

// --------------------------------------------------------------- wheels
// Bits of the first `period` words of the odd-only bitmap with every
// multiple of 3..p_c cleared (c = number of wheel primes including 2).
__global__ __launch_bounds__(256) void k_wheel(PiEntry *wheel, uint32_t period, uint32_t c)
{
  const uint32_t sp[8] = {2, 3, 5, 7, 11, 13, 17, 19};
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < period; w += gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (uint32_t i = 0; i < 32; i++) {
      uint64_t n = 2 * (32ull * w + i) + 1;
      bool ok = true;
      for (uint32_t k = 1; k < c; k++) ok = ok && (n % sp[k] != 0);
      bits |= (uint32_t)ok << i;
    }
    wheel[w].bits = bits;
    wheel[w].count = 0;
  }
}

hipError_t launch_wheel(PiEntry *wheel, uint32_t period, uint32_t c, hipStream_t s)
{
  uint32_t g = (period + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_wheel, dim3(g), dim3(256), 0, s, wheel, period, c);
  return hipGetLastError();
}

hipError_t launch_scan_entries(PiEntry *tab, uint64_t nent, uint32_t *tot, uint64_t *base,
                               hipStream_t s)
{
  return scan_entries(tab, nent, tot, base, s);
}

// ------------------------------------------------- reference-layout pi table
// out[k] = pi(lo + 2k - 1) - base_pi for k in [0, count), lo even; this is the
// u32-per-odd-n layout of src/pitable.cu:172-218 (index (n + 1 - lo) / 2).
__global__ __launch_bounds__(256) void k_expand(const PiEntry *__restrict__ tab, uint32_t *out,
                                                uint64_t lo, uint64_t count, uint64_t base_pi)
{
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t n = lo + 2 * k;
    uint64_t v = n >= 1 ? pi_lookup(tab, n - 1) : 0;
    out[k] = (uint32_t)(v - base_pi);
  }
}

hipError_t launch_expand_pitab_u32(const PiEntry *tab, uint32_t *out, uint64_t lo, uint64_t count,
                                   uint64_t base_pi, hipStream_t s)
{
  uint32_t g = (uint32_t)((count + 255) / 256);
  if (g > 65536) g = 65536;
  if (g == 0) return hipSuccess;
  hipLaunchKernelGGL(k_expand, dim3(g), dim3(256), 0, s, tab, out, lo, count, base_pi);
  return hipGetLastError();
}

// ------------------------------------------------------------- counting
// phi(x, a) for small x (class Phi, src/phi.cu:42-55): number of odd k in
// [lo_idx, hi_idx) (k = 2i + 1) whose least prime factor exceeds pa.
__global__ __launch_bounds__(256) void k_phi_count(const uint32_t *__restrict__ lpf, uint64_t i0,
                                                   uint64_t i1, uint32_t pa,
                                                   unsigned long long *out)
{
  unsigned long long c = 0;
  for (uint64_t i = i0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < i1;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += lpf[i] > pa;
  c = wave_sum(c);
  if (lane_id() == 0) atomicAdd(out, c);
}

hipError_t launch_phi_count_lpf(const uint32_t *lpf, uint64_t i0, uint64_t i1, uint32_t pa,
                                unsigned long long *out, hipStream_t s)
{
  if (i1 <= i0) return hipSuccess;
  uint64_t n = i1 - i0;
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_phi_count, dim3(g), dim3(256), 0, s, lpf, i0, i1, pa, out);
  return hipGetLastError();
}

// count of primes in [lo, hi) from the resident pi table (hi - 1 within table)
__global__ void k_count_range(const PiEntry *__restrict__ tab, uint64_t lo, uint64_t hi,
                              unsigned long long *out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint64_t a = lo >= 1 ? pi_lookup(tab, lo - 1) : 0;
    uint64_t b = hi >= 1 ? pi_lookup(tab, hi - 1) : 0;
    *out = b - a;
  }
}

hipError_t launch_count_range(const PiEntry *tab, uint64_t lo, uint64_t hi,
                              unsigned long long *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_count_range, dim3(1), dim3(64), 0, s, tab, lo, hi, out);
  return hipGetLastError();
}

// Plain segmented-Eratosthenes count of the primes in [lo, hi) (the
// CudaSieve::countPrimes replacement; also the brute-force check of pi(x)).
// Segment k covers integers [base + k*2*SB, ...), base = lo & ~1.
template <int T, int K>
__global__ __launch_bounds__(T) void k_count_sieve(uint64_t lo, uint64_t hi,
                                                   const uint32_t *__restrict__ sp, uint32_t nsp,
                                                   unsigned long long *out)
{
  extern __shared__ uint32_t lds[];
  const uint32_t t = threadIdx.x;
  const uint64_t SB = 32ull * K * T;
  const uint64_t base = lo & ~1ull;
  const uint64_t s0 = base + (uint64_t)blockIdx.x * 2 * SB;
  for (uint32_t g = t; g < (uint32_t)(K * T); g += T) lds[seg_addr(g, K, T)] = 0xffffffffu;
  __syncthreads();
  seg_cross_primes(lds, K, T, s0, SB, sp, 1, nsp);
  __syncthreads();
  // count bits whose integer is an odd prime in [max(lo, 3), hi)
  unsigned long long c = 0;
  for (uint32_t j = 0; j < (uint32_t)K; j++) {
    uint32_t w = lds[own_addr(t, j, K)];
    uint64_t g = (uint64_t)t * K + j;
    uint64_t n0 = s0 + 64 * g + 1;                 // integer of bit 0 of this word
    if (n0 + 62 < lo || n0 >= hi) continue;
    if (n0 < lo || n0 + 62 >= hi || n0 < 3) {
      for (int i = 0; i < 32; i++) {
        uint64_t n = n0 + 2 * i;
        if (((w >> i) & 1u) && n >= lo && n < hi && n >= 3) c++;
      }
    } else {
      c += cpx_popc32(w);
    }
  }
  c = wave_sum(c);
  if (lane_id() == 0) atomicAdd(out, c);
}

hipError_t launch_count_primes_sieve(uint64_t lo, uint64_t hi, const uint32_t *sprimes,
                                     uint32_t nsp, unsigned long long *out, hipStream_t s)
{
  if (hi <= lo) return hipSuccess;
  constexpr int T = 256, K = 64;                   // 2^19 odd bits per segment
  const uint64_t SB = 32ull * K * T;
  uint64_t base = lo & ~1ull;
  uint64_t nseg = (hi - base + 2 * SB - 1) / (2 * SB);
  if (nseg > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_count_sieve<T, K>), dim3((uint32_t)nseg), dim3(T), (size_t)(K + 1) * T * 4, s,
                     lo, hi, sprimes, nsp, out);
  return hipGetLastError();
}

// ------------------------------------------- odd-only prime bit sieve
// Segmented Eratosthenes over words of 64 integers from `base` (a multiple
// of 64): bit i of word j stands for n = base + 64 j + 2 i + 1.  A block
// sieves one segment of RS_K * RS_T words in LDS; word j gets 1 for an odd
// prime n in [lo, hi) (n >= 3) and 0 otherwise, or the complement when
// `composite_ones` (the genDeviceBitSieve convention the reference's hard
// leaves read: a clear bit is a prime, src/sieve/S2_hard_device.cu:248-249).
// seg_cnt[b] = primes in segment b.
#define RS_T 256
#define RS_K 64                                    // 2^20 integers per segment

__global__ __launch_bounds__(RS_T) void k_bit_sieve(uint64_t base, uint64_t lo, uint64_t hi,
                                                    uint64_t nwords, const uint32_t *__restrict__ sp,
                                                    uint32_t nsp, uint32_t composite_ones,
                                                    uint32_t *words, uint32_t *seg_cnt)
{
  extern __shared__ uint32_t lds[];
  __shared__ uint32_t wsum[RS_T / CPX_WAVE];
  const uint32_t t = threadIdx.x;
  constexpr uint32_t NWS = RS_K * RS_T;            // words per segment
  const uint64_t s0 = base + (uint64_t)blockIdx.x * 64 * NWS;
  for (uint32_t g = t; g < NWS; g += RS_T) lds[seg_addr(g, RS_K, RS_T)] = 0xffffffffu;
  __syncthreads();
  seg_cross_primes(lds, RS_K, RS_T, s0, 32ull * NWS, sp, 1, nsp);
  __syncthreads();
  uint32_t c = 0;
  for (uint32_t g = t; g < NWS; g += RS_T) {       // coalesced word order
    const uint64_t gw = (uint64_t)blockIdx.x * NWS + g;
    if (gw >= nwords) break;
    uint32_t w = lds[seg_addr(g, RS_K, RS_T)];
    const uint64_t n0 = s0 + 64 * (uint64_t)g + 1;     // integer of bit 0
    if (n0 < lo || n0 + 62 >= hi || n0 < 3) {
      uint32_t m = 0;
      for (uint32_t i = 0; i < 32; i++) {
        const uint64_t n = n0 + 2 * i;
        m |= (n >= lo && n < hi && n >= 3) ? (1u << i) : 0u;
      }
      w &= m;
    }
    c += cpx_popc32(w);
    if (words) words[gw] = composite_ones ? ~w : w;
  }
  c = wave_sum(c);
  if (lane_id() == 0) wsum[wave_id()] = c;
  __syncthreads();
  if (t == 0) {
    uint32_t tot = 0;
    for (uint32_t w = 0; w < RS_T / CPX_WAVE; w++) tot += wsum[w];
    seg_cnt[blockIdx.x] = tot;
  }
}

// The primes of each segment's words (prime = 1), in order, from out index
// seg_off[b] (the exclusive scan of the segment counts): thread t of block b
// takes words [t*RS_K, (t+1)*RS_K) of the segment, after a block scan of
// their popcounts.
template <typename OutT>
__global__ __launch_bounds__(RS_T) void k_bit_extract(uint64_t base, const uint32_t *__restrict__ words,
                                                      uint64_t nwords, const uint64_t *__restrict__ seg_off,
                                                      OutT *out)
{
  __shared__ uint32_t wtot[RS_T / CPX_WAVE];
  const uint32_t t = threadIdx.x;
  constexpr uint32_t NWS = RS_K * RS_T;
  const uint64_t w0 = (uint64_t)blockIdx.x * NWS + (uint64_t)t * RS_K;
  uint32_t c = 0;
  for (uint32_t j = 0; j < RS_K; j++)
    if (w0 + j < nwords) c += cpx_popc32(words[w0 + j]);
  const uint32_t incl = wave_incl_scan(c);
  if (lane_id() == CPX_WAVE - 1) wtot[wave_id()] = incl;
  __syncthreads();
  uint64_t k = seg_off[blockIdx.x] + (incl - c);
  for (uint32_t w = 0; w < wave_id(); w++) k += wtot[w];
  for (uint32_t j = 0; j < RS_K; j++) {
    if (w0 + j >= nwords) break;
    uint32_t b = words[w0 + j];
    while (b) {
      const uint32_t i = __builtin_ctz(b);
      b &= b - 1;
      out[k++] = (OutT)(base + 64 * (w0 + j) + 2 * i + 1);
    }
  }
}

uint64_t range_sieve_segments(uint64_t base, uint64_t hi)
{
  const uint64_t span = 64ull * RS_K * RS_T;
  return (hi - base + span - 1) / span;
}

size_t range_sieve_lds() { return (size_t)(RS_K + 1) * RS_T * 4; }

hipError_t launch_bit_sieve(uint64_t base, uint64_t lo, uint64_t hi, uint64_t nwords, const uint32_t *sp,
                            uint32_t nsp, bool composite_ones, uint32_t *words, uint32_t *seg_cnt,
                            hipStream_t s)
{
  const uint64_t nseg = range_sieve_segments(base, hi);
  if (nseg == 0) return hipSuccess;
  if (nseg > 0x7fffffffull) return hipErrorInvalidValue;
  (void)hipFuncSetAttribute((const void *)k_bit_sieve, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)range_sieve_lds());
  hipLaunchKernelGGL(k_bit_sieve, dim3((uint32_t)nseg), dim3(RS_T), range_sieve_lds(), s, base, lo, hi,
                     nwords, sp, nsp, composite_ones ? 1u : 0u, words, seg_cnt);
  return hipGetLastError();
}

hipError_t launch_bit_extract(uint64_t base, uint64_t hi, const uint32_t *words, uint64_t nwords,
                              const uint64_t *seg_off, void *out, bool out32, hipStream_t s)
{
  const uint64_t nseg = range_sieve_segments(base, hi);
  if (nseg == 0) return hipSuccess;
  if (nseg > 0x7fffffffull) return hipErrorInvalidValue;
  if (out32)
    hipLaunchKernelGGL(k_bit_extract<uint32_t>, dim3((uint32_t)nseg), dim3(RS_T), 0, s, base, words, nwords,
                       seg_off, (uint32_t *)out);
  else
    hipLaunchKernelGGL(k_bit_extract<uint64_t>, dim3((uint32_t)nseg), dim3(RS_T), 0, s, base, words, nwords,
                       seg_off, (uint64_t *)out);
  return hipGetLastError();
}

// x / y[i] for a 128-bit x, in place (divXbyY, src/P2.cu:135-138).
__global__ __launch_bounds__(256) void k_div_x_by_y(uint64_t xlo, uint64_t xhi, uint64_t *d,
                                                    size_t len)
{
  u128 x = ((u128)xhi << 64) | xlo;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (size_t)gridDim.x * blockDim.x) {
    uint64_t r;
    d[i] = (uint64_t)cpx_div128_64(x, d[i], &r);
  }
}

hipError_t launch_div_x_by_y(uint64_t x_lo, uint64_t x_hi, uint64_t *d_y, size_t len,
                             hipStream_t s)
{
  if (!len) return hipSuccess;
  uint32_t g = (uint32_t)((len + 255) / 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_div_x_by_y, dim3(g), dim3(256), 0, s, x_lo, x_hi, d_y, len);
  return hipGetLastError();
}

// ------------------------------------------------------- phi prefix (Phi::generateRange)
// out[i] = #{ j <= i : lpf[j] > pa } -- for odd k = 2i + 1 this is phi(k, a).
__global__ __launch_bounds__(256) void k_ind_local(const uint32_t *__restrict__ lpf, uint64_t n,
                                                   uint32_t pa, uint32_t *out, uint32_t *tot)
{
  __shared__ uint32_t wsum[4];
  const uint32_t t = threadIdx.x;
  const uint64_t c0 = (uint64_t)blockIdx.x * SCAN_CHUNK;
  const uint32_t per = SCAN_CHUNK / 256;
  const uint64_t b0 = c0 + (uint64_t)t * per;
  uint32_t c = 0;
  for (uint32_t k = 0; k < per; k++) if (b0 + k < n) c += lpf[b0 + k] > pa;
  uint32_t incl = wave_incl_scan(c);
  if (lane_id() == CPX_WAVE - 1) wsum[wave_id()] = incl;
  __syncthreads();
  uint32_t pre = incl - c;
  for (uint32_t w = 0; w < wave_id(); w++) pre += wsum[w];
  for (uint32_t k = 0; k < per; k++) {
    if (b0 + k < n) { pre += lpf[b0 + k] > pa; out[b0 + k] = pre; }
  }
  if (t == 255) tot[blockIdx.x] = pre;
}

__global__ __launch_bounds__(256) void k_add_base_u32(uint32_t *out, uint64_t n, const uint64_t *base)
{
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    out[e] += (uint32_t)base[e / SCAN_CHUNK];
}

hipError_t launch_phi_prefix(const uint32_t *lpf, uint64_t n, uint32_t pa, uint32_t *out,
                             hipStream_t s)
{
  if (!n) return hipSuccess;
  uint64_t nb = (n + SCAN_CHUNK - 1) / SCAN_CHUNK;
  uint32_t *tot = nullptr;
  uint64_t *base = nullptr;
  hipError_t e = hipMallocAsync((void **)&tot, nb * 4, s);
  if (!e) e = hipMallocAsync((void **)&base, nb * 8, s);
  if (e) return e;
  hipLaunchKernelGGL(k_ind_local, dim3((uint32_t)nb), dim3(256), 0, s, lpf, n, pa, out, tot);
  hipLaunchKernelGGL(k_scan_tot, dim3(1), dim3(256), 0, s, tot, nb, base);
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_add_base_u32, dim3(g), dim3(256), 0, s, out, n, base);
  e = hipGetLastError();
  (void)hipFreeAsync(tot, s);
  (void)hipFreeAsync(base, s);
  return e;
}

// ------------------------------------------ primes of [lo, hi) as uint64
__global__ __launch_bounds__(256) void k_extract_range(const PiEntry *__restrict__ tab, uint64_t lo,
                                                       uint64_t hi, uint64_t *out, uint64_t cap)
{
  const uint64_t base = lo >= 1 ? pi_lookup(tab, lo - 1) : 0;     // primes < lo
  const uint64_t w0 = lo >> 6, w1 = (hi - 1) >> 6;
  for (uint64_t w = w0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w <= w1;
       w += (uint64_t)gridDim.x * blockDim.x) {
    PiEntry e = tab[w];
    uint32_t b = e.bits;
    uint64_t k = 1 + (uint64_t)e.count;                           // global index of next odd prime
    while (b) {
      uint32_t i = __builtin_ctz(b);
      uint64_t n = 64 * w + 2 * i + 1;
      b &= b - 1;
      if (n >= lo && n < hi && k - base < cap) out[k - base] = n;
      k++;
    }
    if (w == 0 && lo <= 2 && hi > 2 && cap) out[0] = 2;
  }
}

hipError_t launch_extract_range_u64(const PiEntry *tab, uint64_t lo, uint64_t hi, uint64_t *out,
                                    uint64_t cap, hipStream_t s)
{
  if (hi <= lo) return hipSuccess;
  uint64_t nw = ((hi - 1) >> 6) - (lo >> 6) + 1;
  uint32_t g = (uint32_t)((nw + 255) / 256);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_extract_range, dim3(g), dim3(256), 0, s, tab, lo, hi, out, cap);
  return hipGetLastError();
}
