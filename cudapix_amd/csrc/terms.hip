// terms.hip -- grid-stride kernels for the cheap terms of
//   pi(x) = S0 + (S1 + U) + V + S3 + pi(y) - 1 - P2
// (assembly as src/Deleglise-Rivat/deleglise-rivat.cpp:77), plus the hard-leaf
// carry and the per-query combine.  Every kernel walks device-resident block
// descriptors, so its launch shape is fixed and a captured graph can be
// replayed for any batch that fits the plan capacity.
#include "cpx_terms.hpp"

__global__ void k_zero_u32(uint32_t *p, size_t n)
{
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

hipError_t launch_zero_u32(uint32_t *p, size_t n, hipStream_t s)
{
  uint32_t g = (uint32_t)((n + 255) / 256);
  if (g == 0) g = 1;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_zero_u32, dim3(g), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// --------------------------------------------------------------- setup
// Block per query: pi(sqrt x), pi(p2_e), and the stage table
// {p_b, floor(x / p_b)} with the emulated 128/64 divide.
__global__ __launch_bounds__(256) void k_setup(BatchDev bd)
{
  const uint32_t nq = *bd.nqueries;
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    QueryDev &q = bd.queries[qi];
    const u128 x = ((u128)q.x_hi << 64) | q.x_lo;
    if (threadIdx.x == 0) {
      q.pi_sqrtx = pi_lookup(bd.pitab, q.sqrtx);
      q.pi_e = pi_lookup(bd.pitab, q.p2_e);
    }
    for (uint32_t s = threadIdx.x; s < q.nstage; s += blockDim.x) {
      uint32_t p = bd.primes[q.c + s];
      StageDev d;
      d.p = p;
      d.ylo = (uint32_t)(q.y / p);
      d.inv_p = 1.0 / (double)p;
      d.magic = ~0ull / p + 1;
      uint64_t rem;
      const u128 xq = cpx_div128_64(x, p, &rem);
      d.xp = (uint64_t)xq;
      d.xph = (uint32_t)(xq >> 64);
      d.pad = 0;
      d.xp3 = ~0ull;
      if (q.omega && (uint64_t)p * p > q.y) {
        uint64_t r3;
        const u128 t = cpx_div128_64(xq, (uint64_t)p * p, &r3);
        d.xp3 = t < (u128)q.y ? (uint64_t)t : q.y;
      }
      bd.stages[q.stage_off + s] = d;
    }
  }
}

hipError_t launch_setup(BatchDev b, uint32_t grid, uint32_t, hipStream_t s)
{
  hipLaunchKernelGGL(k_setup, dim3(grid), dim3(256), 0, s, b);
  return hipGetLastError();
}

// ----------------------------------------------------------- ordinary
// S0 (ordinary_part, cpx_terms.hpp), one descriptor per block iteration.
__global__ __launch_bounds__(256) void k_ordinary(BatchDev bd)
{
  __shared__ i128 red[4];
  const uint32_t nd = *bd.n_ord;
  for (uint32_t di = blockIdx.x; di < nd; di += gridDim.x) {
    const BlockDesc d = bd.ord_desc[di];
    i128 t = block_sum_i128(ordinary_part<256>(bd, d, threadIdx.x), red);
    if (threadIdx.x == 0) bd.parts[bd.queries[d.query].part_off + d.part] = t;
  }
}

hipError_t launch_ordinary(BatchDev b, uint32_t grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_ordinary, dim3(grid), dim3(256), 0, s, b);
  return hipGetLastError();
}

// ------------------------------------------------------------ trivial U
// U (trivial_part, cpx_terms.hpp), one descriptor per block iteration.
__global__ __launch_bounds__(256) void k_trivial(BatchDev bd)
{
  __shared__ i128 red[4];
  const uint32_t nd = *bd.n_triv;
  for (uint32_t di = blockIdx.x; di < nd; di += gridDim.x) {
    const BlockDesc d = bd.triv_desc[di];
    i128 t = block_sum_i128(trivial_part<256>(bd, d, threadIdx.x), red);
    if (threadIdx.x == 0) bd.parts[bd.queries[d.query].part_off + d.part] = t;
  }
}

hipError_t launch_trivial(BatchDev b, uint32_t grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_trivial, dim3(grid), dim3(256), 0, s, b);
  return hipGetLastError();
}

// -------------------------------------------------------- hard carry
// hard_fin_desc (cpx_terms.hpp) per carry descriptor: 128-thread blocks,
// whose time is a latency chain over the segment chunks, so many in flight.
#define FIN_T 128          // carry-block threads: segments per chunk
__global__ __launch_bounds__(FIN_T) void k_hard_fin(BatchDev bd)
{
  __shared__ FinShared<FIN_T> fs;
  const uint32_t nd = *bd.n_fin;
  for (uint32_t di = blockIdx.x; di < nd; di += gridDim.x) {
    const BlockDesc d = bd.fin_desc[di];
    hard_fin_desc<FIN_T>(bd, d, fs);
  }
}

hipError_t launch_hard_finalize(BatchDev b, uint32_t grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_hard_fin, dim3(grid), dim3(FIN_T), 0, s, b);
  return hipGetLastError();
}

// ------------------------------------------------------------- combine
// Block per query (combine_query, cpx_terms.hpp).
__global__ __launch_bounds__(256) void k_combine(BatchDev bd)
{
  __shared__ CombShared<256> cs;
  const uint32_t nq = *bd.nqueries;
  for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) combine_query<256>(bd, qi, cs, false);
}

hipError_t launch_combine(BatchDev b, uint32_t grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_combine, dim3(grid), dim3(256), 0, s, b);
  return hipGetLastError();
}

// ------------------------------------------------------------ sigma form
// The reference's second form of the easy leaves (src/Deleglise-Rivat/S2.cu:30-110
// with the sums of src/Deleglise-Rivat/sigma.cu:52-112, 159-214): block b takes
// the prime p = primes[j0 + b] of (x^1/4, x^1/3] and sums over the primes
// p < q <= sqrt(x/p) the weighted pi(x/pq) of A (weight 2 when x/pq < y),
// then thread 0 stores the four per-p parts {A_p, pi(x/(p y)) if p <= sqrt z,
// pi(min(y, x/p^2)) if p > sqrt z, -pi(sqrt(x/p))^2}.  Needs x/p < 2^64,
// the prime list to sqrt(x/p) and the pi table to sqrt(x).
static __device__ __forceinline__ uint64_t isqrt64_dev(uint64_t v)
{
  uint64_t s = (uint64_t)sqrt((double)v);
  while ((u128)s * s > v) s--;
  while ((u128)(s + 1) * (s + 1) <= v) s++;
  return s;
}

__global__ __launch_bounds__(256) void k_sigma(uint64_t x_lo, uint64_t x_hi, uint64_t y, uint64_t sz,
                                               const uint32_t *__restrict__ primes,
                                               const PiEntry *__restrict__ pitab, uint64_t j0,
                                               uint64_t np, i128 *__restrict__ part)
{
  __shared__ i128 scratch[256 / CPX_WAVE];
  const u128 x = ((u128)x_hi << 64) | x_lo;
  for (uint64_t b = blockIdx.x; b < np; b += gridDim.x) {
    const uint64_t j = j0 + b;
    const uint64_t p = primes[j];
    const uint64_t xp = cpx_div128_to64(x, p);
    const uint64_t s = isqrt64_dev(xp);
    const uint64_t kend = pi_lookup(pitab, s);     // primes[k] <= s for k < kend
    const double xpf = (double)xp;
    uint64_t acc = 0;
    for (uint64_t k = j + 1 + threadIdx.x; k < kend; k += blockDim.x) {
      const uint64_t n = cpx_div_leaf(xp, xpf, primes[k]);   // < sqrt x
      acc += (n < y ? 2ull : 1ull) * pi_lookup(pitab, n);
    }
    const i128 a = block_sum_i128((i128)acc, scratch);
    if (threadIdx.x == 0) {
      i128 *o = part + 4 * b;
      o[0] = a;
      o[1] = p <= sz ? (i128)pi_lookup(pitab, xp / y) : (i128)0;
      const uint64_t v = xp / p;
      o[2] = p > sz ? (i128)pi_lookup(pitab, v < y ? v : y) : (i128)0;
      const i128 ps = (i128)pi_lookup(pitab, s);
      o[3] = -ps * ps;
    }
  }
}

hipError_t launch_sigma(u128 x, uint64_t y, uint64_t sz, const uint32_t *primes, const PiEntry *pitab,
                        uint64_t j0, uint64_t np, i128 *part, hipStream_t s)
{
  if (np == 0) return hipSuccess;
  const uint32_t g = (uint32_t)std::min<uint64_t>(np, 1u << 20);
  hipLaunchKernelGGL(k_sigma, dim3(g), dim3(256), 0, s, (uint64_t)x, (uint64_t)(x >> 64), y, sz, primes,
                     pitab, j0, np, part);
  return hipGetLastError();
}
