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
      d.xp = cpx_div128_to64(x, p);
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
