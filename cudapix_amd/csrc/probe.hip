// probe.hip -- microbenchmarks for the roofline peaks of this device: LDS read
// bandwidth and 32-bit integer VALU rate, measured on all CUs with HIP events.
#include "cpx_internal.hpp"

// LDS: every thread streams 16-byte reads over a 32 KiB block array
// (consecutive lanes, consecutive 16-byte slots: no bank conflicts), eight
// reads in flight per iteration; the xor of the data keeps the reads alive.
__global__ __launch_bounds__(256) void k_probe_lds(uint32_t iters, uint32_t *sink)
{
  __shared__ uint4 buf[2048];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < 2048; i += 256) buf[i] = make_uint4(i, i * 3u, i * 5u, i * 7u);
  __syncthreads();
  uint4 acc = make_uint4(0, 0, 0, 0);
  uint32_t base = t;
  for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
    for (uint32_t u = 0; u < 8; u++) {
      const uint4 v = buf[(base + u * 256) & 2047];
      acc.x ^= v.x; acc.y ^= v.y; acc.z ^= v.z; acc.w ^= v.w;
    }
    base += 7;
  }
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u) sink[blockIdx.x] = 1;
}

// VALU: eight independent add/xor chains per lane (16 int32 operations per
// lane per iteration).
__global__ __launch_bounds__(256) void k_probe_valu(uint32_t iters, uint32_t seed, uint32_t *sink)
{
  uint32_t a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = seed + threadIdx.x * 8 + k;
  for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a[k] += it;
      a[k] ^= a[(k + 1) & 7] >> 1;
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) r ^= a[k];
  if (r == 0x9e3779b9u) sink[blockIdx.x] = r;
}

hipError_t probe_lds(uint32_t blocks, uint32_t iters, uint32_t *sink, hipStream_t s)
{
  hipLaunchKernelGGL(k_probe_lds, dim3(blocks), dim3(256), 0, s, iters, sink);
  return hipGetLastError();
}

hipError_t probe_valu(uint32_t blocks, uint32_t iters, uint32_t *sink, hipStream_t s)
{
  hipLaunchKernelGGL(k_probe_valu, dim3(blocks), dim3(256), 0, s, iters, 12345u, sink);
  return hipGetLastError();
}
