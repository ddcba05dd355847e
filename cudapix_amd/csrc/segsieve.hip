// segsieve.hip -- the persistent segment kernel of a batch launch: every
// block takes tiles (segment, easy-leaf and P2-lookup tiles, segtiles.hpp)
// from the batch's list until it is empty.
#include "segtiles.hpp"

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
    if (tile.group == CPX_TILE_EASY) easy_tile<T, K>(bd, (uint32_t)tile.seg, lds, sh);
    else if (tile.group == CPX_TILE_P2) p2tab_tile<T, K>(bd, (uint32_t)tile.seg, lds);
    else hard_segment<T, K>(bd, tile, lds, sh);
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
