// Multi-level exclusive prefix sum over u32 (grid cells, pair offsets, manifold compaction, adjacency, colour buckets).
// 1024-element tiles per 256-thread block (4 per thread): wavefront-wide shuffles, then the 4 wave totals through LDS;
// tile totals are scanned recursively and added back.
#include "kernels.h"

namespace avbd {

constexpr uint32_t SCAN_BLOCK = 256, SCAN_TILE = 1024;

__device__ __forceinline__ uint32_t waveInclusive(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_tiles(uint32_t *data, uint32_t n, uint32_t *sums) {
  __shared__ uint32_t waveTot[SCAN_BLOCK / 64];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = base + k < n ? data[base + k] : 0u;
  uint32_t local = v[0] + v[1] + v[2] + v[3];
  uint32_t inc = waveInclusive(local);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) waveTot[wave] = inc;
  __syncthreads();
  uint32_t waveOff = 0, total = 0;
#pragma unroll
  for (int k = 0; k < int(SCAN_BLOCK / 64); k++) {
    if (k < wave) waveOff += waveTot[k];
    total += waveTot[k];
  }
  uint32_t run = waveOff + inc - local;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_add(uint32_t *data, uint32_t n, const uint32_t *sums) {
  const uint32_t add = sums[blockIdx.x];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (base + k < n) data[base + k] += add;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t value, uint64_t n) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = value;
}

void launch_fill_u32(uint32_t *p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  uint64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  k_fill_u32<<<uint32_t(g), 256, 0, s>>>(p, value, n);
}

uint32_t scan_scratch_words(uint32_t n) {
  uint32_t words = 0;
  while (n > 1) {
    uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    words += tiles;
    n = tiles;
  }
  return words + 1;
}

void launch_scan_exclusive(uint32_t *data, uint32_t n, uint32_t *scratch, hipStream_t s) {
  if (n == 0) return;
  uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  k_scan_tiles<<<tiles, SCAN_BLOCK, 0, s>>>(data, n, scratch);
  if (tiles > 1) {
    launch_scan_exclusive(scratch, tiles, scratch + tiles, s);
    k_scan_add<<<tiles, SCAN_BLOCK, 0, s>>>(data, n, scratch);
  }
}

}  // namespace avbd
