// Exclusive prefix sum over u32 (grid cells, pair offsets, manifold compaction, adjacency, colour buckets) in one
// launch: 1024-element tiles per 256-thread block (4 per thread), wavefront-wide shuffles, then the 4 wave totals
// through LDS; each tile's offset comes from its predecessors' published totals (decoupled look-back).
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

// Single pass with decoupled look-back. Scratch layout: [0] tile ticket, [1] finished tiles, [2] waits given up,
// [3] unused, then one u64 status per tile: flag (bits 62-63: 1 = tile total, 2 = inclusive prefix) | value. Tiles
// take their index from the ticket in the order they start, so a tile waits only for tiles already running. The last
// tile to finish clears the statuses and counters, so the scratch is all zero again for the next scan.
constexpr uint64_t SCAN_TOTAL = 1ull << 62, SCAN_PREFIX = 2ull << 62;
constexpr uint32_t SCAN_SPIN = 1u << 20;

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_onepass(uint32_t *data, uint32_t n, uint32_t tiles,
                                                              uint32_t *state, uint32_t *abortFlag,
                                                              const uint32_t *gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t waveTot[SCAN_BLOCK / 64];
  __shared__ uint32_t tileShared, prefixShared, lastShared;
  uint64_t *status = reinterpret_cast<uint64_t *>(state + 4);
  if (threadIdx.x == 0) tileShared = atomicAdd(&state[0], 1u);
  __syncthreads();
  const uint32_t tile = tileShared;
  const uint32_t base = tile * SCAN_TILE + threadIdx.x * 4;
  uint32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = base + k < n ? data[base + k] : 0u;
  const uint32_t local = v[0] + v[1] + v[2] + v[3];
  const uint32_t inc = waveInclusive(local);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) waveTot[wave] = inc;
  __syncthreads();
  uint32_t waveOff = 0, total = 0;
#pragma unroll
  for (int k = 0; k < int(SCAN_BLOCK / 64); k++) {
    if (k < wave) waveOff += waveTot[k];
    total += waveTot[k];
  }
  if (wave == 0) {
    // Look back 64 tiles at a time: the nearest tile with an inclusive prefix ends the window's sum; until then every
    // tile in between must have published its total (else the window is polled again)
    uint32_t prefix = 0;
    if (tile == 0) {
      if (lane == 0) __hip_atomic_store(&status[0], SCAN_PREFIX | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(&status[tile], SCAN_TOTAL | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int64_t end = int64_t(tile) - 1;  // the window is tiles end, end - 1, ..., end - 63 (lane k reads end - k)
      for (uint32_t spin = 0;;) {
        const int64_t j = end - lane;
        const uint64_t st = j >= 0 ? __hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : SCAN_PREFIX;  // before tile 0: an empty prefix
        const uint64_t hasPrefix = __ballot((st >> 62) == 2u), missing = __ballot((st >> 62) == 0u);
        const uint32_t stop = hasPrefix ? uint32_t(__builtin_ctzll(hasPrefix)) : 64u;
        const uint64_t needed = stop >= 63u ? ~0ull : (2ull << stop) - 1ull;
        if (missing & needed) {
          if (++spin >= SCAN_SPIN) {  // never expected: those tiles are running; the scan is then not valid
            if (lane == 0) atomicOr(abortFlag ? abortFlag : &state[2], 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        uint32_t v = uint32_t(lane) <= stop ? uint32_t(st) : 0u;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
        prefix += v;
        if (stop < 64u) break;
        end -= 64;
      }
      if (lane == 0)
        __hip_atomic_store(&status[tile], SCAN_PREFIX | uint32_t(prefix + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) {
      prefixShared = prefix;
      // Release this tile's statuses before it counts as finished; the last tile acquires them all before it clears
      lastShared =
          __hip_atomic_fetch_add(&state[1], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == tiles - 1 ? 1u : 0u;
    }
  }
  __syncthreads();
  uint32_t run = prefixShared + waveOff + inc - local;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (lastShared) {  // every tile has finished its look-back
    for (uint32_t t = threadIdx.x; t < tiles; t += SCAN_BLOCK)
      __hip_atomic_store(&status[t], uint64_t(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
      state[0] = 0u;
      state[1] = 0u;
    }
  }
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

uint32_t scan_scratch_words(uint32_t n) { return 4u + 2u * ((n + SCAN_TILE - 1) / SCAN_TILE); }

void launch_scan_exclusive(uint32_t *data, uint32_t n, uint32_t *scratch, hipStream_t s, uint32_t *abortFlag,
                           const uint32_t *gate) {
  if (n == 0) return;
  const uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  k_scan_onepass<<<tiles, SCAN_BLOCK, 0, s>>>(data, n, tiles, scratch, abortFlag, gate);
}

}  // namespace avbd
