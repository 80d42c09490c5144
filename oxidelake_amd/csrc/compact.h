// Single-pass stream compaction over tiles: the decoupled look-back of
// Merrill and Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back" (NVIDIA tech report NVR-2016-002), run by one wave64 per tile.
// Tile ids come from an atomic ticket, so every tile a wave waits on has
// already started and will publish: every wait ends. Shared by the filter
// interpreter kernel (filter.hip) and the generated filter kernels
// (filter_jit.h); compiles under hipcc and hipRTC.
#pragma once

#include "scalar.h"

namespace oxr {

constexpr uint64_t kFlagA = 1ull << 62;  // tile aggregate available
constexpr uint64_t kFlagP = 2ull << 62;  // inclusive prefix available
constexpr uint64_t kValMask = (1ull << 62) - 1;

__device__ inline uint64_t atomic_load_u64(uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void atomic_store_u64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

__device__ inline uint64_t wave_sum64(uint64_t x) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// Decoupled look-back, run by one whole wave. Publishes `agg` for `tile`
// and returns the exclusive prefix of the tiles of its batch before it
// (`first` is the batch's first tile: every batch is compacted on its own).
__device__ inline uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t first, uint64_t agg, int lane) {
  if (tile == first) {
    if (lane == 0) atomic_store_u64(&status[tile], kFlagP | agg);
    return 0;
  }
  if (lane == 0) atomic_store_u64(&status[tile], kFlagA | agg);
  uint64_t excl = 0;
  int64_t end = (int64_t)tile - 1;  // nearest predecessor in this window
  const int64_t lo = (int64_t)first;
  while (true) {
    const int64_t idx = end - lane;
    uint64_t st = 0;
    if (idx >= lo) {
      // every predecessor took its ticket before this tile did, so it is
      // running and will publish: the wait ends
      do {
        st = atomic_load_u64(&status[idx]);
      } while ((st >> 62) == 0);
    }
    const uint64_t pmask = __ballot(idx >= lo && (st >> 62) == 2);
    if (pmask) {
      const int nearest = __builtin_ctzll(pmask);
      uint64_t v = (lane <= nearest && idx >= lo) ? (st & kValMask) : 0;
      excl += wave_sum64(v);
      break;
    }
    excl += wave_sum64(idx >= lo ? (st & kValMask) : 0);
    end -= 64;
  }
  if (lane == 0) atomic_store_u64(&status[tile], kFlagP | (excl + agg));
  return excl;
}

}  // namespace oxr
