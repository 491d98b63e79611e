// wave.h -- wave64 collectives used for stream compaction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvc {

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Exclusive prefix sum of v over the wave; total receives the wave sum.
// Every lane of the wave must call it.
__device__ __forceinline__ int wave_excl_scan(int v, int& total) {
  const int lane = lane_id();
  int x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

// Reserve `v` consecutive slots of *counter per lane (one atomic per wave);
// returns this lane's first slot.  Every lane of the wave must call it.
__device__ __forceinline__ int wave_reserve(int v, int* counter) {
  int total;
  const int prefix = wave_excl_scan(v, total);
  int base = 0;
  if (lane_id() == 0 && total > 0) base = atomicAdd(counter, total);
  base = __shfl(base, 0, 64);
  return base + prefix;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const int o = __shfl_xor(v, d, 64);
    v = o < v ? o : v;
  }
  return v;
}

}  // namespace fvc
