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

// Reserve v[c] consecutive slots of *ctr[c] per thread, for each of the N
// counters, with one atomic per block and counter; slot[c] receives the
// thread's first slot.  Slots are consecutive within the block in thread
// order.  Every thread of the block must call it, at most once per kernel;
// blockDim.x must be a multiple of 64, at most 1024.
template <int N>
__device__ __forceinline__ void block_reserve(const int (&v)[N], int* const (&ctr)[N], int (&slot)[N]) {
  __shared__ int s_off[N][16];
  __shared__ int s_base[N];
  const int wave = (int)(threadIdx.x >> 6), n_waves = (int)(blockDim.x >> 6);
  int pre[N];
#pragma unroll
  for (int c = 0; c < N; c++) {
    int tot;
    pre[c] = wave_excl_scan(v[c], tot);
    if (lane_id() == 0) s_off[c][wave] = tot;
  }
  __syncthreads();
  if ((int)threadIdx.x < N) {
    const int c = (int)threadIdx.x;
    int run = 0;
    for (int w = 0; w < n_waves; w++) {
      const int t = s_off[c][w];
      s_off[c][w] = run;
      run += t;
    }
    s_base[c] = run > 0 ? atomicAdd(ctr[c], run) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < N; c++) slot[c] = s_base[c] + s_off[c][wave] + pre[c];
}

// v + v(lane ^ 16) and v + v(lane ^ 32) with the gfx950 row swaps: for a
// value passed as both operands, v_permlane16_swap returns rows {0,0,2,2}
// and {1,1,3,3}, v_permlane32_swap halves {0,0} and {1,1}, so the sum of the
// two results is the xor-partner sum on every lane (same two operands, so
// the same float result as v + partner).
__device__ __forceinline__ float xor16_sum(float v) {
  const unsigned u = __builtin_bit_cast(unsigned, v);
  const auto p = __builtin_amdgcn_permlane16_swap(u, u, false, false);
  return __builtin_bit_cast(float, (unsigned)p[0]) + __builtin_bit_cast(float, (unsigned)p[1]);
}
__device__ __forceinline__ float xor32_sum(float v) {
  const unsigned u = __builtin_bit_cast(unsigned, v);
  const auto p = __builtin_amdgcn_permlane32_swap(u, u, false, false);
  return __builtin_bit_cast(float, (unsigned)p[0]) + __builtin_bit_cast(float, (unsigned)p[1]);
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

// LDS written by lanes of a wave, read by other lanes of the same wave
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace fvc
