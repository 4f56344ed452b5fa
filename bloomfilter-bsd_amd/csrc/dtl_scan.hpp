// dtl_scan.hpp -- wavefront and workgroup prefix sums (64-wide waves).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dtlh {

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(v, d, 64);
    if (lane >= static_cast<uint32_t>(d)) v += t;
  }
  return v;
}

// Sum over the wave, in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Exclusive scan over the workgroup in thread order; `total` receives the sum.
template <int BT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total, uint32_t* sh) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint32_t x = lane < static_cast<uint32_t>(BT / 64) ? sh[lane] : 0u;
    x = wave_incl_scan(x);
    if (lane < static_cast<uint32_t>(BT / 64)) sh[lane] = x;
  }
  __syncthreads();
  const uint32_t before = wave ? sh[wave - 1] : 0u;
  total = sh[BT / 64 - 1];
  __syncthreads();
  return before + incl - v;
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(v, d, 64);
    if (lane >= static_cast<uint32_t>(d)) v += t;
  }
  return v;
}

// 64-bit exclusive scan over the workgroup (e.g. of several packed counters).
template <int BT>
__device__ __forceinline__ uint64_t block_excl_scan64(uint64_t v, uint64_t& total, uint64_t* sh) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t incl = wave_incl_scan64(v);
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint64_t x = lane < static_cast<uint32_t>(BT / 64) ? sh[lane] : 0ull;
    x = wave_incl_scan64(x);
    if (lane < static_cast<uint32_t>(BT / 64)) sh[lane] = x;
  }
  __syncthreads();
  const uint64_t before = wave ? sh[wave - 1] : 0ull;
  total = sh[BT / 64 - 1];
  __syncthreads();
  return before + incl - v;
}

}  // namespace dtlh
