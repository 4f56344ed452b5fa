// dtl_gather.hip -- random-sector read microbenchmark (roofline reference).
//
// Reads n uniformly scattered, aligned sectors of a device buffer the way a
// filter probe reads its blocks, and reports nothing but time: it measures
// the achievable random-access bandwidth the probe kernels are compared
// with.  Access patterns (mode):
//   0  one lane per 64 B sector, four 16-byte loads
//   1  as 0 with non-temporal loads
//   2  four lanes per 64 B sector, one 16-byte load each (one 64 B request)
//   3  as 2 with non-temporal loads
//   4  one lane per 32 B sector, two 16-byte loads
//   5  two lanes per 32 B sector, one 16-byte load each
//   6  as 2, sectors drawn by a full-avalanche hash of the read index
//   7  as 0, sectors drawn by the same hash
//   8  LDS: every workgroup stages the whole buffer (at most 128 KiB) in LDS,
//      then each lane reads pseudo-random 4-byte words of it (the access of
//      a one-word-block filter probed from LDS)
//   9  LDS as 8, but each lane reads a pseudo-random 64 B block as four
//      16-byte pieces, piece p of block b from slot 4 b + ((p + b / 4) mod 4)
//      (the rotated layout in which the probe stages 64-byte blocks)
// Modes 0-5 pick sector (i * 0x9E3779B1) >> shift, a low-discrepancy
// sequence that spreads consecutive reads evenly over the memory channels;
// modes 6-7 pick independent uniform sectors, like the blocks of random keys.
#include <hip/hip_runtime.h>

#include "dtl_hip.h"
#include "dtl_internal.hpp"

namespace dtlh {

constexpr int kGatherPerLane = 8;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t sector_of(uint64_t i, uint32_t shift) {
  return (static_cast<uint32_t>(i) * 0x9E3779B1u) >> shift;
}

// MurmurHash3's 32-bit finalizer (fmix32, Austin Appleby, public domain).
__device__ __forceinline__ uint32_t sector_of_random(uint64_t i, uint32_t shift) {
  uint32_t h = static_cast<uint32_t>(i);
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h >> shift;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_gather(const u32x4* __restrict__ buf, uint32_t shift,
                                                uint64_t n_reads, uint64_t* __restrict__ sink) {
  constexpr bool NT = (MODE == 1 || MODE == 3);
  constexpr bool RANDOM = (MODE == 6 || MODE == 7);
  constexpr int LANES = (MODE == 2 || MODE == 3 || MODE == 6) ? 4 : (MODE == 5 ? 2 : 1);
  constexpr int V = (MODE == 4) ? 2 : ((MODE == 5 || LANES == 4) ? 1 : 4);  // uint4 per lane per sector
  constexpr int SEC_V = (MODE == 4 || MODE == 5) ? 2 : 4;                    // uint4 per sector
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t group = gid / LANES;
  const uint32_t part = static_cast<uint32_t>(gid % LANES);
  u32x4 acc = {0u, 0u, 0u, 0u};
  u32x4 v[kGatherPerLane][V];
#pragma unroll
  for (int r = 0; r < kGatherPerLane; ++r) {
    const uint64_t i = group * kGatherPerLane + r;
    const uint64_t s = i < n_reads ? (RANDOM ? sector_of_random(i, shift) : sector_of(i, shift)) : 0;
    const u32x4* p = buf + s * SEC_V + part * V;
#pragma unroll
    for (int j = 0; j < V; ++j) v[r][j] = NT ? __builtin_nontemporal_load(p + j) : p[j];
  }
#pragma unroll
  for (int r = 0; r < kGatherPerLane; ++r)
#pragma unroll
    for (int j = 0; j < V; ++j) acc ^= v[r][j];
  if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9E3779B1u) sink[0] = gid;  // keeps the loads alive
}

// BLOCKS: 64 B blocks (mode 9) instead of 4-byte words (mode 8); `shift`
// maps a 32-bit draw to a word or block index of the buffer.
template <bool BLOCKS>
__global__ __launch_bounds__(1024) void k_lds_gather(const u32x4* __restrict__ buf, uint32_t buf_bytes,
                                                     uint32_t shift, uint64_t n_reads,
                                                     uint64_t* __restrict__ sink) {
  constexpr int R = BLOCKS ? 4 : 8;  // reads per lane per iteration
  extern __shared__ u32x4 lds_buf[];
  for (uint32_t i = threadIdx.x; i < buf_bytes / 16; i += blockDim.x) lds_buf[i] = buf[i];
  __syncthreads();
  const uint32_t* words = reinterpret_cast<const uint32_t*>(lds_buf);
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x * R;
  // indexes: the top bits of a per-lane linear congruential stream
  // (Numerical Recipes' 1664525 / 1013904223), seeded by a full-avalanche
  // hash of the lane, so that the index arithmetic costs two operations
  uint32_t x = sector_of_random(gid, 0);
  uint32_t acc = 0;
  for (uint64_t base = gid * R; base < n_reads; base += stride) {
    if constexpr (BLOCKS) {
      u32x4 v[R][4];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        x = x * 1664525u + 1013904223u;
        const uint32_t b = x >> shift;
#pragma unroll
        for (int p = 0; p < 4; ++p) v[r][p] = lds_buf[b * 4u + ((static_cast<uint32_t>(p) + (b >> 2)) & 3u)];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const u32x4 t = v[r][0] ^ v[r][1] ^ v[r][2] ^ v[r][3];
        acc ^= (base + r < n_reads) ? (t.x ^ t.y ^ t.z ^ t.w) : 0u;
      }
    } else {
      uint32_t v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        x = x * 1664525u + 1013904223u;
        v[r] = words[x >> shift];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) acc ^= (base + r < n_reads) ? v[r] : 0u;
    }
  }
  if (acc == 0x9E3779B1u) sink[0] = acc;  // keeps the reads alive
}

template <int MODE>
static hipError_t launch_gather(const void* buf, uint32_t shift, uint64_t n, uint64_t* sink,
                                hipStream_t st) {
  constexpr int LANES = (MODE == 2 || MODE == 3 || MODE == 6) ? 4 : (MODE == 5 ? 2 : 1);
  const uint64_t threads = (n + kGatherPerLane - 1) / kGatherPerLane * LANES;
  const uint64_t blocks = (threads + 255) / 256;
  hipLaunchKernelGGL(k_gather<MODE>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, st,
                     static_cast<const u32x4*>(buf), shift, n, sink);
  return hipGetLastError();
}

}  // namespace dtlh

extern "C" int dtl_gather_bench(const void* d_buf, uint64_t buf_bytes, uint32_t mode, uint64_t n_reads,
                                uint64_t* d_sink, dtl_stream_t stream) {
  using namespace dtlh;
  const uint64_t sector = mode == 8 ? 4 : ((mode == 4 || mode == 5) ? 32 : 64);
  if (!d_buf || !d_sink || mode > 9 || buf_bytes < 2 * sector || (buf_bytes & (buf_bytes - 1)))
    return DTL_EINVAL;
  if (mode >= 8 && (buf_bytes < 64 || buf_bytes > (128u << 10))) return DTL_EINVAL;
  if ((n_reads + kGatherPerLane - 1) / kGatherPerLane * 4 / 256 > 0x7FFFFFFFull) return DTL_EINVAL;
  const uint64_t sectors = buf_bytes / sector;
  const uint32_t shift = 32u - (63u - static_cast<uint32_t>(__builtin_clzll(sectors)));
  if (sectors > (1ull << 32)) return DTL_EINVAL;
  (void)hipGetLastError();
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e = hipSuccess;
  if (mode >= 8) {
    int dev = 0, cus = 0;
    uint32_t max_lds = 0;
    const int rc = current_device(&dev, &cus, &max_lds);
    if (rc != DTL_OK) return rc;
    const uint32_t bytes = static_cast<uint32_t>(buf_bytes);
    if (bytes > max_lds) return DTL_EINVAL;
    auto fn = mode == 9 ? k_lds_gather<true> : k_lds_gather<false>;
    e = allow_dynamic_lds(reinterpret_cast<const void*>(fn), bytes);
    int per_cu = 0;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 1024, bytes);
    if (e == hipSuccess) {
      const uint32_t grid = static_cast<uint32_t>(std::max(per_cu, 1) * std::max(cus, 1));
      hipLaunchKernelGGL(fn, dim3(grid), dim3(1024), bytes, st, static_cast<const u32x4*>(d_buf),
                         bytes, shift, n_reads, d_sink);
      e = hipGetLastError();
    }
    if (e != hipSuccess) {
      set_last_error(e);
      return DTL_EHIP;
    }
    return DTL_OK;
  }
  switch (mode) {
    case 0: e = launch_gather<0>(d_buf, shift, n_reads, d_sink, st); break;
    case 1: e = launch_gather<1>(d_buf, shift, n_reads, d_sink, st); break;
    case 2: e = launch_gather<2>(d_buf, shift, n_reads, d_sink, st); break;
    case 3: e = launch_gather<3>(d_buf, shift, n_reads, d_sink, st); break;
    case 4: e = launch_gather<4>(d_buf, shift, n_reads, d_sink, st); break;
    case 5: e = launch_gather<5>(d_buf, shift, n_reads, d_sink, st); break;
    case 6: e = launch_gather<6>(d_buf, shift, n_reads, d_sink, st); break;
    default: e = launch_gather<7>(d_buf, shift, n_reads, d_sink, st); break;
  }
  if (e != hipSuccess) {
    set_last_error(e);
    return DTL_EHIP;
  }
  return DTL_OK;
}
