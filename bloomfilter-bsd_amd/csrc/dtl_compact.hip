// dtl_compact.hip -- order-preserving compaction of the match bitmap.
//
// Phase 2 (scan): one workgroup per task scans the task's tile counts in
// place into exclusive offsets and writes the task's match count.
// Phase 3 (expand): one workgroup per tile (8 tiles per workgroup for single
// lists); 256 threads each own 8 keys (one
// byte of the bitmap), a workgroup exclusive scan of the byte popcounts gives
// every match its rank, and positions i + match_offset are written in key
// order: exactly the ascending selection vector of the reference
// (blocked_bloomfilter_batch_dispatch.hpp:47-67).
#include <hip/hip_runtime.h>

#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"
#include "dtl_scan.hpp"

namespace dtlh {

constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;

__device__ void scan_range(uint32_t* cnt, uint32_t begin, uint32_t count, uint64_t* d_count,
                           uint32_t* sh) {
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < count; c0 += kScanThreads * kScanItems) {
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t idx = c0 + threadIdx.x * kScanItems + i;
      v[i] = idx < count ? cnt[begin + idx] : 0u;
      local += v[i];
    }
    uint32_t tot;
    uint32_t run = carry + block_excl_scan<kScanThreads>(local, tot, sh);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      const uint32_t idx = c0 + threadIdx.x * kScanItems + i;
      if (idx < count) cnt[begin + idx] = run;
      run += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) *d_count = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_single(uint32_t* cnt, uint32_t ntiles,
                                                              uint64_t* d_count) {
  __shared__ uint32_t sh[kScanThreads / 64];
  scan_range(cnt, 0, ntiles, d_count, sh);
}

// Multi-workgroup scan for long count arrays: chunk b = kScanChunk counts.
// Pass 1: partial[b] = sum of chunk b.  Pass 2: workgroup b adds up
// partial[0 .. b) itself (at most kScanTmpWords chunks), scans its chunk with
// that carry in place, and the last chunk writes the total.
constexpr uint32_t kScanChunk = kScanThreads * kScanItems;

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ cnt, uint32_t ntiles,
                                                              uint32_t* __restrict__ partial) {
  __shared__ uint32_t sh[kScanThreads / 64];
  const uint32_t c0 = blockIdx.x * kScanChunk;
  uint32_t local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t idx = c0 + threadIdx.x * kScanItems + i;
    local += idx < ntiles ? cnt[idx] : 0u;
  }
  uint32_t tot;
  (void)block_excl_scan<kScanThreads>(local, tot, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_finish(uint32_t* __restrict__ cnt, uint32_t ntiles,
                                                              const uint32_t* __restrict__ partial,
                                                              uint64_t* __restrict__ d_count) {
  __shared__ uint32_t sh[kScanThreads / 64];
  uint32_t pre = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += kScanThreads) pre += partial[j];
  uint32_t tot;
  (void)block_excl_scan<kScanThreads>(pre, tot, sh);
  uint32_t carry = tot;  // sum of partial[0 .. blockIdx.x)
  const uint32_t c0 = blockIdx.x * kScanChunk;
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t idx = c0 + threadIdx.x * kScanItems + i;
    v[i] = idx < ntiles ? cnt[idx] : 0u;
    local += v[i];
  }
  uint32_t run = carry + block_excl_scan<kScanThreads>(local, tot, sh);
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    const uint32_t idx = c0 + threadIdx.x * kScanItems + i;
    if (idx < ntiles) cnt[idx] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *d_count = carry + tot;
}

// Exclusive scan of tile_cnt[0 .. ntiles) in place, total to *d_count;
// tile_cnt[ntiles .. ntiles + kScanTmpWords) is scratch.
static hipError_t scan_counts(uint32_t* tile_cnt, uint32_t ntiles, uint64_t* d_count, hipStream_t stream) {
  if (ntiles <= 4 * kScanChunk) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanThreads), 0, stream, tile_cnt, ntiles, d_count);
    return hipGetLastError();
  }
  const uint32_t chunks = (ntiles + kScanChunk - 1) / kScanChunk;
  if (chunks > kScanTmpWords) return hipErrorInvalidValue;
  uint32_t* partial = tile_cnt + ntiles;
  hipLaunchKernelGGL(k_scan_reduce, dim3(chunks), dim3(kScanThreads), 0, stream, tile_cnt, ntiles, partial);
  hipLaunchKernelGGL(k_scan_finish, dim3(chunks), dim3(kScanThreads), 0, stream, tile_cnt, ntiles, partial,
                     d_count);
  return hipGetLastError();
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tasks(uint32_t* cnt,
                                                             const BatchTaskDev* tasks) {
  __shared__ uint32_t sh[kScanThreads / 64];
  const BatchTaskDev& t = tasks[blockIdx.x];
  scan_range(cnt, t.tile_begin, t.tile_count, t.count, sh);
}

constexpr int kExpandThreads = 256;
static_assert(kExpandThreads * 8 == kTileKeys, "one bitmap byte per thread");

__device__ __forceinline__ void expand_tile(const uint64_t* __restrict__ bits, uint32_t gtile,
                                            uint32_t ltile, uint32_t offset, uint32_t match_offset,
                                            uint32_t* __restrict__ pos, uint32_t* sh) {
  const uint32_t t = threadIdx.x;
  const uint64_t word = bits[static_cast<uint64_t>(gtile) * kTileWords + (t >> 3)];
  uint32_t b8 = static_cast<uint32_t>(word >> ((t & 7u) * 8u)) & 0xFFu;
  uint32_t tot;
  uint32_t out = offset + block_excl_scan<kExpandThreads>(__popc(b8), tot, sh);
  const uint32_t kb = ltile * kTileKeys + (t >> 3) * 64u + (t & 7u) * 8u;
  while (b8) {
    const uint32_t j = __ffs(b8) - 1u;
    pos[out++] = kb + j + match_offset;
    b8 &= b8 - 1u;
  }
}

// As expand_tile, but writes src[i] (a gathered value) instead of position i.
__device__ __forceinline__ void expand_tile_gather(const uint64_t* __restrict__ bits, uint32_t tile,
                                                   uint32_t offset, const uint32_t* __restrict__ src,
                                                   uint32_t* __restrict__ out, uint32_t* sh) {
  const uint32_t t = threadIdx.x;
  const uint64_t word = bits[static_cast<uint64_t>(tile) * kTileWords + (t >> 3)];
  uint32_t b8 = static_cast<uint32_t>(word >> ((t & 7u) * 8u)) & 0xFFu;
  uint32_t tot;
  uint32_t o = offset + block_excl_scan<kExpandThreads>(__popc(b8), tot, sh);
  const uint64_t kb = static_cast<uint64_t>(tile) * kTileKeys + (t >> 3) * 64u + (t & 7u) * 8u;
  while (b8) {
    const uint32_t j = __ffs(b8) - 1u;
    out[o++] = src[kb + j];
    b8 &= b8 - 1u;
  }
}

// Tiles per workgroup of the single-list expands: up to 8 for long lists,
// fewer for short ones so that the grid still covers every CU several times.
static uint32_t expand_tiles(uint32_t ntiles) {
  return ntiles >= 65536 ? 8u : ntiles >= 16384 ? 4u : ntiles >= 4096 ? 2u : 1u;
}

__global__ __launch_bounds__(kExpandThreads) void k_expand_gather(const uint64_t* __restrict__ bits,
                                                                  const uint32_t* __restrict__ offs,
                                                                  uint32_t ntiles, uint32_t per_wg,
                                                                  const uint32_t* __restrict__ src,
                                                                  uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[kExpandThreads / 64];
  const uint32_t t1 = min(ntiles, (blockIdx.x + 1) * per_wg);
  for (uint32_t t = blockIdx.x * per_wg; t < t1; ++t) expand_tile_gather(bits, t, offs[t], src, out, sh);
}

__global__ __launch_bounds__(kExpandThreads) void k_expand_single(const uint64_t* __restrict__ bits,
                                                                  const uint32_t* __restrict__ offs,
                                                                  uint32_t ntiles, uint32_t per_wg,
                                                                  uint32_t match_offset,
                                                                  uint32_t* __restrict__ pos) {
  __shared__ uint32_t sh[kExpandThreads / 64];
  const uint32_t t1 = min(ntiles, (blockIdx.x + 1) * per_wg);
  for (uint32_t t = blockIdx.x * per_wg; t < t1; ++t) expand_tile(bits, t, t, offs[t], match_offset, pos, sh);
}

__global__ __launch_bounds__(kExpandThreads) void k_expand_tasks(const uint64_t* __restrict__ bits,
                                                                 const uint32_t* __restrict__ offs,
                                                                 const BatchTaskDev* __restrict__ tasks,
                                                                 const uint32_t* __restrict__ tile_task,
                                                                 uint32_t total_tiles, uint32_t per_wg) {
  __shared__ uint32_t sh[kExpandThreads / 64];
  const uint32_t g1 = min(total_tiles, (blockIdx.x + 1) * per_wg);
  for (uint32_t g = blockIdx.x * per_wg; g < g1; ++g) {
    const BatchTaskDev& t = tasks[tile_task[g]];
    expand_tile(bits, g, g - t.tile_begin, offs[g], t.match_offset, t.pos, sh);
  }
}

__global__ void k_zero_count(uint64_t* d_count) { *d_count = 0; }

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }

hipError_t compact_single(const uint64_t* bits, uint32_t* tile_cnt, uint32_t ntiles, uint32_t n,
                          uint32_t match_offset, uint32_t* pos, uint64_t* d_count,
                          hipStream_t stream) {
  (void)n;
  hipError_t e = scan_counts(tile_cnt, ntiles, d_count, stream);
  if (e != hipSuccess) return e;
  const uint32_t per = expand_tiles(ntiles);
  hipLaunchKernelGGL(k_expand_single, dim3((ntiles + per - 1) / per), dim3(kExpandThreads), 0, stream, bits,
                     tile_cnt, ntiles, per, match_offset, pos);
  return hipGetLastError();
}

hipError_t scan_tile_counts(uint32_t* tile_cnt, uint32_t ntiles, uint64_t* d_total, hipStream_t stream) {
  return scan_counts(tile_cnt, ntiles, d_total, stream);
}

hipError_t expand_gather(const uint64_t* bits, const uint32_t* offs, uint32_t ntiles, const uint32_t* src,
                         uint32_t* out, hipStream_t stream) {
  const uint32_t per = expand_tiles(ntiles);
  hipLaunchKernelGGL(k_expand_gather, dim3((ntiles + per - 1) / per), dim3(kExpandThreads), 0, stream, bits, offs,
                     ntiles, per, src, out);
  return hipGetLastError();
}

hipError_t compact_tasks(const uint64_t* bits, uint32_t* tile_cnt, const BatchTaskDev* tasks,
                         const uint32_t* tile_task, uint32_t n_tasks, uint32_t total_tiles,
                         hipStream_t stream) {
  hipLaunchKernelGGL(k_scan_tasks, dim3(n_tasks), dim3(kScanThreads), 0, stream, tile_cnt, tasks);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || total_tiles == 0) return e;
  const uint32_t per = expand_tiles(total_tiles);
  hipLaunchKernelGGL(k_expand_tasks, dim3((total_tiles + per - 1) / per), dim3(kExpandThreads), 0, stream, bits,
                     tile_cnt, tasks, tile_task, total_tiles, per);
  return hipGetLastError();
}

hipError_t write_zero_count(uint64_t* d_count, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, stream, d_count);
  return hipGetLastError();
}

hipError_t zero_u32(uint32_t* p, hipStream_t stream) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, stream, p);
  return hipGetLastError();
}

}  // namespace dtlh
