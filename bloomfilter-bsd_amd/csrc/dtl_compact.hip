// dtl_compact.hip -- order-preserving compaction of the match bitmap.
//
// Phase 2 (scan): one workgroup per task scans the task's tile counts in
// place into exclusive offsets and writes the task's match count (single
// lists of up to 8192 tiles skip it: their expand adds up the counts itself).
// Phase 3 (expand): one workgroup per group of up to 8 tiles; 256 threads
// each own 8 keys (one byte of the bitmap) of every tile of the group, one
// workgroup exclusive scan of the byte popcounts gives every match its rank,
// and positions i + match_offset are written in key order: exactly the
// ascending selection vector of the reference
// (blocked_bloomfilter_batch_dispatch.hpp:47-67).
#include <hip/hip_runtime.h>

#include <algorithm>

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
constexpr int kExpandGroup = 8;  // tiles per workgroup of an expand, at most

// The part of an expand that does not depend on where the matches go: each
// thread's bitmap byte of every tile of the group, and its rank among the
// group's matches (per tile: 16-bit fields, tiles 0-3 in ex[0], 4-7 in ex[1];
// the tiles' totals in tot[0], tot[1] the same way).
struct ExpandPrep {
  uint32_t base[kExpandGroup];
  uint32_t b8[kExpandGroup];
  uint64_t ex[2], tot[2];
  uint32_t kb;
};

template <class OffOf>
__device__ __forceinline__ void expand_prep(const uint64_t* __restrict__ bits, uint32_t t0, uint32_t nt,
                                            OffOf&& off_of, uint64_t* sh, ExpandPrep& P) {
  const uint32_t t = threadIdx.x;
#pragma unroll
  for (int r = 0; r < kExpandGroup; ++r) P.base[r] = off_of(static_cast<uint32_t>(r) < nt ? r : nt - 1u);
  // Every tile slot loads (tiles past the group's end re-read its last tile
  // and are masked): straight-line loads, no branches around them.  A form
  // that skipped the loads of missing tiles behind branches gave wrong
  // bytes for tiles 4..7 of full groups on gfx950 (ROCm 7.2, -O3); this
  // form is exact (tests/test_gpu_parity.py::test_large_batches_exact).
#pragma unroll
  for (int r = 0; r < kExpandGroup; ++r) {
    const uint32_t rr = static_cast<uint32_t>(r) < nt ? static_cast<uint32_t>(r) : nt - 1u;
    const uint64_t w = bits[static_cast<uint64_t>(t0 + rr) * kTileWords + (t >> 3)];
    const uint64_t word = static_cast<uint32_t>(r) < nt ? w : 0ull;
    P.b8[r] = static_cast<uint32_t>(word >> ((t & 7u) * 8u)) & 0xFFu;
  }
  uint64_t lo = 0, hi = 0;
#pragma unroll
  for (int r = 0; r < kExpandGroup; ++r) {
    const uint64_t c = static_cast<uint64_t>(__popc(P.b8[r])) << (16 * (r % 4));
    if (r < 4) lo |= c; else hi |= c;
  }
  P.ex[0] = block_excl_scan64<kExpandThreads>(lo, P.tot[0], sh);
  P.ex[1] = block_excl_scan64<kExpandThreads>(hi, P.tot[1], sh);
  P.kb = (t >> 3) * 64u + (t & 7u) * 8u;
}

__device__ __forceinline__ uint32_t expand_field(const uint64_t* v, int r) {
  return static_cast<uint32_t>((v[r / 4] >> (16 * (r % 4))) & 0xFFFFu);
}

// Expands up to kExpandGroup consecutive tiles [t0, t0 + nt) at once.  Every
// thread loads its bitmap byte of each tile first (all loads in flight
// together), one workgroup scan ranks the matches of all tiles (the 8
// per-tile byte popcounts packed 16 bits each into two 64-bit values; a tile
// has at most 2048 matches, so fields never carry into each other), and
// value(tile, key index inside the tile) is written for every match, in key
// order within its tile, to out_of(tile)[slot].  off_of(r) is the first
// output slot of tile t0 + r; it is read before the scans, so its latency
// overlaps theirs.  STAGE forms: where the group averages at least kStageMin
// matches per tile, each tile's values are first placed in LDS (stage,
// kTileKeys words) and then written out by consecutive lanes to consecutive
// slots; sparser groups store straight from the bit loop (a lane's few
// stores are spread over the tile's output, and the barriers of the staged
// form would cost more than they save).  The expands of the short-list and
// batched compactions do not stage: their 8 KiB of LDS per workgroup made
// the 4096-filter semi-join plan 3.6% slower (the expands share the CUs with
// LDS-staged probe launches of other groups) and the plain calls 1.5% slower.
constexpr uint32_t kStageMin = 256;

template <bool STAGE, class OffOf, class Value, class OutOf>
__device__ __forceinline__ void expand_group(const uint64_t* __restrict__ bits, uint32_t t0, uint32_t nt,
                                             OffOf&& off_of, uint64_t* sh, uint32_t* stage, Value&& value,
                                             OutOf&& out_of) {
  ExpandPrep P;
  expand_prep(bits, t0, nt, off_of, sh, P);
  uint32_t total = 0;
#pragma unroll
  for (int r = 0; r < kExpandGroup; ++r) total += static_cast<uint32_t>(r) < nt ? expand_field(P.tot, r) : 0u;
  if (!STAGE || total < nt * kStageMin) {  // uniform
#pragma unroll
    for (int r = 0; r < kExpandGroup; ++r) {
      if (static_cast<uint32_t>(r) >= nt) break;
      uint32_t* out = out_of(t0 + r);
      uint32_t o = P.base[r] + expand_field(P.ex, r);
      uint32_t b = P.b8[r];
      while (b) {
        const uint32_t j = __ffs(b) - 1u;
        out[o++] = value(t0 + r, P.kb + j);
        b &= b - 1u;
      }
    }
    return;
  }
  if constexpr (!STAGE) return;
#pragma unroll
  for (int r = 0; r < kExpandGroup; ++r) {
    if (static_cast<uint32_t>(r) >= nt) break;  // uniform
    uint32_t o = expand_field(P.ex, r);
    uint32_t b = P.b8[r];
    while (b) {
      const uint32_t j = __ffs(b) - 1u;
      stage[o++] = value(t0 + r, P.kb + j);
      b &= b - 1u;
    }
    __syncthreads();
    const uint32_t cnt = expand_field(P.tot, r);
    uint32_t* dst = out_of(t0 + r) + P.base[r];
    for (uint32_t i = threadIdx.x; i < cnt; i += kExpandThreads) dst[i] = stage[i];
    __syncthreads();
  }
}

// Tiles per workgroup of the single-list expands: up to 8 for long lists,
// fewer for short ones so that the grid still covers every CU several times.
static uint32_t expand_tiles(uint32_t ntiles) {
  return ntiles >= 65536 ? 8u : ntiles >= 16384 ? 4u : ntiles >= 4096 ? 2u : 1u;
}

// Writes src[i] (a gathered value) for every match i instead of i itself,
// at out[*d_base + rank] (d_base null: at out[rank]).
__global__ __launch_bounds__(kExpandThreads) void k_expand_gather(const uint64_t* __restrict__ bits,
                                                                  const uint32_t* __restrict__ offs,
                                                                  uint32_t ntiles, uint32_t per_wg,
                                                                  const uint32_t* __restrict__ src,
                                                                  uint32_t* __restrict__ out,
                                                                  const uint64_t* __restrict__ d_base) {
  __shared__ uint64_t sh[kExpandThreads / 64];
  __shared__ uint32_t stage[kTileKeys];
  const uint32_t t0 = blockIdx.x * per_wg;
  uint32_t* dst = out + (d_base ? *d_base : 0ull);
  expand_group<true>(
      bits, t0, min(per_wg, ntiles - t0), [&](uint32_t r) { return offs[t0 + r]; }, sh, stage,
      [&](uint32_t tile, uint32_t i) { return src[static_cast<uint64_t>(tile) * kTileKeys + i]; },
      [&](uint32_t) { return dst; });
}

__global__ __launch_bounds__(kExpandThreads) void k_expand_single(const uint64_t* __restrict__ bits,
                                                                  const uint32_t* __restrict__ offs,
                                                                  uint32_t ntiles, uint32_t per_wg,
                                                                  uint32_t match_offset,
                                                                  uint32_t* __restrict__ pos) {
  __shared__ uint64_t sh[kExpandThreads / 64];
  __shared__ uint32_t stage[kTileKeys];
  const uint32_t t0 = blockIdx.x * per_wg;
  expand_group<true>(
      bits, t0, min(per_wg, ntiles - t0), [&](uint32_t r) { return offs[t0 + r]; }, sh, stage,
      [&](uint32_t tile, uint32_t i) { return tile * kTileKeys + i + match_offset; }, [&](uint32_t) { return pos; });
}

// Short lists (at most kFusedTiles tiles): the scan folded into the expand.
// Workgroup g adds up the tile counts before its group itself (at most
// kFusedTiles counts, read from L2), so no scan launch runs between the probe
// and the expand; the workgroup holding the last tile writes the total.  The
// counts are read, not overwritten.
constexpr uint32_t kFusedTiles = 8192;

__global__ __launch_bounds__(kExpandThreads) void k_expand_fused(const uint64_t* __restrict__ bits,
                                                                 const uint32_t* __restrict__ cnt,
                                                                 uint32_t ntiles, uint32_t per_wg,
                                                                 uint32_t match_offset,
                                                                 uint32_t* __restrict__ pos,
                                                                 uint64_t* __restrict__ d_count) {
  __shared__ uint64_t sh[kExpandThreads / 64];
  __shared__ uint32_t sh32[kExpandThreads / 64];
  __shared__ uint32_t s_off[kExpandGroup];
  const uint32_t t0 = blockIdx.x * per_wg;
  const uint32_t nt = min(per_wg, ntiles - t0);
  uint32_t part = 0;
#pragma unroll 8
  for (uint32_t j = threadIdx.x; j < t0; j += kExpandThreads) part += cnt[j];
  uint32_t before;
  (void)block_excl_scan<kExpandThreads>(part, before, sh32);  // before = sum of cnt[0 .. t0)
  if (threadIdx.x == 0) {
    uint32_t o = before;
    for (uint32_t r = 0; r < nt; ++r) {
      s_off[r] = o;
      o += cnt[t0 + r];
    }
    if (t0 + nt == ntiles) *d_count = o;
  }
  __syncthreads();
  expand_group<false>(
      bits, t0, nt, [&](uint32_t r) { return s_off[r]; }, sh, nullptr,
      [&](uint32_t tile, uint32_t i) { return tile * kTileKeys + i + match_offset; }, [&](uint32_t) { return pos; });
}

__global__ __launch_bounds__(kExpandThreads) void k_expand_tasks(const uint64_t* __restrict__ bits,
                                                                 const uint32_t* __restrict__ offs,
                                                                 const BatchTaskDev* __restrict__ tasks,
                                                                 const uint32_t* __restrict__ tile_task,
                                                                 uint32_t total_tiles, uint32_t per_wg) {
  __shared__ uint64_t sh[kExpandThreads / 64];
  const uint32_t g0 = blockIdx.x * per_wg;
  expand_group<false>(
      bits, g0, min(per_wg, total_tiles - g0), [&](uint32_t r) { return offs[g0 + r]; }, sh, nullptr,
      [&](uint32_t g, uint32_t i) {
        const BatchTaskDev& t = tasks[tile_task[g]];
        return (g - t.tile_begin) * kTileKeys + i + t.match_offset;
      },
      [&](uint32_t g) { return tasks[tile_task[g]].pos; });
}

__global__ void k_zero_count(uint64_t* d_count) { *d_count = 0; }

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }

hipError_t compact_single(const uint64_t* bits, uint32_t* tile_cnt, uint32_t ntiles, uint32_t n,
                          uint32_t match_offset, uint32_t* pos, uint64_t* d_count,
                          hipStream_t stream) {
  (void)n;
  if (ntiles <= kFusedTiles) {
    // fewer tiles per workgroup for shorter lists; the prefix sums then read
    // at most ntiles^2 / (2 per) counts in all
    const uint32_t per = ntiles >= 4096 ? 8u : ntiles >= 1024 ? 4u : ntiles >= 256 ? 2u : 1u;
    hipLaunchKernelGGL(k_expand_fused, dim3((ntiles + per - 1) / per), dim3(kExpandThreads), 0, stream, bits,
                       tile_cnt, ntiles, per, match_offset, pos, d_count);
    return hipGetLastError();
  }
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
                         uint32_t* out, const uint64_t* d_out_base, hipStream_t stream) {
  const uint32_t per = expand_tiles(ntiles);
  hipLaunchKernelGGL(k_expand_gather, dim3((ntiles + per - 1) / per), dim3(kExpandThreads), 0, stream, bits, offs,
                     ntiles, per, src, out, d_out_base);
  return hipGetLastError();
}

__global__ void k_add_counts(const uint64_t* a, const uint64_t* b, uint64_t* out) { *out = *a + *b; }

hipError_t add_counts(const uint64_t* a, const uint64_t* b, uint64_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(k_add_counts, dim3(1), dim3(1), 0, stream, a, b, out);
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
