// dtl_batched.hpp -- one persistent launch that probes many filters, each
// with its own probe batch (the semi-join pattern: thousands of per-partition
// filters of uneven size).
//
// Work items are tiles of kTileKeys keys; every task's keys form a contiguous
// run of global tiles.  Workgroups grab runs of `chunk` consecutive tiles from
// an atomic counter until the counter passes the last tile, so large tasks are
// spread over the whole GPU and small ones do not leave CUs idle.  Filters up
// to `lds_max` bytes are staged in LDS once per run of tiles of the same task;
// larger filters are probed from global memory (L2 / Infinity Cache / HBM),
// lane-split (split_probe_tile) where the configuration allows.
// Output is the same bitmap + per-tile count as the single-filter kernels, so
// the compaction phases are shared.  HALF: each lane requests half as many
// keys (one-lane tiles) or 16-byte pieces (lane-split tiles) before it tests
// them -- fewer registers, so more waves per SIMD; the autotuner picks.
#pragma once

#include <hip/hip_runtime.h>

#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"

namespace dtlh {

template <class C, int BT, bool HALF>
__global__ __launch_bounds__(BT) void k_bbf_probe_batched(const BatchTaskDev* __restrict__ tasks,
                                                          const uint32_t* __restrict__ tile_task,
                                                          uint32_t total_tiles, uint32_t chunk,
                                                          uint32_t* __restrict__ work_counter,
                                                          uint32_t lds_max,
                                                          uint64_t* __restrict__ bits,
                                                          uint32_t* __restrict__ tile_cnt) {
  using W = typename C::word_t;
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int UF = ROUNDS < kGlobalUnroll ? ROUNDS : kGlobalUnroll;
  constexpr int U = HALF && UF > 1 ? UF / 2 : UF;  // keys per lane requested before the test
  extern __shared__ uint4 lds_filter[];
  __shared__ uint32_t red[BT / 64];
  __shared__ uint32_t s_start;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  uint32_t staged = 0xFFFFFFFFu;
  // per-lane multipliers of the lane-split path (nothing for other configurations)
  [[maybe_unused]] const std::conditional_t<C::split.ok, SplitLane<C>, NoSplitLane> sl(lane & (C::split.lanes - 1));
  for (;;) {
    if (tid == 0) s_start = atomicAdd(work_counter, chunk);
    __syncthreads();
    const uint32_t g0 = s_start;
    __syncthreads();
    if (g0 >= total_tiles) break;
    const uint32_t g1 = min(g0 + chunk, total_tiles);
    for (uint32_t g = g0; g < g1; ++g) {
      const uint32_t ti = tile_task[g];
      const BatchTaskDev& T = tasks[ti];
      const uint32_t* keys = T.keys;
      const uint32_t n = T.n;
      AddrArgs addr;
      addr.magic = T.addr_magic;
      addr.mshift = T.addr_mshift;
      addr.divisor = T.addr_divisor;
      addr.post = T.addr_post;
      const uint32_t fbytes = T.filter_bytes;
      const W* filter = static_cast<const W*>(T.words);
      const bool use_lds = fbytes <= lds_max;
      if (use_lds && ti != staged) {
        LdsLayout<C>::stage(lds_filter, filter, fbytes, tid, BT);
        staged = ti;
        __syncthreads();
      }
      const uint64_t local_base = static_cast<uint64_t>(g - T.tile_begin) * kTileKeys;
      uint32_t cnt = 0;
      if constexpr (C::split.ok) {
        if (!use_lds) {
          // blocks of 32 B or more from L2 / HBM: several lanes per key
          cnt = split_probe_tile<C, BT, HALF ? DTL_SPLIT_UNROLL / 2 : DTL_SPLIT_UNROLL>(
              sl, filter, keys, n, addr, local_base, bits + static_cast<uint64_t>(g) * kTileWords);
          const uint32_t total = block_sum_waves<BT>(cnt, red);
          if (tid == 0) tile_cnt[g] = total;
          __syncthreads();
          continue;
        }
      }
#pragma unroll
      for (int r0 = 0; r0 < ROUNDS; r0 += U) {
        uint32_t key[U];
        bool valid[U];
        bool hit[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint64_t i = local_base + static_cast<uint64_t>(r0 + u) * BT + tid;
          valid[u] = i < n;
          key[u] = load_key(keys, valid[u] ? i : n - 1);  // tasks have n >= 1
        }
        if (use_lds || C::split.ok) {
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t b = block_of(key[u] * kMul.v[0], addr);
            hit[u] = LdsLayout<C>::test(key[u], b, lds_filter) & valid[u];
          }
        } else {
          W blk[U][C::w];
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const uint32_t b = block_of(key[u] * kMul.v[0], addr);
            load_block<C>(filter + static_cast<uint64_t>(b) * C::w, blk[u]);
          }
#pragma unroll
          for (int u = 0; u < U; ++u) hit[u] = test_regs<C>(key[u], blk[u]) & valid[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint64_t m = __ballot(hit[u]);
          if (lane == 0)
            store_bits(bits, static_cast<uint64_t>(g) * kTileWords + ((r0 + u) * BT) / 64 + wave, m);
          cnt += __popcll(m);
        }
      }
      const uint32_t total = block_sum_waves<BT>(cnt, red);
      if (tid == 0) tile_cnt[g] = total;
      __syncthreads();
    }
  }
}

// CUs x workgroups per CU of the launch shape (LDS reservation, workgroup
// size, keys in flight).
template <class C, int BT>
uint32_t batched_full_grid_bt(const BbfBatchLaunch& L) {
  auto fn = L.half_in_flight ? k_bbf_probe_batched<C, BT, true> : k_bbf_probe_batched<C, BT, false>;
  const uint32_t lds_bytes = (L.lds_max_bytes + 15u) & ~15u;
  int per_cu = 0;
  if (allow_dynamic_lds(reinterpret_cast<const void*>(fn), lds_bytes) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BT, lds_bytes) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(L.cu_count > 0 ? L.cu_count : 1);
}

template <class C>
uint32_t batched_full_grid(const BbfBatchLaunch& L) {
  if constexpr (sizeof(typename C::word_t) == 4) {
    if (L.threads == 1024) return batched_full_grid_bt<C, 1024>(L);
  }
  return batched_full_grid_bt<C, 256>(L);
}

template <class C, int BT>
hipError_t launch_probe_batched_bt(const BbfBatchLaunch& L, uint32_t* work_counter, uint32_t chunk) {
  auto fn = L.half_in_flight ? k_bbf_probe_batched<C, BT, true> : k_bbf_probe_batched<C, BT, false>;
  const uint32_t lds_bytes = (L.lds_max_bytes + 15u) & ~15u;
  hipError_t ae = allow_dynamic_lds(reinterpret_cast<const void*>(fn), lds_bytes);
  if (ae != hipSuccess) return ae;
  uint32_t grid = L.grid;
  if (grid == 0) grid = batched_full_grid_bt<C, BT>(L);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(BT), lds_bytes, L.stream, L.tasks, L.tile_task,
                     L.total_tiles, chunk, work_counter, L.lds_max_bytes, L.bits, L.tile_cnt);
  return hipGetLastError();
}

// `work_counter` must be zero when the launch starts; the caller clears it
// with a kernel on the same stream (graph-capturable).
template <class C>
hipError_t launch_probe_batched(const BbfBatchLaunch& L) {
  const uint32_t chunk = L.chunk ? L.chunk : 1u;
  // 1024-thread workgroups only for 32-bit words (as launch_probe's LDS path)
  if constexpr (sizeof(typename C::word_t) == 4) {
    if (L.threads == 1024) return launch_probe_batched_bt<C, 1024>(L, L.work_counter, chunk);
  }
  return launch_probe_batched_bt<C, 256>(L, L.work_counter, chunk);
}

}  // namespace dtlh
