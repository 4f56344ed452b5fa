// dtl_launch.hpp -- host launchers of the templated kernels and the
// compile-time enumeration of every configuration the reference dispatches.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dtl_batched.hpp"
#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"

namespace dtlh {

inline AddrArgs make_addr(uint32_t magic, uint32_t mshift, uint32_t divisor, uint32_t post) {
  AddrArgs a;
  a.magic = magic;
  a.mshift = mshift;
  a.divisor = divisor;
  a.post = post;
  return a;
}

template <class C, int BT>
hipError_t launch_probe_lds(const BbfProbeLaunch& L) {
  using W = typename C::word_t;
  auto fn = L.addr_magic == 0 ? k_bbf_probe_lds<C, BT, true> : k_bbf_probe_lds<C, BT, false>;
  const uint32_t lds_bytes = (L.filter_bytes + 15u) & ~15u;
  hipError_t ae = allow_dynamic_lds(reinterpret_cast<const void*>(fn), lds_bytes);
  if (ae != hipSuccess) return ae;
  uint32_t grid = L.lds_grid;
  if (grid == 0) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BT, lds_bytes);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    grid = static_cast<uint32_t>(per_cu) * static_cast<uint32_t>(std::max(L.cu_count, 1));
  }
  grid = std::max(1u, std::min(grid, L.ntiles));
  hipLaunchKernelGGL(fn, dim3(grid), dim3(BT), lds_bytes, L.stream,
                     static_cast<const W*>(L.words), L.filter_bytes, L.keys, L.n,
                     make_addr(L.addr_magic, L.addr_mshift, L.addr_divisor, L.addr_post),
                     L.bits, L.tile_cnt, L.ntiles);
  return hipGetLastError();
}

template <class C>
hipError_t launch_probe(const BbfProbeLaunch& L) {
  using W = typename C::word_t;
  if (L.use_lds) {
    // 1024-thread workgroups only for 32-bit words: with 64-bit words that
    // shape lost whole 64-key ballot words now and then on gfx950 (ROCm 7.2;
    // cause not found: it persists with the blocks read from global memory
    // and with the bitmap stores deferred), the 256-thread shape never did.
    if constexpr (sizeof(W) == 4) {
      if (L.lds_threads == 1024) return launch_probe_lds<C, 1024>(L);
    }
    return launch_probe_lds<C, 256>(L);
  }
  auto fn = k_bbf_probe_global<C>;
  if constexpr (C::split.ok) {
    if (L.split) fn = k_bbf_probe_split<C>;
  }
  hipLaunchKernelGGL(fn, dim3(L.ntiles), dim3(kGlobalThreads), 0, L.stream,
                     static_cast<const W*>(L.words), L.keys, L.n,
                     make_addr(L.addr_magic, L.addr_mshift, L.addr_divisor, L.addr_post),
                     L.bits, L.tile_cnt, 0u, L.ntiles, static_cast<const uint32_t*>(nullptr), 0u);
  return hipGetLastError();
}

// Probe phase of the partitioned path: the global (or lane-split) kernel
// over the grouped keys (L.ptiles tiles; with capacity bins, each tile's
// keys end at its bin's fill), XCD-ordered, into the partitioned-order bitmap.
template <class C>
hipError_t launch_probe_part(const BbfPartLaunch& L) {
  using W = typename C::word_t;
  const uint32_t per_xcd = (L.ptiles + 7u) / 8u;
  auto fn = k_bbf_probe_global<C>;
  if constexpr (C::split.ok) {
    if (L.split) fn = k_bbf_probe_split<C>;
  }
  const uint32_t* fill = L.bin_cap ? L.cursor : nullptr;
  hipLaunchKernelGGL(fn, dim3(per_xcd * 8u), dim3(kGlobalThreads), 0, L.stream,
                     static_cast<const W*>(L.words), L.keys_p, L.n,
                     make_addr(L.addr_magic, L.addr_mshift, L.addr_divisor, L.addr_post), L.pbits,
                     L.ptile_cnt, per_xcd, L.ptiles, fill, L.bin_cap / kTileKeys);
  return hipGetLastError();
}

template <class C>
hipError_t launch_build(const BbfBuildLaunch& L) {
  using W = typename C::word_t;
  const uint64_t want = (L.n + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(std::max(L.cu_count, 1)) * 64;
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
  hipLaunchKernelGGL(k_bbf_build<C>, dim3(grid), dim3(256), 0, L.stream, static_cast<W*>(L.words),
                     L.keys, L.n,
                     make_addr(L.addr_magic, L.addr_mshift, L.addr_divisor, L.addr_post));
  return hipGetLastError();
}

template <class C>
BbfEntry make_entry() {
  BbfEntry e;
  e.word_bits = C::word_bits;
  e.w = C::w;
  e.s = C::s;
  e.z = C::z;
  e.k = C::k;
  e.probe = &launch_probe<C>;
  e.build = &launch_build<C>;
  e.probe_batched = &launch_probe_batched<C>;
  e.probe_part = &launch_probe_part<C>;
  e.batched_grid = &batched_full_grid<C>;
  e.split_lanes = C::split.ok ? static_cast<uint32_t>(C::split.lanes) : 0u;
  return e;
}

// Every (s, k) the reference's sectorized filter accepts for this word type and w.
template <class W, uint32_t Wc>
void register_sectorized(std::vector<BbfEntry>& out) {
  static_for<5>([&](auto si) {
    constexpr uint32_t S = 1u << decltype(si)::value;
    static_for<16>([&](auto ki) {
      constexpr uint32_t K = decltype(ki)::value + 1;
      if constexpr (valid_sgew_sltw(sizeof(W) * 8, Wc, S, K)) {
        out.push_back(make_entry<BbfCfg<W, Wc, S, 1, K>>());
      }
    });
  });
}

// Every (z, k) the reference's zoned filter accepts for this word type and w.
template <class W, uint32_t Wc>
void register_zoned(std::vector<BbfEntry>& out) {
  static_for<3>([&](auto zi) {
    constexpr uint32_t Z = 2u << decltype(zi)::value;
    static_for<8>([&](auto ki) {
      constexpr uint32_t K = 2 * (decltype(ki)::value + 1);
      if constexpr (valid_zoned(sizeof(W) * 8, Wc, Z, K)) {
        out.push_back(make_entry<BbfCfg<W, Wc, Wc, Z, K>>());
      }
    });
  });
}

}  // namespace dtlh
