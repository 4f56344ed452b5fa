// dtl_kernels.hpp -- gfx950 kernels of the filter probe path (templates).
//
// Work decomposition (all probe kernels):
//   * keys are cut into tiles of kTileKeys consecutive keys; a workgroup
//     handles whole tiles, one key per lane per round, rounds of one key
//     per thread so each wave reads 64 consecutive keys (256 B, coalesced);
//   * each wave turns its 64 hit flags into one 64-bit ballot word of the
//     match bitmap (bit i <-> key i) and the workgroup writes one match count
//     per tile;
//   * compaction (dtl_compact.hip) scans the tile counts and expands the
//     bitmap into ascending positions, so the list comes out in the
//     reference's order without atomics.
// Filter access:
//   * "global" kernels load the key's whole block (<= 128 B) into VGPRs with
//     16-byte loads; the filter stays in HBM / Infinity Cache / L2;
//   * "lds" kernels first stage the whole filter in LDS (filters up to the
//     workgroup LDS limit, 160 KiB on gfx950) and then loop over tiles
//     (persistent grid), reading only the words a key needs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <type_traits>

#include "dtl_internal.hpp"
#include "dtl_sched.hpp"

namespace dtlh {


// Block addressing (block_addressing_logic.hpp:104-110 MAGIC, :205-210 POW2),
// evaluated branch-free as ((h - (mulhi(h, magic) >> mshift) * divisor) >> post):
//   POW2:  magic = divisor = 0, post = 32 - log2(block_cnt)  ->  h >> post
//   MAGIC: post = 0                                          ->  h mod divisor
struct AddrArgs {
  uint32_t magic;
  uint32_t mshift;
  uint32_t divisor;
  uint32_t post;
};

__device__ __forceinline__ uint32_t block_of(uint32_t h0, const AddrArgs& a) {
  const uint32_t q = __umulhi(h0, a.magic) >> a.mshift;
  return (h0 - q * a.divisor) >> a.post;
}

// The hash values a key's chunks are cut from, each computed once:
// h[i] = key * kMul[i] for 1 <= i <= NSRC, h[0] = 0.  One product per hash
// function matters on gfx950, where a 32-bit multiply issues at a quarter of
// the rate of other VALU operations: written as one expression per chunk, the
// compiler narrows each product's constant to the bits that chunk needs
// (different constants) and can no longer share the products.
template <int NSRC>
struct KeyHashes {
  uint32_t h[NSRC + 1];
  __device__ __forceinline__ explicit KeyHashes(uint32_t key) {
    h[0] = 0;
    static_for<NSRC>([&](auto i) { h[decltype(i)::value + 1] = key * kMul.v[decltype(i)::value + 1]; });
  }
  __device__ __forceinline__ uint32_t chunk(Chunk c) const {
    return ((h[c.src] >> c.shift) & ((1u << c.nbits) - 1u)) + c.base;
  }
};

template <class C>
using HashesOf = KeyHashes<C::sched.max_src>;

template <class W>
__device__ __forceinline__ W one_at(uint32_t bit) {
  return static_cast<W>(1) << bit;
}

// Per-group search masks of a key.
template <class C>
__device__ __forceinline__ void key_masks(const HashesOf<C>& H, typename C::word_t (&mask)[C::sched.ngroups]) {
  using W = typename C::word_t;
  static_for<C::sched.ngroups>([&](auto g) { mask[g] = 0; });
  static_for<C::sched.nbits>([&](auto b) {
    constexpr Chunk c = C::sched.bbit[b];
    constexpr int g = C::sched.bgroup[b];
    mask[g] |= one_at<W>(H.chunk(c));
  });
}

template <int BASE, int SPAN, class W, int N>
__device__ __forceinline__ W select_word(const W (&blk)[N], uint32_t off) {
  W r = blk[BASE];
  static_for<SPAN - 1>([&](auto j) {
    constexpr int J = decltype(j)::value + 1;
    r = (off == static_cast<uint32_t>(J)) ? blk[BASE + J] : r;
  });
  return r;
}

// Loads a whole block (w words) into registers with the widest aligned loads.
template <class C>
__device__ __forceinline__ void load_block(const typename C::word_t* __restrict__ p,
                                           typename C::word_t (&blk)[C::w]) {
  using W = typename C::word_t;
  constexpr int BYTES = C::w * sizeof(W);
  if constexpr (BYTES >= 16) {
    uint4 v[BYTES / 16];
    const uint4* q = reinterpret_cast<const uint4*>(p);
    static_for<BYTES / 16>([&](auto i) { v[i] = q[i]; });
    __builtin_memcpy(blk, v, BYTES);
  } else if constexpr (BYTES == 8) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    __builtin_memcpy(blk, &v, 8);
  } else {
    const uint32_t v = *reinterpret_cast<const uint32_t*>(p);
    __builtin_memcpy(blk, &v, 4);
  }
}

// Hit test against a block held in registers.
template <class C>
__device__ __forceinline__ bool test_regs(uint32_t key, const typename C::word_t (&blk)[C::w]) {
  using W = typename C::word_t;
  const HashesOf<C> H(key);
  W mask[C::sched.ngroups];
  key_masks<C>(H, mask);
  bool ok = true;
  static_for<C::sched.ngroups>([&](auto g) {
    constexpr Chunk d = C::sched.gdyn[g];
    constexpr int base = C::sched.gword[g];
    W word;
    if constexpr (d.nbits == 0) {
      word = blk[base];
    } else {
      word = select_word<base, (1 << d.nbits)>(blk, H.chunk(d));
    }
    ok &= (word & mask[g]) == mask[g];
  });
  return ok;
}

// Hit test reading only the needed words (LDS or global pointer to the block).
template <class C, class P>
__device__ __forceinline__ bool test_ptr(uint32_t key, P blk) {
  using W = typename C::word_t;
  const HashesOf<C> H(key);
  W mask[C::sched.ngroups];
  key_masks<C>(H, mask);
  bool ok = true;
  static_for<C::sched.ngroups>([&](auto g) {
    constexpr Chunk d = C::sched.gdyn[g];
    constexpr int base = C::sched.gword[g];
    uint32_t wi = base;
    if constexpr (d.nbits != 0) wi += H.chunk(d);
    const W word = blk[wi];
    ok &= (word & mask[g]) == mask[g];
  });
  return ok;
}

__device__ __forceinline__ uint32_t load_key(const uint32_t* __restrict__ keys, uint64_t i) {
  return __builtin_nontemporal_load(keys + i);
}

// Ordinary store: the bitmap is read back by the compaction right after.
__device__ __forceinline__ void store_bits(uint64_t* __restrict__ bits, uint64_t i, uint64_t v) {
  bits[i] = v;
}

// Sums a per-wave uniform value over the workgroup; result valid in thread 0.
template <int BT>
__device__ __forceinline__ uint32_t block_sum_waves(uint32_t v, uint32_t* red) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  uint32_t s = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < BT / 64; ++i) s += red[i];
  }
  return s;
}

// ---------------------------------------------------------------------------
// Blocked Bloom filter probe, filter read from global memory.
// block = 256 threads (8 rounds x 4 waves = 32 bitmap words per tile).
// tiles_per_xcd == 0: workgroup b probes tile b (grid = ntiles).
// tiles_per_xcd > 0 (partitioned probe, keys grouped by filter slice):
// workgroup b runs on XCD b % 8 (round-robin dispatch) and probes tile
// (b % 8) * tiles_per_xcd + b / 8, so each XCD walks a contiguous range of
// the grouped keys and the filter slices of that range are cached in its
// own L2 one after another (grid = 8 * tiles_per_xcd).
// ---------------------------------------------------------------------------
// Keys grouped in bins of tiles_per_bin tiles (the partitioned probe of
// blocked Bloom filters): bin_fill[b] is the final cursor of bin b (an index
// of the grouped array, possibly past the bin's capacity); keys of the bin
// occupy [b * cap, min(bin_fill[b], (b + 1) * cap)).  Returns that end for
// the bin of `tile`.
__device__ __forceinline__ uint32_t bin_end_of(const uint32_t* __restrict__ bin_fill, uint32_t tile,
                                               uint32_t tiles_per_bin) {
  const uint32_t b = tile / tiles_per_bin;
  const uint32_t cap_end = (b + 1u) * tiles_per_bin * kTileKeys;
  const uint32_t f = bin_fill[b];
  return f < cap_end ? f : cap_end;
}

// A tile with no keys: its ballot words and count are zero.
__device__ __forceinline__ void clear_tile(uint64_t* __restrict__ bits, uint32_t* __restrict__ tile_cnt,
                                           uint32_t tile) {
  if (threadIdx.x < kTileWords) bits[static_cast<uint64_t>(tile) * kTileWords + threadIdx.x] = 0ull;
  if (threadIdx.x == 0) tile_cnt[tile] = 0u;
}

#ifndef DTL_GLOBAL_UNROLL
#define DTL_GLOBAL_UNROLL 8  // keys in flight per lane (build-time tuning knob; 8 measured fastest)
#endif
constexpr int kGlobalThreads = 256;
constexpr int kGlobalUnroll = DTL_GLOBAL_UNROLL;
static_assert(kTileKeys / kGlobalThreads % kGlobalUnroll == 0, "rounds per tile must be a multiple of U");

template <class C>
__global__ __launch_bounds__(kGlobalThreads) void k_bbf_probe_global(
    const typename C::word_t* __restrict__ filter, const uint32_t* __restrict__ keys, uint32_t n,
    AddrArgs addr, uint64_t* __restrict__ bits, uint32_t* __restrict__ tile_cnt,
    uint32_t tiles_per_xcd, uint32_t ntiles, const uint32_t* __restrict__ bin_fill, uint32_t tiles_per_bin) {
  using W = typename C::word_t;
  constexpr int BT = kGlobalThreads;
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int U = kGlobalUnroll;
  __shared__ uint32_t red[BT / 64];
  const uint32_t tile =
      tiles_per_xcd ? (blockIdx.x & 7u) * tiles_per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (tile >= ntiles) return;  // uniform per workgroup, before any barrier
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t tile_base = static_cast<uint64_t>(tile) * kTileKeys;
  if (bin_fill) {  // keys grouped in bins of fixed capacity: this tile's bin ends at its fill
    n = bin_end_of(bin_fill, tile, tiles_per_bin);
    if (n <= tile_base) {  // nothing of the bin in this tile (uniform, before any barrier)
      clear_tile(bits, tile_cnt, tile);
      return;
    }
  }
  uint32_t cnt = 0;
#pragma unroll
  for (int r0 = 0; r0 < ROUNDS; r0 += U) {
    uint32_t key[U];
    bool valid[U];
    W blk[U][C::w];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = tile_base + static_cast<uint64_t>(r0 + u) * BT + tid;
      valid[u] = i < n;
      key[u] = load_key(keys, valid[u] ? i : n - 1);  // n >= 1: clamp, no branch
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t b = block_of(key[u] * kMul.v[0], addr);
      load_block<C>(filter + static_cast<uint64_t>(b) * C::w, blk[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool hit = test_regs<C>(key[u], blk[u]) & valid[u];
      const uint64_t m = __ballot(hit);
      if (lane == 0) store_bits(bits, (tile_base + static_cast<uint64_t>(r0 + u) * BT) / 64 + wave, m);
      cnt += __popcll(m);
    }
  }
  const uint32_t total = block_sum_waves<BT>(cnt, red);
  if (tid == 0) tile_cnt[tile] = total;
}

// ---------------------------------------------------------------------------
// Lane-split probe (configurations with C::split.ok): L = block bytes / 16
// lanes share a key, lane `piece` = lane % L loads the key's 16-byte piece
// `piece` of the block and tests the groups that live there.  One 16-byte
// load instruction of a wave then covers 64 / L whole blocks instead of a
// quarter (64 B block) of 64 blocks, which is what bounds gathers from L2:
// the L1/TA handles an instruction per distinct cache line it touches.
// The key's hit is the AND of its L lanes' ballot bits (and_compress).
// Same grid, tile order and outputs as k_bbf_probe_global.
// ---------------------------------------------------------------------------
#ifndef DTL_SPLIT_UNROLL
#define DTL_SPLIT_UNROLL 16  // 16-byte loads per lane per batch (8 and 32 measured slower)
#endif

// Bit j of the result = AND of bits [j*L, (j+1)*L) of m (64 / L result bits).
template <int L>
__device__ __forceinline__ uint64_t and_compress(uint64_t m) {
  if constexpr (L == 2) {
    m &= m >> 1;
    m &= 0x5555555555555555ull;
    m = (m | (m >> 1)) & 0x3333333333333333ull;
    m = (m | (m >> 2)) & 0x0F0F0F0F0F0F0F0Full;
    m = (m | (m >> 4)) & 0x00FF00FF00FF00FFull;
    m = (m | (m >> 8)) & 0x0000FFFF0000FFFFull;
    m = (m | (m >> 16)) & 0x00000000FFFFFFFFull;
  } else if constexpr (L == 4) {
    m &= m >> 1;
    m &= m >> 2;
    m &= 0x1111111111111111ull;
    m = (m | (m >> 3)) & 0x0303030303030303ull;
    m = (m | (m >> 6)) & 0x000F000F000F000Full;
    m = (m | (m >> 12)) & 0x000000FF000000FFull;
    m = (m | (m >> 24)) & 0x000000000000FFFFull;
  } else {
    static_assert(L == 8, "2, 4 or 8 lanes per key");
    m &= m >> 1;
    m &= m >> 2;
    m &= m >> 4;
    m &= 0x0101010101010101ull;
    m = (m | (m >> 7)) & 0x0003000300030003ull;
    m = (m | (m >> 14)) & 0x0000000F0000000Full;
    m = (m | (m >> 28)) & 0x00000000000000FFull;
  }
  return m;
}

// Per-lane folded multipliers of the lane's piece: fb for the bits, fd for
// the dynamic word offsets (split_fold; piece selected at run time once).
template <class C>
struct SplitLane {
  static constexpr SplitSched P = C::split;
  static constexpr int NB = P.gpp * P.bpg;
  static constexpr int NSETS = P.lanes / P.span;
  uint32_t fb[NB];
  uint32_t fd[P.gpp];
  uint32_t sub;  // this lane's piece within its set (span > 1)
  __device__ __forceinline__ explicit SplitLane(uint32_t piece) {
    const uint32_t set = piece / P.span;
    sub = piece % P.span;
    static_for<NB>([&](auto jj) {
      constexpr int j = decltype(jj)::value, gl = j / P.bpg, i = j % P.bpg;
      uint32_t v = split_fold(C::sched.bbit[gl * P.bpg + i]);
      static_for<NSETS - 1>([&](auto pp) {
        constexpr int p = decltype(pp)::value + 1;
        constexpr uint32_t f = split_fold(C::sched.bbit[(p * P.gpp + gl) * P.bpg + i]);
        v = set == static_cast<uint32_t>(p) ? f : v;
      });
      fb[j] = v;
    });
    static_for<P.gpp>([&](auto gg) {
      constexpr int gl = decltype(gg)::value;
      uint32_t v = 0;
      if constexpr (P.dbits[gl] != 0) {
        v = split_fold(C::sched.gdyn[gl]);
        static_for<NSETS - 1>([&](auto pp) {
          constexpr int p = decltype(pp)::value + 1;
          constexpr uint32_t f = split_fold(C::sched.gdyn[p * P.gpp + gl]);
          v = set == static_cast<uint32_t>(p) ? f : v;
        });
      }
      fd[gl] = v;
    });
  }

  // Test of the groups (span 1) or of the groups' words (span > 1) in this
  // lane's piece of the key's block.
  __device__ __forceinline__ bool test(uint32_t key, const uint4& piece_data) const {
    using W = typename C::word_t;
    W pw[P.wpp];
    __builtin_memcpy(pw, &piece_data, 16);
    bool ok = true;
    static_for<P.gpp>([&](auto gg) {
      constexpr int gl = decltype(gg)::value;
      W mask = 0;
      static_for<P.bpg>([&](auto ii) {
        constexpr int j = gl * P.bpg + decltype(ii)::value;
        const uint32_t bit = P.bbase[j] + ((key * fb[j]) >> (32 - P.bbits[j]));
        mask |= one_at<W>(bit);
      });
      if constexpr (P.span > 1) {
        // d: the group's word within its set, span * wpp candidates
        const uint32_t d = (key * fd[gl]) >> (32 - P.dbits[gl]);
        const W word = select_word<0, P.wpp>(pw, d % P.wpp);
        ok &= (d / P.wpp != sub) | ((word & mask) == mask);
      } else {
        W word;
        if constexpr (P.dbits[gl] == 0) {
          word = pw[P.lword[gl]];
        } else {
          word = select_word<P.lword[gl], (1 << P.dbits[gl])>(pw, (key * fd[gl]) >> (32 - P.dbits[gl]));
        }
        ok &= (word & mask) == mask;
      }
    });
    return ok;
  }
};

#ifndef DTL_SPLIT_WAVES
#define DTL_SPLIT_WAVES 1  // minimum waves per SIMD the register allocation must allow
#endif

// Stand-in for SplitLane in kernels instantiated for every configuration.
struct NoSplitLane {
  __device__ __forceinline__ explicit NoSplitLane(uint32_t) {}
};

// One tile of kTileKeys keys (keys[tile_base ..], n keys in all) through the
// lane-split test, BT threads: ballot words to bits_tile[0 .. kTileWords),
// returns the wave's match count (uniform per wave).  `sl` is the lane's
// SplitLane for piece = lane % L.  SU: 16-byte loads per lane per batch.
template <class C, int BT, int SU = DTL_SPLIT_UNROLL>
__device__ __forceinline__ uint32_t split_probe_tile(const SplitLane<C>& sl, const typename C::word_t* filter,
                                                     const uint32_t* __restrict__ keys, uint32_t n,
                                                     const AddrArgs& addr, uint64_t tile_base,
                                                     uint64_t* __restrict__ bits_tile) {
  using W = typename C::word_t;
  constexpr SplitSched P = C::split;
  static_assert(P.ok, "configuration has no lane-split form");
  constexpr int L = P.lanes;
  constexpr int KPI = 64 / L;                   // keys per wave load instruction
  constexpr int ROUNDS = kTileKeys / BT;        // 64-key groups per wave and tile
  constexpr int U = SU < ROUNDS * L ? SU : ROUNDS * L;
  constexpr int RU = U / L;                     // 64-key groups per batch
  static_assert(U % L == 0 && ROUNDS % RU == 0, "unroll must cover whole 64-key groups");
  constexpr uint32_t BYTES = C::w * sizeof(W);
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  // the L lanes of a key are adjacent: the L1 merges a block's pieces only
  // within neighbouring lanes (pieces 64 / L lanes apart measured 1.17x slower)
  const uint32_t kl = lane / L;
  // 32-bit byte offsets from the uniform base (filters of this path are
  // under 4 GiB), so a load is one address op + saddr form.
  const char* fbase = reinterpret_cast<const char*>(filter);
  const uint32_t poff = (lane & (L - 1)) * 16u;
  uint32_t cnt = 0;
  // The loop is instantiated twice, for POW2 (block = h >> post) and MAGIC
  // addressing; the branch is uniform over the grid.
  auto run = [&](auto pow2) {
    // every key word of the tile is requested up front (one per round)
    uint32_t kr[ROUNDS];
#pragma unroll
    for (int j = 0; j < ROUNDS; ++j) {
      const uint64_t i = tile_base + static_cast<uint64_t>(j) * BT + tid;
      kr[j] = load_key(keys, i < n ? i : n - 1);  // n >= 1: clamp, no branch
    }
    constexpr int NB = ROUNDS / RU;  // batches of U sub-rounds
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      uint32_t key[U];
      uint4 dat[U];
      // sub-round s covers keys (s % L) * KPI .. + KPI of round s / L
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int sr = b * U + u;
        key[u] = __shfl(kr[sr / L], (sr % L) * KPI + kl);
        const uint32_t h = key[u] * kMul.v[0];
        const uint32_t blk = decltype(pow2)::value ? h >> addr.post : block_of(h, addr);
        dat[u] = *reinterpret_cast<const uint4*>(fbase + (blk * BYTES + poff));
      }
#pragma unroll
      for (int j = 0; j < RU; ++j) {
        const int round = b * RU + j;
        const uint64_t g0 = tile_base + static_cast<uint64_t>(round) * BT + wave * 64u;
        uint64_t word = 0;
#pragma unroll
        for (int p = 0; p < L; ++p)
          word |= and_compress<L>(__ballot(sl.test(key[j * L + p], dat[j * L + p]))) << (p * KPI);
        // keys past n (last tile only): clear their bits, uniformly per wave
        if (g0 + 64 > n) word &= g0 >= n ? 0ull : (~0ull >> (64 - (n - g0)));
        if (lane == 0) store_bits(bits_tile, static_cast<uint32_t>(round) * (BT / 64) + wave, word);
        cnt += __popcll(word);
      }
    }
  };
  if (addr.magic == 0) {
    run(std::true_type{});
  } else {
    run(std::false_type{});
  }
  return cnt;
}

template <class C>
__global__ __launch_bounds__(kGlobalThreads, DTL_SPLIT_WAVES) void k_bbf_probe_split(
    const typename C::word_t* __restrict__ filter, const uint32_t* __restrict__ keys, uint32_t n,
    AddrArgs addr, uint64_t* __restrict__ bits, uint32_t* __restrict__ tile_cnt,
    uint32_t tiles_per_xcd, uint32_t ntiles, const uint32_t* __restrict__ bin_fill, uint32_t tiles_per_bin) {
  constexpr int BT = kGlobalThreads;
  __shared__ uint32_t red[BT / 64];
  const uint32_t tile =
      tiles_per_xcd ? (blockIdx.x & 7u) * tiles_per_xcd + (blockIdx.x >> 3) : blockIdx.x;
  if (tile >= ntiles) return;  // uniform per workgroup, before any barrier
  if (bin_fill) {  // keys grouped in bins of fixed capacity: this tile's bin ends at its fill
    n = bin_end_of(bin_fill, tile, tiles_per_bin);
    if (n <= tile * kTileKeys) {  // nothing of the bin in this tile (uniform, before any barrier)
      clear_tile(bits, tile_cnt, tile);
      return;
    }
  }
  const SplitLane<C> sl(threadIdx.x & (C::split.lanes - 1));
  const uint32_t cnt = split_probe_tile<C, BT>(sl, filter, keys, n, addr, static_cast<uint64_t>(tile) * kTileKeys,
                                               bits + static_cast<uint64_t>(tile) * kTileWords);
  const uint32_t total = block_sum_waves<BT>(cnt, red);
  if (threadIdx.x == 0) tile_cnt[tile] = total;
}

// ---------------------------------------------------------------------------
// LDS layout of a staged filter.  Blocks of 64 or 128 bytes start on 64- or
// 128-byte boundaries, so the j-th 16-byte piece of every block falls on one
// of only 4 (2) groups of banks of the 256-byte bank row, and 64 lanes
// reading piece j of random blocks conflict heavily (measured 4.8 conflict
// cycles per active LDS cycle for u64 w8 s8 k8).  Where the test reads
// whole, statically indexed words, the staged copy rotates the pieces of
// block b by (b >> R) mod P (P pieces per block, R = log2(256 / block
// bytes)), which spreads piece j of different blocks over all bank groups;
// the probe reads the pieces back in logical order, so the test is unchanged.
// ---------------------------------------------------------------------------
template <class C>
constexpr bool lds_swizzled() {
  constexpr uint32_t bytes = C::w * sizeof(typename C::word_t);
  if (bytes != 64 && bytes != 128) return false;
  for (int g = 0; g < C::sched.ngroups; ++g)
    if (C::sched.gdyn[g].nbits != 0) return false;
  return true;
}

// Word i of type W at an address holding 16-byte pieces (the staged LDS
// filter): read through a byte pointer with its alignment stated, so that the
// access is one ds_read of the word's width and no object is read through a
// type it was not written as.
template <class W>
struct WordsAt {
  const char* p;
  __device__ __forceinline__ W operator[](uint32_t i) const {
    const void* a = __builtin_assume_aligned(p + static_cast<size_t>(i) * sizeof(W), sizeof(W));
    W v;
    __builtin_memcpy(&v, a, sizeof(W));
    return v;
  }
};

template <class C>
struct LdsLayout {
  static constexpr bool kSwz = lds_swizzled<C>();
  static constexpr uint32_t P = C::w * sizeof(typename C::word_t) / 16;  // 16-byte pieces per block
  static constexpr uint32_t R = P == 8 ? 1u : 2u;
  // LDS piece index of piece p of block b (swizzled layouts only)
  __device__ __forceinline__ static uint32_t slot(uint32_t b, uint32_t p) {
    return b * P + ((p + (b >> R)) & (P - 1));
  }
  // The filter's bytes are copied as 16-byte pieces, and a 4-byte tail as
  // 4-byte words written through a byte pointer (filters are multiples of 4
  // bytes); readers take words with WordsAt or whole pieces.
  __device__ __forceinline__ static void stage(uint4* __restrict__ lds, const void* filter, uint32_t filter_bytes,
                                               uint32_t tid, uint32_t nthreads) {
    const uint32_t n16 = filter_bytes / 16;
    const uint4* src = static_cast<const uint4*>(filter);
    if constexpr (kSwz) {
      for (uint32_t i = tid; i < n16; i += nthreads) lds[slot(i / P, i % P)] = src[i];
    } else {
      for (uint32_t i = tid; i < n16; i += nthreads) lds[i] = src[i];
      char* dst = reinterpret_cast<char*>(lds);
      const char* srcb = static_cast<const char*>(filter);
      for (uint32_t i = n16 * 4 + tid; i < filter_bytes / 4; i += nthreads) {
        uint32_t v;
        __builtin_memcpy(&v, __builtin_assume_aligned(srcb + 4 * i, 4), 4);
        __builtin_memcpy(__builtin_assume_aligned(dst + 4 * i, 4), &v, 4);
      }
    }
  }
  __device__ __forceinline__ static bool test(uint32_t key, uint32_t b, const uint4* __restrict__ lds) {
    using W = typename C::word_t;
    if constexpr (kSwz) {
      // whole pieces back in logical order, then the register test (swizzled
      // layouts have no hash-chosen word index)
      W blk[C::w];
      constexpr uint32_t WPP = 16 / sizeof(W);
#pragma unroll
      for (uint32_t p = 0; p < P; ++p) {
        const uint4 v = lds[slot(b, p)];
        __builtin_memcpy(&blk[p * WPP], &v, 16);
      }
      return test_regs<C>(key, blk);
    } else {
      const WordsAt<W> words{reinterpret_cast<const char*>(lds) + static_cast<size_t>(b) * C::w * sizeof(W)};
      return test_ptr<C>(key, words);
    }
  }
};

// ---------------------------------------------------------------------------
// Blocked Bloom filter probe, filter staged in LDS (persistent grid).
// Workgroup b probes tiles b, b + grid, b + 2 grid, ... in steps of G tiles
// (G x ROUNDS = 16 keys per lane: 8 tiles with the 1024-thread block that
// large filters leave one of per CU, 2 with 256 threads).  The keys of the
// next step are requested before the current step is tested, so their
// latency overlaps the test even at one workgroup per CU.  Lane j of a wave
// keeps the ballot word of the step's j-th round, and the wave stores its 16
// words with one 16-lane store at the end of the step.  A tile past the end
// is probed like the others (its keys are clamped, its hits masked) and only
// its stores are skipped: the loop stays branch-free.  POW2 addressing
// (block = h >> post) and MAGIC (block = h mod d) are separate
// instantiations.
// ---------------------------------------------------------------------------
template <int BT, int G, int ROUNDS>
__device__ __forceinline__ void lds_load_step(const uint32_t* __restrict__ keys, uint32_t n, uint32_t t0,
                                              uint32_t grid, uint32_t tid, uint32_t (&key)[G][ROUNDS]) {
#pragma unroll
  for (int g = 0; g < G; ++g)
#pragma unroll
    for (int r = 0; r < ROUNDS; ++r) {
      const uint64_t t = static_cast<uint64_t>(t0) + static_cast<uint64_t>(g) * grid;
      const uint64_t i = t * kTileKeys + static_cast<uint64_t>(r) * BT + tid;
      key[g][r] = load_key(keys, i < n ? i : n - 1);  // n >= 1: clamp, no branch
    }
}

template <class C, int BT, bool POW2>
__global__ __launch_bounds__(BT) void k_bbf_probe_lds(const typename C::word_t* __restrict__ filter,
                                                      uint32_t filter_bytes,
                                                      const uint32_t* __restrict__ keys, uint32_t n,
                                                      AddrArgs addr, uint64_t* __restrict__ bits,
                                                      uint32_t* __restrict__ tile_cnt,
                                                      uint32_t ntiles) {
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int G = ROUNDS >= 16 ? 1 : 16 / ROUNDS;  // tiles per step
  constexpr int NW = BT / 64;
  static_assert(G * ROUNDS <= 64, "one ballot word per lane");
  extern __shared__ uint4 lds_filter[];
  __shared__ uint32_t red[G][NW];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t grid = gridDim.x;
  LdsLayout<C>::stage(lds_filter, filter, filter_bytes, tid, BT);
  uint32_t key[G][ROUNDS];
  lds_load_step<BT, G, ROUNDS>(keys, n, blockIdx.x, grid, tid, key);
  __syncthreads();
  for (uint32_t t0 = blockIdx.x; t0 < ntiles; t0 += G * grid) {
    uint32_t next[G][ROUNDS];
    lds_load_step<BT, G, ROUNDS>(keys, n, t0 + G * grid < ntiles ? t0 + G * grid : t0, grid, tid, next);
    uint64_t word = 0;  // lane j: ballot word of round j of the step
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint64_t tile_base = (static_cast<uint64_t>(t0) + static_cast<uint64_t>(g) * grid) * kTileKeys;
      uint32_t cnt = 0;
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) {
        const uint64_t i = tile_base + static_cast<uint64_t>(r) * BT + tid;
        const uint32_t h0 = key[g][r] * kMul.v[0];
        const uint32_t b = POW2 ? h0 >> addr.post : block_of(h0, addr);
        const bool hit = LdsLayout<C>::test(key[g][r], b, lds_filter) & (i < n);
        const uint64_t m = __ballot(hit);
        word = lane == static_cast<uint32_t>(g * ROUNDS + r) ? m : word;
        cnt += __popcll(m);
      }
      if (lane == 0) red[g][wave] = cnt;
    }
    {
      const uint32_t g = lane / ROUNDS, r = lane % ROUNDS;
      const uint64_t t = static_cast<uint64_t>(t0) + static_cast<uint64_t>(g) * grid;
      if (lane < static_cast<uint32_t>(G * ROUNDS) && t < ntiles)
        store_bits(bits, (t * kTileKeys + static_cast<uint64_t>(r) * BT) / 64 + wave, word);
    }
    __syncthreads();
    if (tid < G) {
      const uint64_t t = static_cast<uint64_t>(t0) + static_cast<uint64_t>(tid) * grid;
      if (t < ntiles) {
        uint32_t s = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) s += red[tid][w];
        tile_cnt[t] = s;
      }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int r = 0; r < ROUNDS; ++r) key[g][r] = next[g][r];
  }
}

// ---------------------------------------------------------------------------
// Blocked Bloom filter build: one key per thread, one atomicOr per group.
// ---------------------------------------------------------------------------
template <class W>
__device__ __forceinline__ void atomic_or_word(W* p, W v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class C>
__global__ __launch_bounds__(256) void k_bbf_build(typename C::word_t* __restrict__ filter,
                                                   const uint32_t* __restrict__ keys, uint64_t n,
                                                   AddrArgs addr) {
  using W = typename C::word_t;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint32_t key = load_key(keys, i);
    const uint32_t b = block_of(key * kMul.v[0], addr);
    W* blk = filter + static_cast<uint64_t>(b) * C::w;
    const HashesOf<C> H(key);
    W mask[C::sched.ngroups];
    key_masks<C>(H, mask);
    static_for<C::sched.ngroups>([&](auto g) {
      constexpr Chunk d = C::sched.gdyn[g];
      uint32_t wi = C::sched.gword[g];
      if constexpr (d.nbits != 0) wi += H.chunk(d);
      atomic_or_word<W>(blk + wi, mask[g]);
    });
  }
}

}  // namespace dtlh
