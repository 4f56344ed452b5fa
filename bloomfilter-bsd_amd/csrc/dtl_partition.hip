// dtl_partition.hip -- partitioned probing of blocked Bloom filters that do
// not fit in LDS or L2.
//
// A probe of a large filter reads one random block per key; on gfx950 every
// such read fills a whole 128-byte L2 line from HBM, so the direct kernel is
// bound by random-line throughput.  The partitioned probe first groups the
// keys by filter slice (slice = 2^slice_shift consecutive blocks, 512 KiB to
// 2 MiB by default), then probes slice after slice, so that a slice is read from HBM
// into one XCD's L2 once and every key of the partition hits it there.
// Only key streams touch HBM: keys are read twice (histogram, scatter), the
// (key, position) pairs written and read once.
//
//   1. k_fill_zero_u32  clears the partition counts;
//   2. k_part_hist      LDS histogram of partition ids, added to global counts;
//   3. k_part_scan      exclusive scan -> partition offsets (and cursors);
//   4. k_part_scatter   per tile of 16 keys per thread: counting sort in LDS, one global
//                       atomicAdd per non-empty bin reserves the tile's run in
//                       each partition, runs written out coalesced;
//   5. k_bbf_probe_split / k_bbf_probe_global (per configuration,
//                       dtl_kernels.hpp) probes the grouped keys in XCD order
//                       and writes the match bitmap in grouped order (ballot
//                       words, no atomics);
//   6. reordering of the matches (only matches move, about 1-2% of keys):
//      scan + k_expand_gather list the original positions of the matches,
//      k_hit_count / k_hit_colscan / k_part_scan / k_hit_place bucket them by
//      position (2^18 positions per bucket or more), and k_hit_bits builds
//      each bucket's piece of the key-order bitmap in LDS and writes either
//      the bucket's positions in ascending order (blocked Bloom filters: a key
//      matches once) or the bitmap piece with its tile counts (cuckoo).
// The order of keys inside a partition depends on atomic timing, but the
// key-order bitmap does not, so the match list is the same ascending list as
// the direct probe's.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dtl_cuckoo.hpp"
#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"
#include "dtl_scan.hpp"

namespace dtlh {

constexpr int kHistThreads = 256;
constexpr int kHistUnroll = 8;
constexpr int kHitUnroll = 8;  // positions in flight per thread in the hit sort
// Keys per thread of the scatter (tile = threads x keys per thread): 16, and
// 32 for blocked Bloom filters of 512 to 1024 partitions (partition_keys).
// The histogram form always uses 16 (its lanes follow the scatter's tiles).
constexpr uint32_t kHistKptLog2 = 4;

constexpr int kPartScanThreads = 1024;

// Partition function: partition = bucket >> slice_shift, where bucket is the
// key's block (blocked Bloom) or first bucket (cuckoo: the same block_of of
// h0), or (kind 1) the key's alternate cuckoo bucket.
struct PartFn {
  AddrArgs a;
  uint32_t slice_shift;
  uint32_t tag_mask;
  CfArgs cf;
};

// KIND is a template parameter of the kernels (the branch stays out of their
// unrolled per-key loops; a runtime branch there cost the scatter ~10%).
template <int KIND>
__device__ __forceinline__ uint32_t part_of(uint32_t key, const PartFn& f) {
  uint32_t b = block_of(key * kMul.v[0], f.a);
  if constexpr (KIND == 1) b = cf_alt(b, cf_tag(key, f.tag_mask), f.cf);
  return b >> f.slice_shift;
}

static PartFn make_part_fn(const BbfPartLaunch& L) {
  PartFn f;
  f.a = AddrArgs{L.addr_magic, L.addr_mshift, L.addr_divisor, L.addr_post};
  f.slice_shift = L.slice_shift;
  f.tag_mask = L.tag_mask;
  f.cf.magic = L.addr_magic;
  f.cf.mshift = L.addr_mshift;
  f.cf.divisor = L.addr_divisor;
  f.cf.post = L.addr_post;
  f.cf.mode = L.cf_mode;
  f.cf.bucket_mask = L.cf_bucket_mask;
  f.cf.bucket_cnt = L.cf_bucket_cnt;
  f.cf.pad = 0;
  f.cf.victim_offset = L.cf_victim_offset;
  return f;
}

__global__ void k_fill_zero_u32(uint32_t* __restrict__ p, uint64_t count) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    p[i] = 0u;
}

// Bins are (partition, lane) pairs, bin = partition * kLanes + lane, where
// lane = (scatter tile index) % kLanes: each partition's range is split into
// kLanes sub-ranges, one per XCD the scatter tiles of that lane run on
// (round-robin dispatch), so the runs one XCD writes are adjacent and merge
// into whole lines in its L2.  The split only affects speed: the hist and
// scatter kernels derive the lane from the tile index the same way.
constexpr uint32_t kLanes = kPartLanes;

__device__ __forceinline__ uint32_t lane_of(uint64_t i, uint32_t tile_log2) {
  return static_cast<uint32_t>(i >> tile_log2) & (kLanes - 1u);
}

// Each workgroup counts keys [blockIdx * per_wg, +per_wg) (per_wg is a
// multiple of kHistThreads * kHistUnroll) into nparts * kLanes bins.  In LDS
// the bins are lane-major (lane * nparts + partition): the keys of one wave
// share a lane, so their atomics spread over all banks.
template <int KIND>
__global__ __launch_bounds__(kHistThreads) void k_part_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                            PartFn pf,
                                                            uint32_t nparts, uint32_t per_wg,
                                                            uint32_t tile_log2,
                                                            uint32_t* __restrict__ counts) {
  extern __shared__ uint32_t hist[];
  const uint32_t tid = threadIdx.x;
  const uint32_t nbins = nparts * kLanes;
  for (uint32_t i = tid; i < nbins; i += kHistThreads) hist[i] = 0u;
  __syncthreads();
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * per_wg;
  const uint64_t end = std::min<uint64_t>(begin + per_wg, n);
  for (uint64_t base = begin; base < end; base += kHistThreads * kHistUnroll) {
    uint32_t key[kHistUnroll];
    bool valid[kHistUnroll];
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const uint64_t i = base + static_cast<uint64_t>(u) * kHistThreads + tid;
      valid[u] = i < end;
      key[u] = load_key(keys, valid[u] ? i : begin);
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const uint64_t i = base + static_cast<uint64_t>(u) * kHistThreads + tid;
      if (valid[u]) atomicAdd(&hist[lane_of(i, tile_log2) * nparts + part_of<KIND>(key[u], pf)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nbins; i += kHistThreads) {
    const uint32_t c = hist[(i % kLanes) * nparts + i / kLanes];  // global bin i = partition * kLanes + lane
    if (c) atomicAdd(&counts[i], c);
  }
}

// Both partition functions of the cuckoo probe in one read of the keys: the
// slice of the first bucket into counts0, of the alternate bucket into
// counts1 (LDS: the two lane-major bin arrays one after the other).
__global__ __launch_bounds__(kHistThreads) void k_part_hist2(const uint32_t* __restrict__ keys, uint32_t n,
                                                             PartFn pf, uint32_t nparts, uint32_t per_wg,
                                                             uint32_t tile_log2, uint32_t* __restrict__ counts0,
                                                             uint32_t* __restrict__ counts1) {
  extern __shared__ uint32_t hist[];
  const uint32_t tid = threadIdx.x;
  const uint32_t nbins = nparts * kLanes;
  for (uint32_t i = tid; i < 2 * nbins; i += kHistThreads) hist[i] = 0u;
  __syncthreads();
  const uint64_t begin = static_cast<uint64_t>(blockIdx.x) * per_wg;
  const uint64_t end = std::min<uint64_t>(begin + per_wg, n);
  for (uint64_t base = begin; base < end; base += kHistThreads * kHistUnroll) {
    uint32_t key[kHistUnroll];
    bool valid[kHistUnroll];
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const uint64_t i = base + static_cast<uint64_t>(u) * kHistThreads + tid;
      valid[u] = i < end;
      key[u] = load_key(keys, valid[u] ? i : begin);
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const uint64_t i = base + static_cast<uint64_t>(u) * kHistThreads + tid;
      if (valid[u]) {
        const uint32_t row = lane_of(i, tile_log2) * nparts;
        atomicAdd(&hist[row + part_of<0>(key[u], pf)], 1u);
        atomicAdd(&hist[nbins + row + part_of<1>(key[u], pf)], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = tid; i < nbins; i += kHistThreads) {
    const uint32_t j = (i % kLanes) * nparts + i / kLanes;  // global bin i = partition * kLanes + lane
    const uint32_t c0 = hist[j], c1 = hist[nbins + j];
    if (c0) atomicAdd(&counts0[i], c0);
    if (c1) atomicAdd(&counts1[i], c1);
  }
}

// One workgroup: offsets[p] = sum of counts[< p], offsets[nparts] = total;
// cursor[p] = offsets[p].
__global__ __launch_bounds__(kPartScanThreads) void k_part_scan(const uint32_t* __restrict__ counts,
                                                                uint32_t nparts,
                                                                uint32_t* __restrict__ offsets,
                                                                uint32_t* __restrict__ cursor) {
  __shared__ uint32_t sh[kPartScanThreads / 64];
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nparts; c0 += kPartScanThreads) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < nparts ? counts[i] : 0u;
    uint32_t tot;
    const uint32_t ex = carry + block_excl_scan<kPartScanThreads>(v, tot, sh);
    if (i < nparts) {
      offsets[i] = ex;
      cursor[i] = ex;
    }
    carry += tot;
  }
  if (threadIdx.x == 0) offsets[nparts] = carry;
}

// LDS layout (dynamic): hist | lbase | gbase (nparts u32 each) | buf (tile
// u32).  The runs are written with ordinary (cached) stores: each tile's run
// in a partition is ~64 keys and shares its first and last 128-byte lines
// with the neighbouring runs of the same lane, which L2 merges; non-temporal
// stores of the same runs made the scatter 1.7x slower on MI355X.  The tile goes out in two passes through the same buffer, keys first,
// then positions, so a workgroup needs ~4 bytes of LDS per key and two
// workgroups fit on a CU: one loads and sorts while the other writes (the
// launch bound asks for 2 workgroups' waves per SIMD, <= 64 VGPRs at 1024
// threads).
// Capacity bins: the keys of a tile that do not fit in their (full) bins --
// none for uniform keys, many when a batch repeats a few keys -- are spilled
// to other bins.  The workgroup reserves all of them at once, in up to
// kSpillSegs chunks taken from bins walked from a start chosen by the
// workgroup (one atomicAdd per chunk; a chunk may fit only partly, and only
// the slots below the bin's capacity are used, all of them written); keys
// beyond those chunks take one slot at a time (spill_slot).  The bins'
// capacities add up to more than n, so free slots exist; the bounds on the
// walks only keep the loops finite.  Where a key sits decides locality only:
// the probe computes every key's block from the key.
constexpr uint32_t kSpillSegs = 32;
constexpr uint32_t kSpillMark = 0x80000000u;  // dst of a spilled key: mark | its index among the tile's spills

// One slot for a spilled key beyond the workgroup's chunks.
__device__ __forceinline__ uint32_t spill_slot(uint32_t* __restrict__ cursor, uint32_t nbins, uint32_t cap,
                                               uint32_t key, uint32_t pos) {
  uint32_t b = ((key ^ (pos * 0x9E3779B9u)) * 0x85EBCA6Bu) % nbins;
  for (uint32_t t = 0; t < 2u * nbins; ++t) {
    const uint32_t s = atomicAdd(&cursor[b], 1u);
    if (s < (b + 1u) * cap) return s;
    b = b + 1u == nbins ? 0u : b + 1u;
  }
  return 0xFFFFFFFFu;
}

// Keys a scatter leaves out (SKIP forms only): those whose bit is set in a
// key-order bitmap, while a device count exceeds a threshold (the cuckoo
// probe's pass 2 drops the keys pass 1 found, once enough of them did).
struct KeySkip {
  const uint64_t* bits;
  const uint64_t* gate;
  uint64_t thr;
};

template <int kScatterThreads, int KPT, int KIND, bool CAP, bool SKIP = false>
__global__ __launch_bounds__(kScatterThreads, kScatterThreads / (8 * KPT)) void k_part_scatter(const uint32_t* __restrict__ keys,
                                                                  uint32_t n, PartFn pf, uint32_t nparts,
                                                                  uint32_t* __restrict__ cursor,
                                                                  uint32_t* __restrict__ keys_p,
                                                                  uint32_t* __restrict__ pos_p,
                                                                  uint32_t bin_cap, KeySkip sk) {
  // SKIP: a left-out key's slot is kDropSlot, and its LDS slot in ps has bit
  // 15 set (slots stay below 2^15: tiles of at most 32K keys)
  constexpr uint32_t kDropSlot = 0xFFFFFFFFu, kDropLds = 0x8000u;
  static_assert(kScatterThreads * KPT <= 32768, "LDS slots below 2^15");
  constexpr uint32_t kScatterTile = kScatterThreads * KPT;
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t red[kScatterThreads / 64];
  uint32_t* hist = sm;
  uint32_t* lbase = sm + nparts;
  uint32_t* gbase = sm + 2 * nparts;
  uint32_t* fit = sm + 3 * nparts;  // CAP: keys of the tile's run that fit in the bin
  __shared__ uint32_t s_nspill, s_nseg, s_placed;
  __shared__ uint32_t s_seg[kSpillSegs][3];  // CAP: (first spill index, grouped slot, count)
  uint32_t* buf = sm + (CAP ? 4 : 3) * nparts;
  const uint32_t tid = threadIdx.x;
  const uint64_t tile_base = static_cast<uint64_t>(blockIdx.x) * kScatterTile;
  const uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(kScatterTile, n - tile_base));
  for (uint32_t i = tid; i < nparts; i += kScatterThreads) hist[i] = 0u;
  uint32_t key[KPT], slot[KPT];
  bool skip_on = false;
  uint32_t dropm = 0;  // SKIP: bit r set = leave out this thread's key r
  if constexpr (SKIP) {
    // the tile's skip bits (cnt / 8 bytes) go to the front of buf, which
    // holds nothing else until the keys are placed: one load per thread,
    // issued before the keys, instead of one dependent load per key
    skip_on = *sk.gate > sk.thr;  // uniform
    const uint32_t nw = (cnt + 31u) / 32u;
    uint32_t sw = 0;
    if (skip_on && tid < nw)
      __builtin_memcpy(&sw, reinterpret_cast<const char*>(sk.bits) + (tile_base / 32u + tid) * 4u, 4);
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
      const uint32_t j = r * kScatterThreads + tid;
      key[r] = load_key(keys, tile_base + (j < cnt ? j : cnt - 1));
    }
    if (skip_on && tid < nw) buf[tid] = sw;
    __syncthreads();
    if (skip_on) {
#pragma unroll
      for (int r = 0; r < KPT; ++r) {
        const uint32_t j = r * kScatterThreads + tid;  // words past the tile's end: read, never used
        dropm |= ((buf[j >> 5] >> (j & 31u)) & 1u) << r;
      }
    }
  } else {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
      const uint32_t j = r * kScatterThreads + tid;
      key[r] = load_key(keys, tile_base + (j < cnt ? j : cnt - 1));
    }
  }
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t j = r * kScatterThreads + tid;
    const uint32_t pid = part_of<KIND>(key[r], pf);
    bool drop = false;
    if constexpr (SKIP) drop = j < cnt && ((dropm >> r) & 1u) != 0u;
    // rank in the partition (low 16 bits, tile < 2^16) and partition id
    slot[r] = drop ? kDropSlot : j < cnt ? (atomicAdd(&hist[pid], 1u) | (pid << 16)) : 0u;
  }
  __syncthreads();
  // exclusive scan of the histogram (each thread owns `per` consecutive bins)
  // and reservation of this tile's run in every non-empty partition
  const uint32_t per = (nparts + kScatterThreads - 1) / kScatterThreads;
  const uint32_t b0 = tid * per, b1 = std::min(nparts, b0 + per);
  uint32_t local = 0;
  for (uint32_t b = b0; b < b1; ++b) local += hist[b];
  uint32_t tot;
  uint32_t run = block_excl_scan<kScatterThreads>(local, tot, red);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t c = hist[b];
    lbase[b] = run;
    if (c) {
      const uint32_t bin = b * kLanes + (blockIdx.x & (kLanes - 1u));
      const uint32_t g = atomicAdd(&cursor[bin], c);
      gbase[b] = g;
      if constexpr (CAP) {
        const uint32_t end = (bin + 1u) * bin_cap;
        fit[b] = g >= end ? 0u : min(c, end - g);
      }
    }
    run += c;
  }
  __syncthreads();
  // keys in the tile's LDS buffer: all cnt, or those a SKIP form keeps
  const uint32_t m = SKIP ? tot : cnt;
  uint32_t ps[KPT / 2];  // LDS slots (< 2^16), two per register
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t j = r * kScatterThreads + tid;
    const bool drop = SKIP && slot[r] == kDropSlot;
    const uint32_t sv = drop ? kDropLds : lbase[slot[r] >> 16] + (slot[r] & 0xFFFFu);
    if (j < cnt && !drop) buf[sv] = key[r];
    ps[r / 2] = (r & 1) ? (ps[r / 2] | (sv << 16)) : sv;
  }
  __syncthreads();
  uint32_t dst[KPT];
  if constexpr (CAP) {
    if (tid == 0) s_nspill = 0u;
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t sl = r * kScatterThreads + tid;
    if (sl < m) {
      const uint32_t k = buf[sl];
      const uint32_t p = part_of<KIND>(k, pf);
      const uint32_t o = sl - lbase[p];
      dst[r] = gbase[p] + o;
      if constexpr (CAP) {
        if (o >= fit[p]) dst[r] = kSpillMark | atomicAdd(&s_nspill, 1u);
      }
      if (!CAP || !(dst[r] & kSpillMark)) keys_p[dst[r]] = k;
    }
  }
  if constexpr (CAP) {
    __syncthreads();
    const uint32_t nspill = s_nspill;  // uniform
    if (nspill) {
      if (tid == 0) {
        const uint32_t nbins = nparts * kLanes;
        uint32_t b = (blockIdx.x * 0x9E3779B1u) % nbins, done = 0, nseg = 0;
        for (uint32_t t = 0; t < 2u * nbins && done < nspill && nseg < kSpillSegs; ++t) {
          const uint32_t want = nspill - done;
          const uint32_t g = atomicAdd(&cursor[b], want);
          const uint32_t end = (b + 1u) * bin_cap;
          const uint32_t f = g >= end ? 0u : min(want, end - g);
          if (f) {
            s_seg[nseg][0] = done;
            s_seg[nseg][1] = g;
            s_seg[nseg][2] = f;
            ++nseg;
            done += f;
          }
          b = b + 1u == nbins ? 0u : b + 1u;
        }
        s_nseg = nseg;
        s_placed = done;
      }
      __syncthreads();
      const uint32_t nseg = s_nseg, placed = s_placed;
#pragma unroll
      for (int r = 0; r < KPT; ++r) {
        const uint32_t sl = r * kScatterThreads + tid;
        if (sl < m && (dst[r] & kSpillMark)) {
          const uint32_t idx = dst[r] & ~kSpillMark;
          const uint32_t k = buf[sl];
          uint32_t d = 0xFFFFFFFFu;
          if (idx < placed) {
            for (uint32_t q = 0; q < nseg; ++q)
              if (idx >= s_seg[q][0] && idx < s_seg[q][0] + s_seg[q][2]) d = s_seg[q][1] + (idx - s_seg[q][0]);
          } else {
            d = spill_slot(cursor, nparts * kLanes, bin_cap, k, static_cast<uint32_t>(tile_base) + sl);
          }
          dst[r] = d;
          if (d != 0xFFFFFFFFu) keys_p[d] = k;
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t j = r * kScatterThreads + tid;
    const uint32_t sv = (ps[r / 2] >> (16 * (r & 1))) & 0xFFFFu;
    if (j < cnt && (!SKIP || !(sv & kDropLds))) buf[sv] = static_cast<uint32_t>(tile_base) + j;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t sl = r * kScatterThreads + tid;
    if (sl < m && (!CAP || dst[r] != 0xFFFFFFFFu)) pos_p[dst[r]] = buf[sl];
  }
}

// Capacity bins: cursor[b] = b * cap, the start of bin b.
__global__ void k_part_init_bins(uint32_t* __restrict__ cursor, uint32_t nbins, uint32_t cap) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbins) cursor[b] = b * cap;
}

// Hit sort: the matches' original positions hp[0 .. *d_hits) (any order)
// are bucketed by position without global atomics.  Workgroup g of G owns
// the slice [h * g / G, h * (g + 1) / G) of hp:
//   k_hit_count    LDS histogram of its slice -> hcounts[g][bucket];
//   k_hit_colscan  per bucket, exclusive scan over g in place, and the
//                  bucket totals (then k_part_scan -> bucket offsets);
//   k_hit_place    writes its positions to bucket offset + hcounts[g][b] +
//                  LDS rank (order inside a bucket is irrelevant: k_hit_bits
//                  turns each bucket into a bitmap).
// The part of hp a hit-sort stage works on (HitSortStage): hp[first ..
// *d_hits), first = *gate for the second stage of an open gate, else 0.
struct HitSel {
  const uint64_t* d_hits;
  const uint64_t* gate;  // null for kHitSortAll
  uint64_t thr;
  uint32_t stage;
};

__device__ __forceinline__ bool hit_gate_open(const HitSel& s) { return s.gate != nullptr && *s.gate > s.thr; }

// false when the stage has nothing to do (kHitSortEarly, gate closed); the
// answer is the same for every thread of the grid
__device__ __forceinline__ bool hit_stage_runs(const HitSel& s) {
  return s.stage != kHitSortEarly || hit_gate_open(s);
}

// this workgroup's share [b0, b1) of the stage's range of hp
__device__ __forceinline__ void hit_slice(const HitSel& s, uint64_t* b0, uint64_t* b1) {
  const uint64_t first = (s.stage == kHitSortRest && hit_gate_open(s)) ? *s.gate : 0;
  const uint64_t h = *s.d_hits - first;
  *b0 = first + h * blockIdx.x / gridDim.x;
  *b1 = first + h * (blockIdx.x + 1) / gridDim.x;
}

__global__ __launch_bounds__(256) void k_hit_count(const uint32_t* __restrict__ hp, HitSel sel, uint32_t shift,
                                                   uint32_t nbuckets, uint32_t* __restrict__ hcounts) {
  extern __shared__ uint32_t hist[];
  if (!hit_stage_runs(sel)) return;  // uniform, before any barrier
  const uint32_t tid = threadIdx.x;
  for (uint32_t b = tid; b < nbuckets; b += 256) hist[b] = 0u;
  __syncthreads();
  uint64_t b0, b1;
  hit_slice(sel, &b0, &b1);
  // kHitUnroll positions in flight per thread, then their LDS atomics
  uint64_t i = b0 + tid;
  for (; i + (kHitUnroll - 1) * 256 < b1; i += kHitUnroll * 256) {
    uint32_t p[kHitUnroll];
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) p[u] = hp[i + u * 256];
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) atomicAdd(&hist[p[u] >> shift], 1u);
  }
  for (; i < b1; i += 256) atomicAdd(&hist[hp[i] >> shift], 1u);
  __syncthreads();
  uint32_t* row = hcounts + static_cast<uint64_t>(blockIdx.x) * nbuckets;
  for (uint32_t b = tid; b < nbuckets; b += 256) row[b] = hist[b];
}

// Column scan of hcounts: a workgroup takes kColB buckets, each split into
// kColS segments of consecutive workgroup rows (one thread per (bucket,
// segment)), so every thread has its segment's loads in flight at once.
// 8 x 32 rather than 16 x 16: twice the workgroups (128 for 1024 buckets),
// half the rows per thread; 5-16 us less per 2^28-key call, same box.
constexpr uint32_t kColB = 8, kColS = 32;
static_assert(kColB * kColS == 256, "one thread per (bucket, segment)");

__global__ __launch_bounds__(256) void k_hit_colscan(uint32_t* __restrict__ hcounts, uint32_t groups,
                                                     uint32_t nbuckets, uint32_t* __restrict__ totals, HitSel sel) {
  __shared__ uint32_t seg[kColS][kColB + 1];
  if (!hit_stage_runs(sel)) return;
  const uint32_t bl = threadIdx.x % kColB, sg = threadIdx.x / kColB;
  const uint32_t b = blockIdx.x * kColB + bl;
  const uint32_t per = (groups + kColS - 1) / kColS;
  const uint32_t g0 = min(groups, sg * per), g1 = min(groups, g0 + per);
  uint32_t sum = 0;
  if (b < nbuckets) {
    for (uint32_t g = g0; g < g1; ++g) sum += hcounts[static_cast<uint64_t>(g) * nbuckets + b];
  }
  seg[sg][bl] = sum;
  __syncthreads();
  if (sg == 0) {
    uint32_t run = 0;
    for (uint32_t k = 0; k < kColS; ++k) {
      const uint32_t v = seg[k][bl];
      seg[k][bl] = run;
      run += v;
    }
    if (b < nbuckets) totals[b] = run;
  }
  __syncthreads();
  uint32_t run = seg[sg][bl];
  if (b < nbuckets) {
    for (uint32_t g = g0; g < g1; ++g) {
      uint32_t* c = hcounts + static_cast<uint64_t>(g) * nbuckets + b;
      const uint32_t v = *c;
      *c = run;
      run += v;
    }
  }
}

__global__ __launch_bounds__(256) void k_hit_place(const uint32_t* __restrict__ hp, HitSel sel, uint32_t shift,
                                                   uint32_t nbuckets, const uint32_t* __restrict__ hcounts,
                                                   const uint32_t* __restrict__ boffsets,
                                                   uint32_t* __restrict__ out) {
  extern __shared__ uint32_t sm[];
  if (!hit_stage_runs(sel)) return;
  uint32_t* rank = sm;
  uint32_t* base = sm + nbuckets;
  const uint32_t tid = threadIdx.x;
  const uint32_t* row = hcounts + static_cast<uint64_t>(blockIdx.x) * nbuckets;
  for (uint32_t b = tid; b < nbuckets; b += 256) {
    rank[b] = 0u;
    base[b] = boffsets[b] + row[b];
  }
  __syncthreads();
  uint64_t b0, b1;
  hit_slice(sel, &b0, &b1);
  uint64_t i = b0 + tid;
  for (; i + (kHitUnroll - 1) * 256 < b1; i += kHitUnroll * 256) {
    uint32_t p[kHitUnroll];
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) p[u] = hp[i + u * 256];
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) {
      const uint32_t b = p[u] >> shift;
      out[base[b] + atomicAdd(&rank[b], 1u)] = p[u];
    }
  }
  for (; i < b1; i += 256) {
    const uint32_t p = hp[i];
    const uint32_t b = p >> shift;
    out[base[b] + atomicAdd(&rank[b], 1u)] = p;
  }
}

// One workgroup per bucket of 2^shift positions: its matches set bits of an
// LDS bitmap (dynamic, 2^shift bits), which is written out whole (zeros
// included) together with the match count of each of its tiles.  The second
// stage of an open gate starts from the bitmap the first stage wrote.  With
// out_pos (no position can occur twice, so boffsets[b] is the number of
// matches before the bucket), the bucket's positions are written in
// ascending order at out_pos + boffsets[b] instead, and workgroup 0 writes
// the total to *d_count_out.
__global__ __launch_bounds__(256) void k_hit_bits(const uint32_t* __restrict__ sorted,
                                                  const uint32_t* __restrict__ boffsets, uint32_t shift,
                                                  uint32_t ntiles, uint64_t* __restrict__ bits,
                                                  uint32_t* __restrict__ tile_cnt, HitSel sel,
                                                  uint32_t* __restrict__ out_pos, uint32_t match_offset,
                                                  uint64_t* __restrict__ d_count_out) {
  extern __shared__ uint32_t lbits[];
  __shared__ uint32_t ctot[kHitEmitTilesMax];
  if (!hit_stage_runs(sel)) return;
  const bool acc = sel.stage == kHitSortRest && hit_gate_open(sel);
  const uint32_t words32 = (1u << shift) / 32u;
  const uint32_t btiles = (1u << shift) / kTileKeys;
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  const uint32_t tile0 = b * btiles;
  const uint32_t tiles = std::min(btiles, ntiles - tile0);
  if (acc) {
    for (uint32_t w = tid; w < words32 / 2u; w += 256) {
      const uint64_t v = w < tiles * kTileWords ? bits[static_cast<uint64_t>(tile0) * kTileWords + w] : 0ull;
      lbits[2 * w] = static_cast<uint32_t>(v);
      lbits[2 * w + 1] = static_cast<uint32_t>(v >> 32);
    }
  } else {
    for (uint32_t i = tid; i < words32; i += 256) lbits[i] = 0u;
  }
  __syncthreads();
  const uint32_t beg = boffsets[b], end = boffsets[b + 1];
  const uint32_t lmask = (1u << shift) - 1u;
  uint32_t i = beg + tid;
  for (; i + (kHitUnroll - 1) * 256 < end; i += kHitUnroll * 256) {
    uint32_t p[kHitUnroll];
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) p[u] = sorted[i + u * 256] & lmask;
#pragma unroll
    for (int u = 0; u < kHitUnroll; ++u) atomicOr(&lbits[p[u] >> 5], 1u << (p[u] & 31u));
  }
  for (; i < end; i += 256) {
    const uint32_t p = sorted[i] & lmask;
    atomicOr(&lbits[p >> 5], 1u << (p & 31u));
  }
  __syncthreads();
  static_assert(2 * kTileWords == 64, "one LDS word of a tile per lane");
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  if (out_pos) {  // uniform
    // tile c of the bucket: lane l holds its LDS word l; tile totals, one
    // exclusive scan of them (wave 0, 4 tiles per lane), then every wave
    // writes its tiles' positions at the tile's base + the lane's rank
    for (uint32_t c = wave; c < tiles; c += 256 / 64) {
      const uint32_t v = wave_sum(__popc(lbits[c * 64u + lane]));
      if (lane == 0) ctot[c] = v;
    }
    __syncthreads();
    if (wave == 0) {
      static_assert(kHitEmitTilesMax == 4 * 64, "4 tile totals per lane");
      uint32_t a[4], sum = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t c = lane * 4u + k;
        a[k] = c < tiles ? ctot[c] : 0u;
        sum += a[k];
      }
      uint32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t c = lane * 4u + k;
        if (c < tiles) ctot[c] = run;
        run += a[k];
      }
    }
    __syncthreads();
    const uint32_t base = boffsets[b];
    for (uint32_t c = wave; c < tiles; c += 256 / 64) {
      uint32_t w = lbits[c * 64u + lane];
      const uint32_t pc = __popc(w);
      uint32_t o = base + ctot[c] + wave_incl_scan(pc) - pc;
      const uint32_t p0 = (tile0 + c) * kTileKeys + lane * 32u + match_offset;
      while (w) {
        const uint32_t j = __ffs(w) - 1u;
        out_pos[o++] = p0 + j;
        w &= w - 1u;
      }
    }
    if (b == 0 && tid == 0) *d_count_out = *sel.d_hits;
    return;
  }
  for (uint32_t w = tid; w < tiles * kTileWords; w += 256) {
    const uint64_t v = static_cast<uint64_t>(lbits[2 * w]) | (static_cast<uint64_t>(lbits[2 * w + 1]) << 32);
    bits[static_cast<uint64_t>(tile0) * kTileWords + w] = v;
  }
  // a tile is 2 kTileWords = 64 LDS words: one per lane of a wave (a thread
  // summing its own tile's words would have every lane of a wave on one bank)
  for (uint32_t t = wave; t < tiles; t += 256 / 64) {
    const uint32_t c = wave_sum(__popc(lbits[t * 64u + lane]));
    if (lane == 0) tile_cnt[tile0 + t] = c;
  }
}

static uint32_t fill_grid(uint64_t count) {
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>((count + 255) / 256, 8192)));
}

static uint32_t part_scatter_lds_bytes(uint32_t nparts, uint32_t tile, bool cap) {
  return nparts * (cap ? 16u : 12u) + tile * 4u;
}

template <int BT, int KPT, int KIND, bool CAP, bool SKIP = false>
static hipError_t launch_scatter(const BbfPartLaunch& L, const PartFn& pf) {
  constexpr uint32_t tile = BT * KPT;
  const uint32_t lds = part_scatter_lds_bytes(L.nparts, tile, CAP);
  auto fn = k_part_scatter<BT, KPT, KIND, CAP, SKIP>;
  hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(fn), lds);
  if (e != hipSuccess) return e;
  KeySkip sk;
  sk.bits = L.skip_bits;
  sk.gate = L.gate_count;
  sk.thr = L.gate_thr;
  const uint32_t grid = static_cast<uint32_t>((static_cast<uint64_t>(L.n) + tile - 1) / tile);
  hipLaunchKernelGGL(fn, dim3(grid), dim3(BT), lds, L.stream, L.keys, L.n, pf, L.nparts, L.cursor, L.keys_p,
                     L.pos_p, L.bin_cap, sk);
  return hipGetLastError();
}

// Launch shape of the histogram (about 8 workgroups per CU, each a multiple
// of kHistThreads * kHistUnroll keys) and of the scatter.
struct PartShape {
  uint32_t per_wg, hist_grid, threads, tile_log2;
};

static PartShape part_shape(const BbfPartLaunch& L) {
  PartShape S;
  const uint64_t unit = kHistThreads * kHistUnroll;
  const uint64_t want_wg = static_cast<uint64_t>(std::max(L.cu_count, 1)) * 8;
  uint64_t per_wg = (L.n + want_wg - 1) / want_wg;
  per_wg = std::max<uint64_t>(unit, (per_wg + unit - 1) / unit * unit);
  S.per_wg = static_cast<uint32_t>(per_wg);
  S.hist_grid = static_cast<uint32_t>((L.n + per_wg - 1) / per_wg);
  S.threads = L.scatter_threads >= 1024 ? 1024u : L.scatter_threads >= 512 ? 512u : 256u;
  S.tile_log2 = (S.threads == 1024 ? 10u : S.threads == 512 ? 9u : 8u) + kHistKptLog2;
  return S;
}

// Scan of L.counts (already histogrammed) and the scatter.
template <int KIND>
static hipError_t scan_and_scatter(const BbfPartLaunch& L, const PartFn& pf, const PartShape& S) {
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(kPartScanThreads), 0, L.stream, L.counts, L.nparts * kLanes,
                     L.offsets, L.cursor);
  switch (S.threads) {
    case 1024: return launch_scatter<1024, 16, KIND, false>(L, pf);
    case 512: return launch_scatter<512, 16, KIND, false>(L, pf);
    default: return launch_scatter<256, 16, KIND, false>(L, pf);
  }
}

template <int KIND>
static hipError_t partition_keys_kind(const BbfPartLaunch& L, const PartFn& pf) {
  const hipStream_t st = L.stream;
  const uint32_t nbins = L.nparts * kLanes;
  const PartShape S = part_shape(L);
  if (!L.counts_ready) {
    const uint32_t nclear = L.counts_next ? 2 * nbins : nbins;  // counts_next follows counts
    hipLaunchKernelGGL(k_fill_zero_u32, dim3(fill_grid(nclear)), dim3(256), 0, st, L.counts,
                       static_cast<uint64_t>(nclear));
    if (L.counts_next) {
      const hipError_t ae = allow_dynamic_lds(reinterpret_cast<const void*>(k_part_hist2), 2 * nbins * 4u);
      if (ae != hipSuccess) return ae;
      hipLaunchKernelGGL(k_part_hist2, dim3(S.hist_grid), dim3(kHistThreads), 2 * nbins * 4u, st, L.keys, L.n,
                         pf, L.nparts, S.per_wg, S.tile_log2, L.counts, L.counts_next);
    } else {
      hipLaunchKernelGGL(k_part_hist<KIND>, dim3(S.hist_grid), dim3(kHistThreads), nbins * 4u, st, L.keys, L.n,
                         pf, L.nparts, S.per_wg, S.tile_log2, L.counts);
    }
  }
  return scan_and_scatter<KIND>(L, pf, S);
}

hipError_t partition_keys(const BbfPartLaunch& L) {
  const PartFn pf = make_part_fn(L);
  if (L.bin_cap) {  // capacity bins: no histogram pass
    const uint32_t nbins = L.nparts * kLanes;
    hipLaunchKernelGGL(k_part_init_bins, dim3((nbins + 255) / 256), dim3(256), 0, L.stream, L.cursor, nbins,
                       L.bin_cap);
    const uint32_t thr = part_shape(L).threads;
    if (L.part_kind == 1) {
      if (L.skip_bits) {
        if (!L.gate_count) return hipErrorInvalidValue;
        return thr == 1024 ? launch_scatter<1024, 16, 1, true, true>(L, pf)
               : thr == 512 ? launch_scatter<512, 16, 1, true, true>(L, pf)
                            : launch_scatter<256, 16, 1, true, true>(L, pf);
      }
      return thr == 1024 ? launch_scatter<1024, 16, 1, true>(L, pf)
             : thr == 512 ? launch_scatter<512, 16, 1, true>(L, pf)
                          : launch_scatter<256, 16, 1, true>(L, pf);
    }
    // 32K-key tiles (32 keys per thread, one workgroup per CU) for blocked
    // Bloom filters of 512..1024 partitions: the runs per bin double to ~64
    // keys, fewer partial lines; measured on one box, interleaved: 1 GiB
    // (512 partitions) 2.443 against 2.485 ms per call, but 128 MiB (64
    // partitions) 2.18 against 2.06 ms and the 512 MiB cuckoo table (256) 4.80
    // against 4.59 ms, where the runs are long already and the second
    // workgroup per CU counts more.  1024 partitions is the most whose
    // cursors and run bases fit in LDS beside the 128 KiB tile.
    if (thr == 1024 && L.cf_which == 0 && L.nparts >= 512 && L.nparts <= 1024)
      return launch_scatter<1024, 32, 0, true>(L, pf);
    return thr == 1024 ? launch_scatter<1024, 16, 0, true>(L, pf)
           : thr == 512 ? launch_scatter<512, 16, 0, true>(L, pf)
                        : launch_scatter<256, 16, 0, true>(L, pf);
  }
  return L.part_kind == 1 ? partition_keys_kind<1>(L, pf) : partition_keys_kind<0>(L, pf);
}

hipError_t part_gather_matches(const BbfPartLaunch& L, uint32_t* hp, const uint64_t* d_base, uint64_t* d_count) {
  hipError_t e = scan_tile_counts(L.ptile_cnt, L.ptiles, d_count, L.stream);
  if (e == hipSuccess) e = expand_gather(L.pbits, L.ptile_cnt, L.ptiles, L.pos_p, hp, d_base, L.stream);
  return e;
}

hipError_t part_sort_matches(const BbfPartLaunch& L, const uint32_t* hp, const uint64_t* d_hits, HitSortStage stage,
                             const uint64_t* gate, uint64_t thr) {
  const hipStream_t st = L.stream;
  const uint32_t G = L.hit_groups, nb = L.nbuckets;
  if (stage != kHitSortAll && gate == nullptr) return hipErrorInvalidValue;
  HitSel sel;
  sel.d_hits = d_hits;
  sel.gate = stage == kHitSortAll ? nullptr : gate;
  sel.thr = thr;
  sel.stage = stage;
  hipError_t e = allow_dynamic_lds(reinterpret_cast<const void*>(k_hit_place), nb * 8u);
  if (e == hipSuccess) e = allow_dynamic_lds(reinterpret_cast<const void*>(k_hit_bits), (1u << L.bucket_shift) / 8u);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_hit_count, dim3(G), dim3(256), nb * 4u, st, hp, sel, L.bucket_shift, nb, L.hcounts);
  hipLaunchKernelGGL(k_hit_colscan, dim3((nb + kColB - 1) / kColB), dim3(256), 0, st, L.hcounts, G, nb, L.bcounts,
                     sel);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(kPartScanThreads), 0, st, L.bcounts, nb, L.boffsets,
                     L.bcursor);
  // bucketed positions overwrite the grouped positions
  hipLaunchKernelGGL(k_hit_place, dim3(G), dim3(256), nb * 8u, st, hp, sel, L.bucket_shift, nb,
                     L.hcounts, L.boffsets, L.pos_p);
  uint32_t* out_pos = stage == kHitSortAll ? L.out_pos : nullptr;
  if (out_pos && ((1u << L.bucket_shift) / kTileKeys > kHitEmitTilesMax || !L.d_count_out))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_hit_bits, dim3(nb), dim3(256), (1u << L.bucket_shift) / 8u, st, L.pos_p, L.boffsets,
                     L.bucket_shift, L.ntiles, L.bits, L.tile_cnt, sel, out_pos, L.match_offset, L.d_count_out);
  return hipGetLastError();
}

hipError_t part_reorder_matches(const BbfPartLaunch& L) {
  // hp (the matches' original positions) overwrites the grouped keys
  hipError_t e = part_gather_matches(L, L.keys_p, nullptr, L.d_hits);
  if (e == hipSuccess) e = part_sort_matches(L, L.keys_p, L.d_hits);
  return e;
}

}  // namespace dtlh
