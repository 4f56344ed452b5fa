// dtl_cuckoo.hip -- cuckoo filter probe and build on gfx950.
//
// One key per lane: i1 from h0, tag from h1 (0 -> 1), i2 = alt(i1, tag);
// both buckets are read with aligned loads (one 8-byte load each for the
// 16-bit / 4-slot table) and every slot is compared with the tag; the victim
// cache entry is read once per workgroup.  Semantics follow
// cuckoofilter_logic.hpp:346-362 (contains) with the bucket search of
// cuckoofilter_table_scalar.hpp:53-81 (equivalent to a per-slot compare).
// Output is the shared bitmap + tile-count format (dtl_kernels.hpp).
#include <hip/hip_runtime.h>

#include "dtl_cuckoo.hpp"
#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace dtlh {

template <int T, int B>
__global__ __launch_bounds__(kGlobalThreads) void k_cf_probe(const uint8_t* __restrict__ buf,
                                                             const uint32_t* __restrict__ keys,
                                                             uint32_t n, CfArgs a,
                                                             uint64_t* __restrict__ bits,
                                                             uint32_t* __restrict__ tile_cnt) {
  constexpr int BT = kGlobalThreads;
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int U = kGlobalUnroll;  // keys in flight per lane
  constexpr uint32_t kMask = static_cast<uint32_t>((1ull << T) - 1ull);
  __shared__ uint32_t red[BT / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  // victim cache {u32 index; u32 tag; bool used} right after the table
  const uint32_t* vp = reinterpret_cast<const uint32_t*>(buf + a.victim_offset);
  const uint32_t v_index = vp[0];
  const uint32_t v_tag = vp[1];
  const bool v_used = (vp[2] & 0xFFu) != 0;
  AddrArgs addr;
  addr.magic = a.magic;
  addr.mshift = a.mshift;
  addr.divisor = a.divisor;
  addr.post = a.post;
  const uint64_t tile_base = static_cast<uint64_t>(blockIdx.x) * kTileKeys;
  uint32_t cnt = 0;
#pragma unroll
  for (int r0 = 0; r0 < ROUNDS; r0 += U) {
    uint32_t i1[U], i2[U], tag[U];
    bool valid[U];
    Bucket<T, B> b1[U], b2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = tile_base + static_cast<uint64_t>(r0 + u) * BT + tid;
      valid[u] = i < n;
      const uint32_t key = load_key(keys, valid[u] ? i : n - 1);
      i1[u] = block_of(key * kMul.v[0], addr);
      tag[u] = cf_tag(key, kMask);
      i2[u] = cf_alt(i1[u], tag[u], a);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      b1[u].load(buf, i1[u]);
      b2[u].load(buf, i2[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bool hit = v_used & (tag[u] == v_tag) & ((i1[u] == v_index) | (i2[u] == v_index));
      hit |= b1[u].has(tag[u]);
      hit |= b2[u].has(tag[u]);
      hit &= valid[u];
      const uint64_t m = __ballot(hit);
      if (lane == 0) store_bits(bits, (tile_base + static_cast<uint64_t>(r0 + u) * BT) / 64 + wave, m);
      cnt += __popcll(m);
    }
  }
  const uint32_t total = block_sum_waves<BT>(cnt, red);
  if (tid == 0) tile_cnt[blockIdx.x] = total;
}

template <int T, int B>
hipError_t launch_cf_probe(const CfProbeLaunch& L) {
  CfArgs a;
  a.mode = L.addr_mode;
  if (L.addr_mode == 0) {
    a.magic = 0;
    a.mshift = 0;
    a.divisor = 0;
    a.post = L.shift;
  } else {
    a.magic = L.magic;
    a.mshift = L.mshift;
    a.divisor = L.bucket_cnt;
    a.post = 0;
  }
  a.bucket_mask = L.bucket_mask;
  a.bucket_cnt = L.bucket_cnt;
  a.pad = 0;
  a.victim_offset = L.victim_offset;
  hipLaunchKernelGGL((k_cf_probe<T, B>), dim3(L.ntiles), dim3(kGlobalThreads), 0, L.stream,
                     static_cast<const uint8_t*>(L.buf), L.keys, L.n, a, L.bits, L.tile_cnt);
  return hipGetLastError();
}

// One pass of the partitioned cuckoo probe over keys grouped by the slice of
// one of their buckets (XCD-ordered tiles, as k_bbf_probe_split): pass 1
// (which = 1, keys grouped by first bucket) tests the first bucket and the
// victim cache, pass 2 (keys grouped by alternate bucket) the alternate
// bucket.  A key is found iff one of its two passes finds it, as in
// cuckoofilter_logic.hpp:346-362.
template <int T, int B>
__global__ __launch_bounds__(kGlobalThreads) void k_cf_probe_part(const uint8_t* __restrict__ buf,
                                                                  const uint32_t* __restrict__ keys,
                                                                  uint32_t n, CfArgs a, uint32_t which,
                                                                  uint64_t* __restrict__ bits,
                                                                  uint32_t* __restrict__ tile_cnt,
                                                                  uint32_t tiles_per_xcd, uint32_t ntiles,
                                                                  const uint32_t* __restrict__ bin_fill,
                                                                  uint32_t tiles_per_bin) {
  constexpr int BT = kGlobalThreads;
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int U = ROUNDS;  // every key of the tile in flight (one bucket load each)
  constexpr uint32_t kMask = static_cast<uint32_t>((1ull << T) - 1ull);
  __shared__ uint32_t red[BT / 64];
  const uint32_t tile = (blockIdx.x & 7u) * tiles_per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;  // uniform per workgroup, before any barrier
  if (bin_fill) {  // keys grouped in bins of fixed capacity: this tile's bin ends at its fill
    n = bin_end_of(bin_fill, tile, tiles_per_bin);
    if (n <= tile * kTileKeys) {  // nothing of the bin in this tile (uniform, before any barrier)
      clear_tile(bits, tile_cnt, tile);
      return;
    }
  }
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t* vp = reinterpret_cast<const uint32_t*>(buf + a.victim_offset);
  const uint32_t v_index = vp[0];
  const uint32_t v_tag = vp[1];
  const bool v_used = which == 1 && (vp[2] & 0xFFu) != 0;
  AddrArgs addr;
  addr.magic = a.magic;
  addr.mshift = a.mshift;
  addr.divisor = a.divisor;
  addr.post = a.post;
  const uint64_t tile_base = static_cast<uint64_t>(tile) * kTileKeys;
  uint32_t cnt = 0;
  uint32_t i1[U], i2[U], tag[U];
  Bucket<T, B> bk[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t i = tile_base + static_cast<uint64_t>(u) * BT + tid;
    const uint32_t key = load_key(keys, i < n ? i : n - 1);
    i1[u] = block_of(key * kMul.v[0], addr);
    tag[u] = cf_tag(key, kMask);
    i2[u] = cf_alt(i1[u], tag[u], a);
  }
#pragma unroll
  for (int u = 0; u < U; ++u) bk[u].load(buf, which == 1 ? i1[u] : i2[u]);
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t g0 = tile_base + static_cast<uint64_t>(u) * BT + wave * 64u;
    bool hit = v_used & (tag[u] == v_tag) & ((i1[u] == v_index) | (i2[u] == v_index));
    hit |= bk[u].has(tag[u]);
    uint64_t m = __ballot(hit);
    if (g0 + 64 > n) m &= g0 >= n ? 0ull : (~0ull >> (64 - (n - g0)));
    if (lane == 0) store_bits(bits, g0 / 64, m);
    cnt += __popcll(m);
  }
  const uint32_t total = block_sum_waves<BT>(cnt, red);
  if (tid == 0) tile_cnt[tile] = total;
}

template <int T, int B>
hipError_t launch_cf_probe_part(const BbfPartLaunch& L) {
  CfArgs a;
  a.magic = L.addr_magic;
  a.mshift = L.addr_mshift;
  a.divisor = L.addr_divisor;
  a.post = L.addr_post;
  a.mode = L.cf_mode;
  a.bucket_mask = L.cf_bucket_mask;
  a.bucket_cnt = L.cf_bucket_cnt;
  a.pad = 0;
  a.victim_offset = L.cf_victim_offset;
  const uint32_t per_xcd = (L.ptiles + 7u) / 8u;
  const uint32_t* fill = L.bin_cap ? L.cursor : nullptr;
  hipLaunchKernelGGL((k_cf_probe_part<T, B>), dim3(per_xcd * 8u), dim3(kGlobalThreads), 0, L.stream,
                     static_cast<const uint8_t*>(L.words), L.keys_p, L.n, a, L.cf_which, L.pbits, L.ptile_cnt,
                     per_xcd, L.ptiles, fill, L.bin_cap / kTileKeys);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Build: concurrent partial-key cuckoo insertion.
//
// The reference inserts sequentially on the host (cuckoofilter_logic.hpp:
// 291-322): put the tag in the first empty slot of i1, else of i2, else kick
// a random slot of the current bucket and carry the evicted tag to its
// alternate bucket, at most 500 times, then park it in the victim cache.
// Here every lane runs that walk for its own keys at the same time.  Each
// step reads the bucket and changes one slot with one compare-and-swap on
// the aligned 4- or 8-byte word that holds it, so a tag is never lost or
// duplicated: a failed CAS (another lane changed the word) costs the step
// and the walk re-reads the same bucket.  Every tag therefore stays in one
// of the two buckets of the key it came from (alt() maps a tag's two buckets
// onto each other), which is all a probe relies on.  Which slots end up
// holding which tags depends on the interleaving, so tables differ from the
// host build's; the probe semantics do not.
//
// Lanes are independent state machines: a lane whose walk ends takes its
// next key at once, so a long walk holds up only its own lane.  A walk that
// reaches max_kicks steps, or whose alternate bucket is bucket_cnt (the
// MAGIC corner case of cuckoofilter_logic.hpp:241-250, which has no bucket
// in the table), claims the victim cache; if that is taken the tag has no
// place and is counted in *failed.
// ---------------------------------------------------------------------------
template <int T, int B>
struct BucketWords {
  static constexpr int BPB = (T * B + 7) / 8;
  static_assert(BPB == 1 || BPB == 2 || BPB == 4 || BPB == 8 || BPB == 16, "bucket straddles words");
  using W = std::conditional_t<(BPB >= 8), unsigned long long, uint32_t>;
  static constexpr int NW = BPB == 16 ? 2 : 1;   // CAS words per bucket
  static constexpr int SPW = B / NW;             // slots per CAS word
  static constexpr uint32_t kMask = static_cast<uint32_t>((1ull << T) - 1ull);
  W* p;          // first CAS word of the bucket
  uint32_t bo;   // bit offset of the bucket inside its word
  __device__ __forceinline__ void at(uint8_t* base, uint32_t i) {
    const uint64_t off = static_cast<uint64_t>(i) * BPB;
    if constexpr (BPB >= 8) {
      p = reinterpret_cast<W*>(base + off);
      bo = 0;
    } else {
      p = reinterpret_cast<W*>(base + (off & ~static_cast<uint64_t>(3)));
      bo = static_cast<uint32_t>(off & 3u) * 8u;
    }
  }
  __device__ __forceinline__ static int word_of(int j) { return j / SPW; }
  __device__ __forceinline__ uint32_t shift_of(int j) const { return bo + T * (j % SPW); }
};

__device__ __forceinline__ uint32_t cf_kick_slot(uint32_t key, uint32_t step, uint32_t seed) {
  uint32_t h = key * 0x9E3779B1u ^ (step * 0x85EBCA6Bu + seed);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  return h;
}

template <int T, int B>
__global__ __launch_bounds__(256) void k_cf_build(uint8_t* __restrict__ buf, const uint32_t* __restrict__ keys,
                                                  uint64_t n, CfArgs a, uint32_t max_kicks, uint32_t seed,
                                                  unsigned long long* __restrict__ failed) {
  using BW = BucketWords<T, B>;
  using W = typename BW::W;
  constexpr uint32_t kMask = BW::kMask;
  AddrArgs addr;
  addr.magic = a.magic;
  addr.mshift = a.mshift;
  addr.divisor = a.divisor;
  addr.post = a.post;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t key = 0, cur_i = 0, cur_tag = 0, step = 0;
  if (k < n) {
    key = keys[k];
    cur_i = block_of(key * kMul.v[0], addr);
    cur_tag = cf_tag(key, kMask);
  }
  while (k < n) {
    bool done = false;
    bool to_victim = step >= max_kicks;
    if (!to_victim) {
      BW bw;
      bw.at(buf, cur_i);
      W v[BW::NW];
#pragma unroll
      for (int w = 0; w < BW::NW; ++w) v[w] = __hip_atomic_load(bw.p + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int jfree = -1;
#pragma unroll
      for (int j = B - 1; j >= 0; --j)
        if (((v[BW::word_of(j)] >> bw.shift_of(j)) & kMask) == 0) jfree = j;
      if (jfree >= 0 || step > 0) {
        const int j = jfree >= 0 ? jfree : static_cast<int>(cf_kick_slot(key, step, seed) & (B - 1));
        const int w = BW::word_of(j);
        const uint32_t sh = bw.shift_of(j);
        W expect = v[w];
        const uint32_t old = static_cast<uint32_t>(expect >> sh) & kMask;
        const W desired = (expect & ~(static_cast<W>(kMask) << sh)) | (static_cast<W>(cur_tag) << sh);
        const bool swapped = __hip_atomic_compare_exchange_strong(bw.p + w, &expect, desired, __ATOMIC_RELAXED,
                                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (swapped) {
          if (jfree >= 0) {
            done = true;
          } else {
            cur_tag = old;  // carry the evicted tag to its other bucket
            cur_i = cf_alt(cur_i, cur_tag, a);
          }
        }
      } else {
        cur_i = cf_alt(cur_i, cur_tag, a);  // first step, both slots of i1 taken: try i2 without a kick
      }
      ++step;
      if (!done && a.mode != 0 && cur_i >= a.bucket_cnt) to_victim = true;
    }
    if (to_victim) {
      uint32_t* vp = reinterpret_cast<uint32_t*>(buf + a.victim_offset);  // {u32 index; u32 tag; u8 used}
      uint32_t zero = 0;
      if (__hip_atomic_compare_exchange_strong(vp + 2, &zero, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        __hip_atomic_store(vp, cur_i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(vp + 1, cur_tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        __hip_atomic_fetch_add(failed, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      done = true;
    }
    if (done) {
      k += stride;
      step = 0;
      if (k < n) {
        key = keys[k];
        cur_i = block_of(key * kMul.v[0], addr);
        cur_tag = cf_tag(key, kMask);
      }
    }
  }
}

template <int T, int B>
hipError_t launch_cf_build(const CfBuildLaunch& L) {
  CfArgs a;
  a.mode = L.addr_mode;
  if (L.addr_mode == 0) {
    a.magic = 0;
    a.mshift = 0;
    a.divisor = 0;
    a.post = L.shift;
  } else {
    a.magic = L.magic;
    a.mshift = L.mshift;
    a.divisor = L.bucket_cnt;
    a.post = 0;
  }
  a.bucket_mask = L.bucket_mask;
  a.bucket_cnt = L.bucket_cnt;
  a.pad = 0;
  a.victim_offset = L.victim_offset;
  const uint64_t want = (L.n + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(std::max(L.cu_count, 1)) * 8;
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
  hipLaunchKernelGGL((k_cf_build<T, B>), dim3(grid), dim3(256), 0, L.stream, static_cast<uint8_t*>(L.buf), L.keys,
                     L.n, a, L.max_kicks, L.seed, reinterpret_cast<unsigned long long*>(L.failed));
  return hipGetLastError();
}

using CfBuildFn = hipError_t (*)(const CfBuildLaunch&);

template <int T, int B>
constexpr CfBuildFn cf_build_fn() {
  constexpr int BPB = (T * B + 7) / 8;
  if constexpr (BPB == 1 || BPB == 2 || BPB == 4 || BPB == 8 || BPB == 16) return &launch_cf_build<T, B>;
  else return nullptr;
}

#define DTL_CF_ENTRY(t, b) {t, b, &launch_cf_probe<t, b>, &launch_cf_probe_part<t, b>, cf_build_fn<t, b>()}
static const CfEntry kCfEntries[] = {
    DTL_CF_ENTRY(4, 1),  DTL_CF_ENTRY(4, 2),  DTL_CF_ENTRY(4, 4),  DTL_CF_ENTRY(8, 1),  DTL_CF_ENTRY(8, 2),
    DTL_CF_ENTRY(8, 4),  DTL_CF_ENTRY(12, 1), DTL_CF_ENTRY(12, 2), DTL_CF_ENTRY(12, 4), DTL_CF_ENTRY(16, 1),
    DTL_CF_ENTRY(16, 2), DTL_CF_ENTRY(16, 4), DTL_CF_ENTRY(32, 1), DTL_CF_ENTRY(32, 2), DTL_CF_ENTRY(32, 4),
};
#undef DTL_CF_ENTRY

const CfEntry* cf_find(uint32_t tag_bits, uint32_t tags_per_bucket) {
  for (const CfEntry& e : kCfEntries)
    if (e.tag_bits == tag_bits && e.tags_per_bucket == tags_per_bucket) return &e;
  return nullptr;
}

}  // namespace dtlh
