// dtl_cuckoo.hip -- cuckoo filter probe on gfx950.
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
                                                                  uint32_t tiles_per_xcd, uint32_t ntiles) {
  constexpr int BT = kGlobalThreads;
  constexpr int ROUNDS = kTileKeys / BT;
  constexpr int U = ROUNDS;  // every key of the tile in flight (one bucket load each)
  constexpr uint32_t kMask = static_cast<uint32_t>((1ull << T) - 1ull);
  __shared__ uint32_t red[BT / 64];
  const uint32_t tile = (blockIdx.x & 7u) * tiles_per_xcd + (blockIdx.x >> 3);
  if (tile >= ntiles) return;  // uniform per workgroup, before any barrier
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
  const uint32_t per_xcd = (L.ntiles + 7u) / 8u;
  hipLaunchKernelGGL((k_cf_probe_part<T, B>), dim3(per_xcd * 8u), dim3(kGlobalThreads), 0, L.stream,
                     static_cast<const uint8_t*>(L.words), L.keys_p, L.n, a, L.cf_which, L.pbits, L.ptile_cnt,
                     per_xcd, L.ntiles);
  return hipGetLastError();
}

#define DTL_CF_ENTRY(t, b) {t, b, &launch_cf_probe<t, b>, &launch_cf_probe_part<t, b>}
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
