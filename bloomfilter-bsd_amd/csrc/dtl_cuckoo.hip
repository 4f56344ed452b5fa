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

#include "dtl_internal.hpp"
#include "dtl_kernels.hpp"

namespace dtlh {

struct CfArgs {
  uint32_t magic, mshift, divisor, post;  // bucket addressing, as AddrArgs
  uint32_t mode;                          // 0 = POW2, 1 = MAGIC
  uint32_t bucket_mask;                   // POW2
  uint32_t bucket_cnt;                    // MAGIC
  uint32_t pad;
  uint64_t victim_offset;
};

// alt_index (cuckoofilter_logic.hpp:214-252)
__device__ __forceinline__ uint32_t cf_alt(uint32_t i, uint32_t tag, const CfArgs& a) {
  const uint32_t s = tag * 0x5bd1e995u;
  if (a.mode == 0) return (i ^ s) & a.bucket_mask;
  const uint32_t q = __umulhi(s, a.magic) >> a.mshift;
  const uint32_t s_mod_m = s - q * a.divisor;
  const uint32_t x = a.bucket_cnt - s_mod_m;
  const uint32_t mask = (x < i) ? 0xFFFFFFFFu : 0u;
  return (mask & a.bucket_cnt) + (x - i);
}

// Little-endian bits of bucket i (bytes_per_bucket = (T*B+7)/8 <= 16).
template <int T, int B>
struct Bucket {
  static constexpr int BPB = (T * B + 7) / 8;
  uint64_t lo, hi;
  __device__ __forceinline__ void load(const uint8_t* __restrict__ base, uint32_t i) {
    const uint64_t off = static_cast<uint64_t>(i) * BPB;
    hi = 0;
    if constexpr (BPB == 16) {
      const uint4 v = *reinterpret_cast<const uint4*>(base + off);
      lo = (static_cast<uint64_t>(v.y) << 32) | v.x;
      hi = (static_cast<uint64_t>(v.w) << 32) | v.z;
    } else if constexpr (BPB == 8) {
      const uint2 v = *reinterpret_cast<const uint2*>(base + off);
      lo = (static_cast<uint64_t>(v.y) << 32) | v.x;
    } else if constexpr (BPB == 4) {
      lo = *reinterpret_cast<const uint32_t*>(base + off);
    } else if constexpr (BPB == 2) {
      lo = *reinterpret_cast<const uint16_t*>(base + off);
    } else if constexpr (BPB == 1) {
      lo = base[off];
    } else {
      // 3- or 6-byte buckets: two aligned dwords cover the bucket.
      const uint64_t a = off & ~static_cast<uint64_t>(3);
      const uint32_t sh = static_cast<uint32_t>(off & 3u) * 8u;
      const uint32_t* p = reinterpret_cast<const uint32_t*>(base + a);
      const uint64_t v = (static_cast<uint64_t>(p[1]) << 32) | p[0];
      lo = (v >> sh) & ((1ull << (BPB * 8)) - 1ull);
    }
  }
  __device__ __forceinline__ bool has(uint32_t tag) const {
    bool f = false;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int bit = T * j;
      const uint64_t word = bit < 64 ? lo : hi;
      const int sh = bit & 63;
      const uint32_t slot = static_cast<uint32_t>(word >> sh) & static_cast<uint32_t>((1ull << T) - 1ull);
      f |= slot == tag;
    }
    return f;
  }
};

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
      uint32_t t = (key * kMul.v[1]) & kMask;
      tag[u] = t + (t == 0);
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

static const CfEntry kCfEntries[] = {
    {4, 1, &launch_cf_probe<4, 1>},   {4, 2, &launch_cf_probe<4, 2>},   {4, 4, &launch_cf_probe<4, 4>},
    {8, 1, &launch_cf_probe<8, 1>},   {8, 2, &launch_cf_probe<8, 2>},   {8, 4, &launch_cf_probe<8, 4>},
    {12, 1, &launch_cf_probe<12, 1>}, {12, 2, &launch_cf_probe<12, 2>}, {12, 4, &launch_cf_probe<12, 4>},
    {16, 1, &launch_cf_probe<16, 1>}, {16, 2, &launch_cf_probe<16, 2>}, {16, 4, &launch_cf_probe<16, 4>},
    {32, 1, &launch_cf_probe<32, 1>}, {32, 2, &launch_cf_probe<32, 2>}, {32, 4, &launch_cf_probe<32, 4>},
};

const CfEntry* cf_find(uint32_t tag_bits, uint32_t tags_per_bucket) {
  for (const CfEntry& e : kCfEntries)
    if (e.tag_bits == tag_bits && e.tags_per_bucket == tags_per_bucket) return &e;
  return nullptr;
}

}  // namespace dtlh
