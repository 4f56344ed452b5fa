// dtl_cuckoo.hpp -- cuckoo filter addressing and bucket test, shared by the
// probe kernels (dtl_cuckoo.hip) and the partitioned probe's partition
// function (dtl_partition.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

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

// Tag of a key: T low bits of h1, 0 replaced by 1 (cuckoofilter_logic.hpp:180-186, 204-207).
__device__ __forceinline__ uint32_t cf_tag(uint32_t key, uint32_t tag_mask) {
  const uint32_t t = (key * kMul.v[1]) & tag_mask;
  return t + (t == 0);
}

}  // namespace dtlh
