// dtl_cuckoo_host.cpp -- host build of the cuckoo filter table.
//
// The reference builds cuckoo tables on the host (partial-key cuckoo hashing
// with up to 500 kicks and a one-entry victim cache); the GPU only probes
// them.  This is that insert, writing the same bucket layout the probe
// kernels read:
//   insert              cuckoofilter_logic.hpp:291-322
//   insert_tag_to_bucket cuckoofilter_table.hpp:185-202
//   read_tag/write_tag   cuckoofilter_table.hpp:111-178 (little-endian packing)
//   alt_index            cuckoofilter_logic.hpp:214-252
// The kick slot comes from std::mt19937 like the reference's per-thread
// dtl::this_thread::rand32 (thread.hpp:71-79), here with an explicit seed so
// tables are reproducible.
#include <cstring>
#include <random>

#include "dtl_hip.h"
#include "dtl_sched.hpp"

namespace {

struct CfHost {
  const dtl_cf_desc* d;
  uint8_t* buf;
  uint32_t tag_mask;

  uint32_t bucket_index(uint32_t h0) const {
    if (d->addr_mode == DTL_ADDR_POW2) return h0 >> d->bucket_shift;
    const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(h0) * d->magic) >> 32) >> d->magic_shift;
    return h0 - q * d->bucket_cnt;
  }

  uint32_t alt(uint32_t i, uint32_t tag) const {
    const uint32_t s = tag * 0x5bd1e995u;
    if (d->addr_mode == DTL_ADDR_POW2) return (i ^ s) & d->bucket_mask;
    const uint32_t c = d->bucket_cnt;
    const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(s) * d->magic) >> 32) >> d->magic_shift;
    const uint32_t x = c - (s - q * c);
    return ((x < i) ? c : 0u) + (x - i);
  }

  uint32_t read_tag(uint64_t i, uint32_t j) const {
    const uint8_t* p = buf + i * d->bytes_per_bucket;
    uint32_t v = 0;
    switch (d->tag_bits) {
      case 4: v = p[j >> 1] >> ((j & 1u) << 2); break;
      case 8: v = p[j]; break;
      case 12: {
        uint16_t x;
        std::memcpy(&x, p + j + (j >> 1), 2);
        v = x >> ((j & 1u) << 2);
      } break;
      case 16: {
        uint16_t x;
        std::memcpy(&x, p + 2 * j, 2);
        v = x;
      } break;
      case 32: std::memcpy(&v, p + 4 * j, 4); break;
    }
    return v & tag_mask;
  }

  void write_tag(uint64_t i, uint32_t j, uint32_t t) const {
    uint8_t* p = buf + i * d->bytes_per_bucket;
    const uint32_t tag = t & tag_mask;
    switch (d->tag_bits) {
      case 4: {
        uint8_t& b = p[j >> 1];
        b = (j & 1u) ? static_cast<uint8_t>((b & 0x0f) | (tag << 4))
                     : static_cast<uint8_t>((b & 0xf0) | tag);
      } break;
      case 8: p[j] = static_cast<uint8_t>(tag); break;
      case 12: {
        uint16_t x;
        std::memcpy(&x, p + j + (j >> 1), 2);
        x = (j & 1u) ? static_cast<uint16_t>((x & 0x000f) | (tag << 4))
                     : static_cast<uint16_t>((x & 0xf000) | tag);
        std::memcpy(p + j + (j >> 1), &x, 2);
      } break;
      case 16: {
        const uint16_t x = static_cast<uint16_t>(tag);
        std::memcpy(p + 2 * j, &x, 2);
      } break;
      case 32: std::memcpy(p + 4 * j, &tag, 4); break;
    }
  }

  bool insert_tag_to_bucket(uint32_t i, uint32_t tag, bool kickout, uint32_t& oldtag,
                            std::mt19937& rng) const {
    for (uint32_t j = 0; j < d->tags_per_bucket; ++j) {
      if (read_tag(i, j) == 0) {
        write_tag(i, j, tag);
        return true;
      }
    }
    if (kickout) {
      const uint32_t r = static_cast<uint32_t>(rng() % d->tags_per_bucket);
      oldtag = read_tag(i, r);
      write_tag(i, r, tag);
    }
    return false;
  }

  bool insert(uint32_t key, std::mt19937& rng) const {
    uint8_t* victim = buf + d->victim_offset;  // {u32 index; u32 tag; u8 used}
    if (victim[8]) return false;
    const uint32_t i = bucket_index(key * dtlh::kMul.v[0]);
    uint32_t tag = (key * dtlh::kMul.v[1]) & tag_mask;
    tag += (tag == 0);
    uint32_t cur_i = i, cur_tag = tag;
    for (uint32_t count = 0; count < 500; ++count) {
      const bool kickout = count > 0;
      uint32_t oldtag = 0;
      if (insert_tag_to_bucket(cur_i, cur_tag, kickout, oldtag, rng)) return true;
      if (kickout) cur_tag = oldtag;
      cur_i = alt(cur_i, cur_tag);
    }
    std::memcpy(victim, &cur_i, 4);
    std::memcpy(victim + 4, &cur_tag, 4);
    victim[8] = 1;
    return true;
  }
};

}  // namespace

extern "C" int cf_build_host(const dtl_cf_desc* d, void* h_buf, const uint32_t* h_keys, uint64_t n,
                             uint32_t seed, uint64_t* inserted) {
  if (!d || !h_buf || (n && !h_keys) || d->bucket_cnt < 2) return DTL_EINVAL;
  CfHost h{d, static_cast<uint8_t*>(h_buf), static_cast<uint32_t>((1ull << d->tag_bits) - 1)};
  std::mt19937 rng(seed);
  uint64_t ok = 0;
  for (uint64_t i = 0; i < n; ++i) ok += h.insert(h_keys[i], rng) ? 1 : 0;
  if (inserted) *inserted = ok;
  return DTL_OK;
}
