// dtl/filter.hpp -- C++ host API of libdtl_hip, mirroring the reference's
// filter classes (namespace dtl) on top of the C ABI in dtl_hip.h.
//
//   dtl::blocked_bloomfilter<Tw>          blocked_bloomfilter.hpp:27-393
//   dtl::zoned_blocked_bloomfilter<Tw>    zoned_blocked_bloomfilter.hpp:27-381
//   dtl::cuckoofilter::cuckoofilter       cuckoofilter/cuckoofilter.hpp:34-346
//   dtl::filter::filter_base              filter_base.hpp:11-68
//   dtl::bbf_32 / bbf_64 / zbbf_32 / zbbf_64 / cf   bbf_32.cpp, bbf_64.cpp, ..., cf.cpp
//   dtl::filter::filter (+ probe_t)       filter.hpp:35-261
//
// Same names, arguments and error behaviour (std::invalid_argument,
// std::logic_error, std::runtime_error) as the reference.  Differences that
// follow from the GPU:
//   * filter words, keys and match-position buffers of the blocked Bloom
//     filters are DEVICE pointers (HBM); every operation runs on the GPU;
//   * the cuckoo filter is built on the host (insert / batch_insert take host
//     table memory, as the reference builds it) or on the GPU
//     (batch_insert_device, concurrent inserts) and probed on the GPU
//     (contains / batch_contains take the device copy of the table);
//   * the table buffer holds the victim cache after the buckets: allocate
//     buffer_size_in_bytes(), not size_in_bytes();
//   * batch_contains returns after the GPU work finished (it reads the match
//     count back); batch_contains_async leaves the count on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../dtl_hip.h"

namespace dtl {

using u1 = bool;
using u32 = uint32_t;
using u64 = uint64_t;

enum class block_addressing : u32 { POWER_OF_TWO, MAGIC, DYNAMIC };

namespace detail {

inline void throw_status(int rc, const std::string& what) {
  switch (rc) {
    case DTL_OK: return;
    case DTL_EINVAL: throw std::invalid_argument(what);
    case DTL_ELOGIC: throw std::logic_error("The block count must not exceed 2^32.");
    case DTL_EHIP: throw std::runtime_error(what + ": " + dtl_status_string(rc) + ": " + dtl_last_error());
    default: throw std::runtime_error(what + ": " + dtl_status_string(rc));
  }
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// A small device buffer per thread for single-key calls and match counts.
struct device_scratch {
  void* ptr = nullptr;
  int dev = -1;
  ~device_scratch() {
    if (ptr) (void)hipFree(ptr);
  }
  void* get() {
    int d = 0;
    check_hip(hipGetDevice(&d), "hipGetDevice");
    if (!ptr || d != dev) {
      if (ptr) (void)hipFree(ptr);
      check_hip(hipMalloc(&ptr, 64), "hipMalloc");
      dev = d;
    }
    return ptr;
  }
};

inline void* scratch64() {
  thread_local device_scratch s;
  return s.get();
}

inline u64 read_count(const u64* d_count, hipStream_t stream) {
  u64 c = 0;
  check_hip(hipMemcpyAsync(&c, d_count, sizeof(c), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
  check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return c;
}

// Common part of the blocked Bloom filter front-ends.
class bbf_handle {
 public:
  dtl_bbf_desc desc{};

  void insert(void* d_filter, u32 key) const {
    u32* k = static_cast<u32*>(scratch64());
    check_hip(hipMemcpy(k, &key, 4, hipMemcpyHostToDevice), "hipMemcpy");
    throw_status(bbf_build(&desc, d_filter, k, 1, nullptr), "bbf_build");
    check_hip(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
  }
  void batch_insert(void* d_filter, const u32* d_keys, u32 key_cnt, hipStream_t stream = nullptr) const {
    throw_status(bbf_build(&desc, d_filter, d_keys, key_cnt, stream), "bbf_build");
    check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
  u1 contains(const void* d_filter, u32 key) const {
    char* s = static_cast<char*>(scratch64());
    u32* k = reinterpret_cast<u32*>(s);
    u32* pos = reinterpret_cast<u32*>(s + 8);
    u64* cnt = reinterpret_cast<u64*>(s + 16);
    check_hip(hipMemcpy(k, &key, 4, hipMemcpyHostToDevice), "hipMemcpy");
    throw_status(bbf_probe(&desc, d_filter, k, 1, 0, pos, cnt, nullptr), "bbf_probe");
    return read_count(cnt, nullptr) != 0;
  }
  u64 batch_contains(const void* d_filter, const u32* d_keys, u32 key_cnt, u32* d_match_positions,
                     u32 match_offset, hipStream_t stream = nullptr) const {
    u64* cnt = reinterpret_cast<u64*>(static_cast<char*>(scratch64()) + 24);
    throw_status(bbf_probe(&desc, d_filter, d_keys, key_cnt, match_offset, d_match_positions, cnt, stream),
                 "bbf_probe");
    return read_count(cnt, stream);
  }
  void batch_contains_async(const void* d_filter, const u32* d_keys, u32 key_cnt, u32* d_match_positions,
                            u32 match_offset, u64* d_count, hipStream_t stream) const {
    throw_status(bbf_probe(&desc, d_filter, d_keys, key_cnt, match_offset, d_match_positions, d_count, stream),
                 "bbf_probe");
  }
};

}  // namespace detail

//===----------------------------------------------------------------------===//
// blocked_bloomfilter_config (blocked_bloomfilter_config.hpp:11-51)
struct blocked_bloomfilter_config {
  u32 k = 8;
  u32 word_size = 4;  // [byte]
  u32 word_cnt_per_block = 1;
  u32 sector_cnt = 1;
  block_addressing addr_mode = block_addressing::POWER_OF_TWO;
  u32 zone_cnt = 1;

  bool operator<(const blocked_bloomfilter_config& o) const {
    if (k != o.k) return k < o.k;
    if (word_size != o.word_size) return word_size < o.word_size;
    if (word_cnt_per_block != o.word_cnt_per_block) return word_cnt_per_block < o.word_cnt_per_block;
    if (sector_cnt != o.sector_cnt) return sector_cnt < o.sector_cnt;
    if (addr_mode != o.addr_mode) return addr_mode < o.addr_mode;
    return zone_cnt < o.zone_cnt;
  }
  bool operator==(const blocked_bloomfilter_config& o) const {
    return k == o.k && word_size == o.word_size && word_cnt_per_block == o.word_cnt_per_block &&
           sector_cnt == o.sector_cnt && addr_mode == o.addr_mode && zone_cnt == o.zone_cnt;
  }
  bool operator!=(const blocked_bloomfilter_config& o) const { return !(*this == o); }
  void print(std::ostream& os) const {
    os << "bbf,k=" << k << ",word_size=" << word_size << ",word_cnt=" << word_cnt_per_block
       << ",sector_cnt=" << sector_cnt << ",addr=" << (addr_mode == block_addressing::POWER_OF_TWO ? "pow2" : "magic")
       << ",zone_cnt=" << zone_cnt;
  }
};

//===----------------------------------------------------------------------===//
// dtl::blocked_bloomfilter<Tw>: sectorized blocked Bloom filter.
template <typename Tw = u32>
struct blocked_bloomfilter : detail::bbf_handle {
  static_assert(sizeof(Tw) == 4 || sizeof(Tw) == 8, "word type must be 32 or 64 bits");
  using key_t = u32;
  using hash_value_t = u32;
  using word_t = Tw;

  u64 m;
  u64 m_actual;
  u32 k;
  u32 word_cnt_per_block;
  u32 sector_cnt;

  blocked_bloomfilter(std::size_t m, u32 k, u32 word_cnt_per_block = 1, u32 sector_cnt = 1)
      : m(m), k(k), word_cnt_per_block(word_cnt_per_block), sector_cnt(sector_cnt) {
    detail::throw_status(dtl_bbf_desc_init(&desc, sizeof(Tw) * 8, m, k, word_cnt_per_block, sector_cnt, 1),
                         "Invalid configuration: k=" + std::to_string(k) + ", w=" +
                             std::to_string(word_cnt_per_block) + ", s=" + std::to_string(sector_cnt));
    m_actual = desc.m_actual;
  }

  void insert(word_t* d_filter, key_t key) { bbf_handle::insert(d_filter, key); }
  void batch_insert(word_t* d_filter, const key_t* d_keys, u32 key_cnt) {
    bbf_handle::batch_insert(d_filter, d_keys, key_cnt);
  }
  u1 contains(const word_t* d_filter, key_t key) const { return bbf_handle::contains(d_filter, key); }
  u64 batch_contains(const word_t* d_filter, const key_t* d_keys, u32 key_cnt, u32* d_match_positions,
                     u32 match_offset) const {
    return bbf_handle::batch_contains(d_filter, d_keys, key_cnt, d_match_positions, match_offset);
  }

  block_addressing get_addressing_mode() const {
    return desc.addr_mode == DTL_ADDR_POW2 ? block_addressing::POWER_OF_TWO : block_addressing::MAGIC;
  }
  std::size_t size_in_bytes() const noexcept { return (m_actual + 7) / 8; }
  std::size_t size() const noexcept { return (m_actual + sizeof(Tw) * 8 - 1) / (sizeof(Tw) * 8); }
  std::string name() const {
    return "{\"name\":\"blocked_bloom_multiword\",\"size\":" + std::to_string(size_in_bytes()) +
           ",\"word_size\":" + std::to_string(sizeof(word_t)) + ",\"k\":" + std::to_string(k) +
           ",\"w\":" + std::to_string(word_cnt_per_block) + ",\"s\":" + std::to_string(sector_cnt) +
           ",\"u\":0,\"e\":0,\"addr\":" +
           (get_addressing_mode() == block_addressing::POWER_OF_TWO ? "\"pow2\"" : "\"magic\"") + "}";
  }
};

//===----------------------------------------------------------------------===//
// dtl::zoned_blocked_bloomfilter<Tw>
template <typename Tw = u32>
struct zoned_blocked_bloomfilter : detail::bbf_handle {
  using key_t = u32;
  using hash_value_t = u32;
  using word_t = Tw;

  u64 m;
  u64 m_actual;
  u32 k;
  u32 word_cnt_per_block;
  u32 zone_cnt;

  zoned_blocked_bloomfilter(std::size_t m, u32 k, u32 word_cnt_per_block, u32 zone_cnt)
      : m(m), k(k), word_cnt_per_block(word_cnt_per_block), zone_cnt(zone_cnt) {
    if (word_cnt_per_block <= zone_cnt)
      throw std::invalid_argument("Invalid configuration: w=" + std::to_string(word_cnt_per_block) +
                                  ", z=" + std::to_string(zone_cnt));
    detail::throw_status(dtl_bbf_desc_init(&desc, sizeof(Tw) * 8, m, k, word_cnt_per_block, 0, zone_cnt),
                         "Invalid configuration: k=" + std::to_string(k) + ", w=" +
                             std::to_string(word_cnt_per_block) + ", z=" + std::to_string(zone_cnt));
    m_actual = desc.m_actual;
  }

  void insert(word_t* d_filter, key_t key) { bbf_handle::insert(d_filter, key); }
  void batch_insert(word_t* d_filter, const key_t* d_keys, u32 key_cnt) {
    bbf_handle::batch_insert(d_filter, d_keys, key_cnt);
  }
  u1 contains(const word_t* d_filter, key_t key) const { return bbf_handle::contains(d_filter, key); }
  u64 batch_contains(const word_t* d_filter, const key_t* d_keys, u32 key_cnt, u32* d_match_positions,
                     u32 match_offset) const {
    return bbf_handle::batch_contains(d_filter, d_keys, key_cnt, d_match_positions, match_offset);
  }
  block_addressing get_addressing_mode() const {
    return desc.addr_mode == DTL_ADDR_POW2 ? block_addressing::POWER_OF_TWO : block_addressing::MAGIC;
  }
  std::size_t size_in_bytes() const noexcept { return (m_actual + 7) / 8; }
  std::size_t size() const noexcept { return (m_actual + sizeof(Tw) * 8 - 1) / (sizeof(Tw) * 8); }
  std::string name() const {
    return "{\"name\":\"blocked_bloom_multiword\",\"size\":" + std::to_string(size_in_bytes()) +
           ",\"word_size\":" + std::to_string(sizeof(word_t)) + ",\"k\":" + std::to_string(k) +
           ",\"w\":" + std::to_string(word_cnt_per_block) + ",\"s\":" + std::to_string(word_cnt_per_block) +
           ",\"z\":" + std::to_string(zone_cnt) + ",\"u\":0,\"e\":0,\"addr\":" +
           (get_addressing_mode() == block_addressing::POWER_OF_TWO ? "\"pow2\"" : "\"magic\"") + "}";
  }
};

//===----------------------------------------------------------------------===//
namespace cuckoofilter {

// cuckoofilter_config.hpp:13-44
struct config {
  u32 bits_per_tag = 8;
  u32 tags_per_bucket = 2;
  block_addressing addr_mode = block_addressing::POWER_OF_TWO;
  bool operator<(const config& o) const {
    if (bits_per_tag != o.bits_per_tag) return bits_per_tag < o.bits_per_tag;
    if (tags_per_bucket != o.tags_per_bucket) return tags_per_bucket < o.tags_per_bucket;
    return addr_mode < o.addr_mode;
  }
  bool operator==(const config& o) const {
    return bits_per_tag == o.bits_per_tag && tags_per_bucket == o.tags_per_bucket && addr_mode == o.addr_mode;
  }
  bool operator!=(const config& o) const { return !(*this == o); }
};

struct cuckoofilter {
  using key_t = u32;
  using word_t = u32;

  u64 m;
  u64 m_actual;
  u32 bits_per_tag;
  u32 tags_per_bucket;
  dtl_cf_desc desc{};
  std::mt19937 rng;

  cuckoofilter(std::size_t m, u32 bits_per_tag, u32 tags_per_bucket, u32 seed = 5489u)
      : m(m), bits_per_tag(bits_per_tag), tags_per_bucket(tags_per_bucket), rng(seed) {
    const int rc = dtl_cf_desc_init(&desc, m, bits_per_tag, tags_per_bucket);
    if (rc == DTL_EINVAL) {
      const bool tag_ok = bits_per_tag == 4 || bits_per_tag == 8 || bits_per_tag == 12 ||
                          bits_per_tag == 16 || bits_per_tag == 32;
      throw std::invalid_argument(tag_ok ? "The given bucket size (associativity) is not supported."
                                         : "The given tag size is not supported.");
    }
    detail::throw_status(rc, "cuckoofilter");
    m_actual = desc.m_actual;
  }

  // Host build (h_table: buffer_size_in_bytes() zeroed host bytes).
  u1 insert(word_t* h_table, key_t key) { return batch_insert(h_table, &key, 1); }
  u1 batch_insert(word_t* h_table, const key_t* h_keys, u32 key_cnt) {
    u64 inserted = 0;
    detail::throw_status(cf_build_host(&desc, h_table, h_keys, key_cnt, static_cast<u32>(rng()), &inserted),
                         "cf_build_host");
    return inserted == key_cnt;
  }

  // GPU build (d_table: buffer_size_in_bytes() zeroed device bytes; d_keys
  // device keys): every key inserted at once (cf_build).  True if every walk
  // found a place, i.e. the table holds every key (the reference's
  // batch_insert returns true while its victim cache takes the overflow).
  // Not available for 12-bit tags with 2 or 4 slots (std::invalid_argument).
  u1 batch_insert_device(word_t* d_table, const key_t* d_keys, u64 key_cnt, hipStream_t stream = nullptr) {
    u64* failed = reinterpret_cast<u64*>(static_cast<char*>(detail::scratch64()) + 32);
    detail::check_hip(hipMemsetAsync(failed, 0, sizeof(u64), stream), "hipMemsetAsync");
    detail::throw_status(cf_build(&desc, d_table, d_keys, key_cnt, static_cast<u32>(rng()), 0, failed, stream),
                         "cf_build");
    return detail::read_count(failed, stream) == 0;
  }

  // GPU probe (d_table: device copy of the table buffer).
  u1 contains(const word_t* d_table, key_t key) const {
    char* s = static_cast<char*>(detail::scratch64());
    u32* k = reinterpret_cast<u32*>(s);
    u32* pos = reinterpret_cast<u32*>(s + 8);
    u64* cnt = reinterpret_cast<u64*>(s + 16);
    detail::check_hip(hipMemcpy(k, &key, 4, hipMemcpyHostToDevice), "hipMemcpy");
    detail::throw_status(cf_probe(&desc, d_table, k, 1, 0, pos, cnt, nullptr), "cf_probe");
    return detail::read_count(cnt, nullptr) != 0;
  }
  u64 batch_contains(const word_t* d_table, const key_t* d_keys, u32 key_cnt, u32* d_match_positions,
                     u32 match_offset, hipStream_t stream = nullptr) const {
    u64* cnt = reinterpret_cast<u64*>(static_cast<char*>(detail::scratch64()) + 24);
    detail::throw_status(cf_probe(&desc, d_table, d_keys, key_cnt, match_offset, d_match_positions, cnt, stream),
                         "cf_probe");
    return detail::read_count(cnt, stream);
  }
  void batch_contains_async(const word_t* d_table, const key_t* d_keys, u32 key_cnt, u32* d_match_positions,
                            u32 match_offset, u64* d_count, hipStream_t stream) const {
    detail::throw_status(cf_probe(&desc, d_table, d_keys, key_cnt, match_offset, d_match_positions, d_count, stream),
                         "cf_probe");
  }

  // cuckoofilter_logic.hpp:495-499 (host table; slot j of a bucket holds
  // bits [j*t, (j+1)*t) of the bucket's little-endian bytes)
  u64 count_occupied_slots(const word_t* h_table) const {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(h_table);
    const u64 mask = (1ull << bits_per_tag) - 1;
    u64 c = 0;
    for (u64 i = 0; i < desc.bucket_cnt; ++i) {
      for (u32 j = 0; j < tags_per_bucket; ++j) {
        const u64 bit = i * desc.bytes_per_bucket * 8 + static_cast<u64>(bits_per_tag) * j;
        u64 v = 0;
        for (u32 b = 0; b < 5; ++b) v |= static_cast<u64>(p[bit / 8 + b]) << (8 * b);
        c += ((v >> (bit % 8)) & mask) != 0;
      }
    }
    return c + (p[desc.victim_offset + 8] != 0);
  }
  u64 get_bucket_count() const { return desc.bucket_cnt; }
  block_addressing get_addressing_mode() const {
    return desc.addr_mode == DTL_ADDR_POW2 ? block_addressing::POWER_OF_TWO : block_addressing::MAGIC;
  }
  std::size_t size_in_bytes() const noexcept { return (m_actual + 7) / 8; }
  std::size_t size() const noexcept { return (m_actual + 31) / 32; }
  std::size_t buffer_size_in_bytes() const noexcept { return desc.buffer_bytes; }
  std::string name() const {
    return "{\"name\":\"cuckoo\",\"size\":" + std::to_string(size_in_bytes()) +
           ",\"tag_bits\":" + std::to_string(bits_per_tag) + ",\"associativity\":" +
           std::to_string(tags_per_bucket) + ",\"delete_support\":false,\"u\":0,\"addr\":" +
           (get_addressing_mode() == block_addressing::POWER_OF_TWO ? "\"pow2\"" : "\"magic\"") + "}";
  }
};

}  // namespace cuckoofilter

//===----------------------------------------------------------------------===//
// Type-erased filters (filter_base.hpp:11-68) over u64 word buffers.
namespace filter {

class filter_base {
 public:
  using key_t = u32;
  using word_t = u64;
  virtual u1 insert(word_t* filter_data, key_t key) = 0;
  virtual u1 batch_insert(word_t* filter_data, const key_t* keys, u32 key_cnt) = 0;
  virtual u1 contains(const word_t* filter_data, key_t key) const = 0;
  virtual u64 batch_contains(const word_t* filter_data, const key_t* keys, u32 key_cnt, u32* match_positions,
                             u32 match_offset) const = 0;
  virtual std::string name() const = 0;
  virtual std::size_t size_in_bytes() const = 0;
  /// Size in 64-bit words (of the buffer to allocate on the device).
  virtual std::size_t size() const = 0;
  filter_base() = default;
  virtual ~filter_base() = default;
};

}  // namespace filter

template <class Impl>
class bbf_wrapper : public filter::filter_base {
 protected:
  Impl impl;

 public:
  template <class... A>
  explicit bbf_wrapper(A... a) : impl(a...) {}
  u1 insert(word_t* filter_data, key_t key) override {
    impl.insert(reinterpret_cast<typename Impl::word_t*>(filter_data), key);
    return true;  // inserts never fail (bbf_32.cpp:38)
  }
  u1 batch_insert(word_t* filter_data, const key_t* keys, u32 key_cnt) override {
    impl.batch_insert(reinterpret_cast<typename Impl::word_t*>(filter_data), keys, key_cnt);
    return true;
  }
  u1 contains(const word_t* filter_data, key_t key) const override {
    return impl.contains(reinterpret_cast<const typename Impl::word_t*>(filter_data), key);
  }
  u64 batch_contains(const word_t* filter_data, const key_t* keys, u32 key_cnt, u32* match_positions,
                     u32 match_offset) const override {
    return impl.batch_contains(reinterpret_cast<const typename Impl::word_t*>(filter_data), keys, key_cnt,
                               match_positions, match_offset);
  }
  std::string name() const override { return impl.name(); }
  std::size_t size_in_bytes() const override { return impl.size_in_bytes(); }
  std::size_t size() const override { return (impl.size_in_bytes() + 7) / 8; }
  const Impl& instance() const { return impl; }
};

// bbf_32.hpp / bbf_64.hpp / zbbf_32.hpp / zbbf_64.hpp
class bbf_32 : public bbf_wrapper<blocked_bloomfilter<u32>> {
 public:
  bbf_32(std::size_t m, u32 k, u32 word_cnt_per_block = 1, u32 sector_cnt = 1)
      : bbf_wrapper(m, k, word_cnt_per_block, sector_cnt) {}
};
class bbf_64 : public bbf_wrapper<blocked_bloomfilter<u64>> {
 public:
  bbf_64(std::size_t m, u32 k, u32 word_cnt_per_block = 1, u32 sector_cnt = 1)
      : bbf_wrapper(m, k, word_cnt_per_block, sector_cnt) {}
};
class zbbf_32 : public bbf_wrapper<zoned_blocked_bloomfilter<u32>> {
 public:
  zbbf_32(std::size_t m, u32 k, u32 word_cnt_per_block, u32 zone_cnt)
      : bbf_wrapper(m, k, word_cnt_per_block, zone_cnt) {}
};
class zbbf_64 : public bbf_wrapper<zoned_blocked_bloomfilter<u64>> {
 public:
  zbbf_64(std::size_t m, u32 k, u32 word_cnt_per_block, u32 zone_cnt)
      : bbf_wrapper(m, k, word_cnt_per_block, zone_cnt) {}
};

// cf.hpp: probe side of the cuckoo filter behind filter_base.  The device
// buffer must be buffer_size_in_bytes() long; insert/batch_insert build into
// a host shadow table and upload it, so they take the DEVICE buffer too.
class cf : public filter::filter_base {
  cuckoofilter::cuckoofilter impl;
  std::vector<uint8_t> shadow;

  void upload(word_t* d) {
    detail::check_hip(hipMemcpy(d, shadow.data(), shadow.size(), hipMemcpyHostToDevice), "hipMemcpy");
  }

 public:
  cf(std::size_t m, u32 bits_per_tag = 16, u32 tags_per_bucket = 4)
      : impl(m, bits_per_tag, tags_per_bucket), shadow(impl.buffer_size_in_bytes(), 0) {}
  u1 insert(word_t* filter_data, key_t key) override {
    const u1 ok = impl.insert(reinterpret_cast<u32*>(shadow.data()), key);
    upload(filter_data);
    return ok;
  }
  u1 batch_insert(word_t* filter_data, const key_t* h_keys, u32 key_cnt) override {
    const u1 ok = impl.batch_insert(reinterpret_cast<u32*>(shadow.data()), h_keys, key_cnt);
    upload(filter_data);
    return ok;
  }
  u1 contains(const word_t* filter_data, key_t key) const override {
    return impl.contains(reinterpret_cast<const u32*>(filter_data), key);
  }
  u64 batch_contains(const word_t* filter_data, const key_t* keys, u32 key_cnt, u32* match_positions,
                     u32 match_offset) const override {
    return impl.batch_contains(reinterpret_cast<const u32*>(filter_data), keys, key_cnt, match_positions,
                               match_offset);
  }
  std::string name() const override { return impl.name(); }
  std::size_t size_in_bytes() const override { return impl.size_in_bytes(); }
  std::size_t size() const override { return (impl.buffer_size_in_bytes() + 7) / 8; }
  u64 count_occupied_slots() const { return impl.count_occupied_slots(reinterpret_cast<const u32*>(shadow.data())); }
  u64 get_bucket_count() const { return impl.get_bucket_count(); }
};

namespace filter {
namespace model {

// Stands in for the reference's model::calibration_data + skyline matrix
// (model/calibration_data.hpp:34-100, skyline_matrix.hpp:118-155): the
// GPU-calibrated cost tables and the learned MLP that choose a blocked Bloom
// filter configuration from (n, tw).  tw is the work per qualifying tuple in
// nanoseconds.
class cost_model {
  dtl_cost_model* m = nullptr;
  explicit cost_model(dtl_cost_model* p) : m(p) {}

 public:
  cost_model(cost_model&& o) noexcept : m(o.m) { o.m = nullptr; }
  cost_model& operator=(cost_model&& o) noexcept {
    std::swap(m, o.m);
    return *this;
  }
  cost_model(const cost_model&) = delete;
  cost_model& operator=(const cost_model&) = delete;
  ~cost_model() {
    if (m) dtl_model_destroy(m);
  }

  /// Runs the calibration sweep on the current GPU (params = nullptr: the
  /// full sweep, 2^13..2^33 bits) and, if `train`, trains the MLP on 25% of
  /// the calibrated cells.
  static cost_model calibrate(const dtl_calib_params* params = nullptr, bool train = true, u32 samples = 4000000,
                              u32 epochs = 16) {
    dtl_cost_model* p = nullptr;
    detail::throw_status(dtl_model_calibrate(params, &p, nullptr), "calibration");
    cost_model cm(p);
    if (train) detail::throw_status(dtl_model_train(p, 0.25, samples, epochs, 1, nullptr), "model training");
    return cm;
  }
  static cost_model load(const std::string& path) {
    dtl_cost_model* p = nullptr;
    if (dtl_model_load(path.c_str(), &p) != DTL_OK)
      throw std::runtime_error("Failed to load the filter cost model from " + path + ".");
    return cost_model(p);
  }
  void save(const std::string& path) const { detail::throw_status(dtl_model_save(m, path.c_str()), "model save"); }

  /// Process-wide model loaded from $DTL_HIP_MODEL or ~/.dtl-hip-model.bin
  /// (the reference reads ~/.dtl-filter.dat, calibration_data.hpp:92-100);
  /// throws if there is none (write one with calibrate().save(path)).
  static const cost_model& get_default_instance() {
    static const cost_model inst = []() {
      const char* env = std::getenv("DTL_HIP_MODEL");
      const char* home = std::getenv("HOME");
      const std::string path = env ? std::string(env) : std::string(home ? home : ".") + "/.dtl-hip-model.bin";
      dtl_cost_model* p = nullptr;
      if (dtl_model_load(path.c_str(), &p) != DTL_OK)
        throw std::runtime_error("Failed to determine filter configuration due to missing calibration data (" +
                                 path + ").");
      return cost_model(p);
    }();
    return inst;
  }

  /// Best (configuration, m) for n keys and tw ns per qualifying tuple: the
  /// MLP's argmin, or with use_mlp = false the calibrated table's (exhaustive).
  dtl_bbf_choice choose(u64 n, u64 tw_ns, bool use_mlp = true) const {
    dtl_bbf_choice ch;
    detail::throw_status(dtl_model_choose(m, n, static_cast<double>(tw_ns), use_mlp ? 1 : 0, &ch, nullptr),
                         "model choose");
    return ch;
  }
  const dtl_cost_model* handle() const { return m; }
};

}  // namespace model

// filter.hpp:35-261: owns the filter data (one device allocation instead of
// NUMA replicas); keys and match buffers may live on the host or the device
// (host buffers are staged through device memory).
class filter {
 public:
  using word_t = u64;

 private:
  struct state {
    std::shared_ptr<filter_base> instance;
    u64* data = nullptr;
    ~state() {
      if (data) (void)hipFree(data);
    }
  };
  std::shared_ptr<state> st;

  static std::shared_ptr<filter_base> construct(const blocked_bloomfilter_config& c, u64 m) {
    switch (c.word_size) {
      case 4:
        if (c.zone_cnt == 1) return std::make_shared<bbf_32>(m, c.k, c.word_cnt_per_block, c.sector_cnt);
        return std::make_shared<zbbf_32>(m, c.k, c.word_cnt_per_block, c.zone_cnt);
      case 8:
        if (c.zone_cnt == 1) return std::make_shared<bbf_64>(m, c.k, c.word_cnt_per_block, c.sector_cnt);
        return std::make_shared<zbbf_64>(m, c.k, c.word_cnt_per_block, c.zone_cnt);
      default: throw std::runtime_error("Illegal configuration. Word size must be either 4 or 8.");
    }
  }

  void allocate() {
    const std::size_t bytes = st->instance->size() * sizeof(u64);
    detail::check_hip(hipMalloc(&st->data, bytes), "hipMalloc");
    detail::check_hip(hipMemset(st->data, 0, bytes), "hipMemset");
  }

  static bool on_device(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
  }

 public:
  filter(const blocked_bloomfilter_config& config, u64 m) : st(std::make_shared<state>()) {
    st->instance = construct(config, m);
    allocate();
  }
  filter(const cuckoofilter::config& config, u64 m) : st(std::make_shared<state>()) {
    st->instance = std::make_shared<cf>(m, config.bits_per_tag, config.tags_per_bucket);
    allocate();
  }
  /// filter.hpp:170-173: the configuration and size are chosen from (n, tw)
  /// by the learned cost model (tw in ns per qualifying tuple); the time
  /// budget is accepted and, as in the reference, not used.
  filter(u64 n, u64 tw, u64 time_budget_millis = 1000)
      : filter(n, tw, model::cost_model::get_default_instance(), time_budget_millis) {}
  filter(u64 n, u64 tw, const model::cost_model& cost_model, u64 time_budget_millis = 1000)
      : st(std::make_shared<state>()) {
    (void)time_budget_millis;
    const dtl_bbf_choice ch = cost_model.choose(n, tw);
    blocked_bloomfilter_config c;
    c.k = ch.k;
    c.word_size = ch.word_bits / 8;
    c.word_cnt_per_block = ch.w;
    c.sector_cnt = ch.z > 1 ? ch.w : ch.s;
    c.zone_cnt = ch.z;
    st->instance = construct(c, ch.m);
    allocate();
  }

  const filter_base& instance() const { return *st->instance; }
  u64* data() const { return st->data; }

  /// Inserts [begin, end).  Keys may be host or device memory for blocked
  /// Bloom filters; cuckoo filters take host keys (host build).
  u1 insert(const u32* begin, const u32* end) {
    const u32 n = static_cast<u32>(end - begin);
    if (dynamic_cast<cf*>(st->instance.get()) || on_device(begin))
      return st->instance->batch_insert(st->data, begin, n);
    u32* d = nullptr;
    detail::check_hip(hipMalloc(&d, std::max<std::size_t>(n, 1) * 4), "hipMalloc");
    detail::check_hip(hipMemcpy(d, begin, std::size_t(n) * 4, hipMemcpyHostToDevice), "hipMemcpy");
    const u1 ok = st->instance->batch_insert(st->data, d, n);
    (void)hipFree(d);
    return ok;
  }

  class probe_t {
    std::shared_ptr<state> st;

   public:
    explicit probe_t(std::shared_ptr<state> s) : st(std::move(s)) {}
    /// Probes [begin, end) and writes ascending positions (+ match_offset)
    /// to match_pos; both may be host or device memory.
    u64 operator()(const u32* begin, const u32* end, u32* match_pos, u32 match_offset = 0) const {
      const u32 n = static_cast<u32>(end - begin);
      const bool dk = on_device(begin), dp = on_device(match_pos);
      const u32* keys = begin;
      u32* pos = match_pos;
      u32 *tk = nullptr, *tp = nullptr;
      if (!dk) {
        detail::check_hip(hipMalloc(&tk, std::max<std::size_t>(n, 1) * 4), "hipMalloc");
        detail::check_hip(hipMemcpy(tk, begin, std::size_t(n) * 4, hipMemcpyHostToDevice), "hipMemcpy");
        keys = tk;
      }
      if (!dp) {
        detail::check_hip(hipMalloc(&tp, std::max<std::size_t>(n, 1) * 4), "hipMalloc");
        pos = tp;
      }
      const u64 c = st->instance->batch_contains(st->data, keys, n, pos, match_offset);
      if (!dp) detail::check_hip(hipMemcpy(match_pos, tp, c * 4, hipMemcpyDeviceToHost), "hipMemcpy");
      if (tk) (void)hipFree(tk);
      if (tp) (void)hipFree(tp);
      return c;
    }
  };

  probe_t probe() const { return probe_t(st); }
};

}  // namespace filter
}  // namespace dtl
