// dtl_capi.cpp -- the C ABI of libdtl_hip (declared in include/dtl_hip.h).
//
// Host-side sizing and validation mirror the reference's constructors; all
// filter arithmetic runs in the gfx950 kernels.  There is no CPU fallback:
// without a usable GPU every entry point that would launch work returns
// DTL_ENODEVICE / DTL_EHIP.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <vector>

#include "dtl_hip.h"
#include "dtl_internal.hpp"

namespace dtlh {

// ---------------------------------------------------------------------------
// registry
// ---------------------------------------------------------------------------
const std::vector<BbfEntry>& bbf_registry() {
  static const std::vector<BbfEntry> reg = [] {
    std::vector<BbfEntry> v;
    dtl_reg_u32_w1(v);
    dtl_reg_u32_w2(v);
    dtl_reg_u32_w4(v);
    dtl_reg_u32_w8(v);
    dtl_reg_u32_w16(v);
    dtl_reg_u64_w1(v);
    dtl_reg_u64_w2(v);
    dtl_reg_u64_w4(v);
    dtl_reg_u64_w8(v);
    dtl_reg_u64_w16(v);
    dtl_reg_z32_w4(v);
    dtl_reg_z32_w8(v);
    dtl_reg_z32_w16(v);
    dtl_reg_z64_w4(v);
    dtl_reg_z64_w8(v);
    dtl_reg_z64_w16(v);
    return v;
  }();
  return reg;
}

const BbfEntry* bbf_find(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k) {
  for (const BbfEntry& e : bbf_registry())
    if (e.word_bits == word_bits && e.w == w && e.s == s && e.z == z && e.k == k) return &e;
  return nullptr;
}

// ---------------------------------------------------------------------------
// sizing helpers (math.hpp:13-46, div.hpp:29-57, block_addressing_logic.hpp)
// ---------------------------------------------------------------------------
static bool is_pow2(uint64_t x) { return x == 1 ? true : (x && ((x & (x - 1)) == 0)); }
static uint64_t next_pow2(uint64_t v) { return 1ull << (64 - __builtin_clzll(v - 1)); }
static uint32_t floor_log2(uint64_t v) { return 63u - static_cast<uint32_t>(__builtin_clzll(v)); }

struct Magic {
  uint32_t magic, shift, divisor;
};

// Smallest d >= max(n, 2), d not a power of two, whose round-up magic number
// needs no add step (libdivide's "algorithm 1": e = d - 2^(32+L) mod d < 2^L).
static bool cheap_magic(uint32_t n, Magic* out) {
  const uint32_t d_max = 0xFFFFFFFFu - 2u;
  for (uint32_t d = std::max(n, 2u); d <= d_max; ++d) {
    if ((d & (d - 1)) == 0) continue;
    const uint32_t L = floor_log2(d);
    const unsigned __int128 p = static_cast<unsigned __int128>(1) << (32 + L);
    const uint64_t q = static_cast<uint64_t>(p / d);
    const uint64_t r = static_cast<uint64_t>(p % d);
    if (d - r < (1ull << L)) {
      out->magic = static_cast<uint32_t>(q + 1);
      out->shift = L;
      out->divisor = d;
      return true;
    }
  }
  return false;
}

struct Addressing {
  uint32_t mode, block_cnt, shift, magic, mshift, mask;
};

static int make_addressing(uint64_t m, uint64_t unit_bits, Addressing* a) {
  const uint64_t max_blocks = 1ull << 31;
  const uint64_t desired = (m + unit_bits - 1) / unit_bits;
  std::memset(a, 0, sizeof(*a));
  a->mode = is_pow2(m) ? DTL_ADDR_POW2 : DTL_ADDR_MAGIC;
  if (desired > max_blocks) return DTL_ELOGIC;
  if (desired < 2) return DTL_EINVAL;  // a single block is undefined in the reference
  if (a->mode == DTL_ADDR_POW2) {
    const uint64_t bc = next_pow2(desired);
    if (next_pow2(bc) > max_blocks) return DTL_ELOGIC;
    a->block_cnt = static_cast<uint32_t>(bc);
    a->shift = 32u - floor_log2(bc);
    a->mask = static_cast<uint32_t>(bc - 1);
  } else {
    Magic mg;
    if (!cheap_magic(static_cast<uint32_t>(desired), &mg)) return DTL_ELOGIC;
    if (next_pow2(mg.divisor) > max_blocks) return DTL_ELOGIC;
    a->block_cnt = mg.divisor;
    a->magic = mg.magic;
    a->mshift = mg.shift;
    a->mask = static_cast<uint32_t>(next_pow2(mg.divisor) - 1);
  }
  return DTL_OK;
}

// ---------------------------------------------------------------------------
// device info + per-stream scratch
// ---------------------------------------------------------------------------
struct DeviceInfo {
  int cu_count = 0;
  uint32_t max_lds = 0;
};

static int current_device_info(DeviceInfo* out, int* dev_out) {
  static std::mutex mu;
  static std::map<int, DeviceInfo> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DTL_ENODEVICE;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(dev);
  if (it == cache.end()) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return DTL_ENODEVICE;
    DeviceInfo di;
    di.cu_count = prop.multiProcessorCount;
    size_t lds = prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor > lds) lds = prop.maxSharedMemoryPerMultiProcessor;
    di.max_lds = static_cast<uint32_t>(std::min<size_t>(lds, 160 * 1024));
    it = cache.emplace(dev, di).first;
  }
  *out = it->second;
  if (dev_out) *dev_out = dev;
  return DTL_OK;
}

int current_device(int* dev, int* cu_count, uint32_t* max_lds) {
  DeviceInfo di;
  const int rc = current_device_info(&di, dev);
  if (rc != DTL_OK) return rc;
  if (cu_count) *cu_count = di.cu_count;
  if (max_lds) *max_lds = di.max_lds;
  return DTL_OK;
}

struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  // set once a call on this stream was captured into a graph: the graph
  // holds ptr, so the buffer may neither move nor be freed from then on
  bool captured = false;
};

static std::mutex g_scratch_mu;
static std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;

static bool stream_capturing(hipStream_t stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Scratch device memory per (device, stream), grown on demand.  Growth
// synchronizes the stream before freeing the old buffer.  A call captured
// into a graph pins the buffer: the graph's kernels hold its address, so a
// later call that needs more than it holds fails with DTL_ERUNTIME (before
// any work is enqueued) instead of freeing memory a replay would use, until
// dtl_scratch_release() is called for the stream.  Capturing therefore
// needs a warm-up call of at least the same size on the stream.
static int get_scratch(int dev, hipStream_t stream, size_t bytes, void** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  Scratch& s = g_scratch[{dev, stream}];
  const bool capturing = stream_capturing(stream);
  if (s.bytes < bytes) {
    if (capturing || s.captured) return DTL_ERUNTIME;
    if (s.ptr) {
      if (hipStreamSynchronize(stream) != hipSuccess) return DTL_EHIP;
      (void)hipFree(s.ptr);
      s.ptr = nullptr;
      s.bytes = 0;
    }
    const size_t want = std::max<size_t>(bytes + bytes / 4, 1 << 20);
    if (hipMalloc(&s.ptr, want) != hipSuccess) {
      s.ptr = nullptr;
      return DTL_ENOMEM;
    }
    s.bytes = want;
  }
  if (capturing) s.captured = true;
  *out = s.ptr;
  return DTL_OK;
}

// Frees the scratch of (current device, stream) and clears its pin.
static int release_scratch(hipStream_t stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return DTL_ENODEVICE;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch.find({dev, stream});
  if (it == g_scratch.end()) return DTL_OK;
  if (stream_capturing(stream)) return DTL_ERUNTIME;
  if (it->second.ptr) {
    if (hipStreamSynchronize(stream) != hipSuccess) return DTL_EHIP;
    (void)hipFree(it->second.ptr);
  }
  g_scratch.erase(it);
  return DTL_OK;
}

static size_t scratch_bytes(hipStream_t stream, int* captured) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch.find({dev, stream});
  if (it == g_scratch.end()) {
    if (captured) *captured = 0;
    return 0;
  }
  if (captured) *captured = it->second.captured ? 1 : 0;
  return it->second.bytes;
}

static thread_local char g_last_error[256] = "";

void set_last_error(hipError_t e) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s (%d)", hipGetErrorString(e), static_cast<int>(e));
}

void set_last_error_text(const char* msg) { std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg); }

static int hip_status(hipError_t e) {
  if (e == hipSuccess) return DTL_OK;
  set_last_error(e);
  return DTL_EHIP;
}

// Errors of earlier, unrelated HIP calls stay in the runtime's last-error
// slot; clear it so that a launch check reports only this call's failures.
static void clear_stale_error() { (void)hipGetLastError(); }

static void fill_addr(const dtl_bbf_desc* d, uint32_t* magic, uint32_t* mshift, uint32_t* divisor,
                      uint32_t* post) {
  if (d->addr_mode == DTL_ADDR_POW2) {
    *magic = 0;
    *mshift = 0;
    *divisor = 0;
    *post = d->block_shift;
  } else {
    *magic = d->magic;
    *mshift = d->magic_shift;
    *divisor = d->block_cnt;
    *post = 0;
  }
}

static bool desc_ok(const dtl_bbf_desc* d) {
  return d && d->block_cnt >= 2 && d->word_cnt == static_cast<uint64_t>(d->block_cnt) * d->w &&
         (d->word_bits == 32 || d->word_bits == 64);
}

static uint32_t block_bytes(const dtl_bbf_desc* d) { return d->w * d->word_bits / 8; }

// Default choice of the partitioned probe (Bloom and cuckoo filters): the
// batch must be large enough for the partition passes to pay for
// themselves, and the smaller the filter, the larger the batch it takes.
// Measured on MI355X with fixed-capacity bins (u64 w8 s8 k8 Bloom filter;
// 16-bit / 4-slot cuckoo table at 95% load; 1% hits), partitioned vs direct
// G probes/s:
//   Bloom  8 MiB: 2^24 82 / 94, 2^25 105 / 101;  16 MiB: 2^23 61 / 62,
//          2^24 82 / 67;  32 MiB: 2^22 41 / 48, 2^23 62 / 53;
//          64 MiB: 2^22 40 / 43, 2^23 61 / 50;
//   cuckoo 8 MiB: 2^24 47 / 52, 2^25 55 / 54;  16 MiB: 2^22 26 / 32,
//          2^23 38 / 34;  32 MiB: 2^22 25 / 27, 2^23 37 / 29;
//          64 MiB: 2^22 24 / 25, 2^23 36 / 26.
static bool partition_pays(uint64_t filter_bytes, uint64_t n, bool cuckoo) {
  if (filter_bytes >= (32ull << 20)) return n >= (1ull << 23);
  if (filter_bytes >= (16ull << 20)) return n >= (cuckoo ? (1ull << 23) : (1ull << 24));
  if (filter_bytes >= (8ull << 20)) return n >= (1ull << 25);
  return false;
}

// Capacity of each (partition, lane) bin of the partitioned probe, or 0 to
// size the partitions with a histogram pass.  A full slice of
// `slice_units` blocks (buckets) out of `total_units` receives n *
// slice_units / total_units keys of a uniform batch, an eighth of them per
// lane; the capacity is that + 3% + 256 keys (a bin's count has a standard
// deviation of sqrt(expected), 0.4% at the 2^28-key benchmark), in whole
// tiles.  The last slice may be partial, so the bins of the full ones must
// cover their own share, not n / bins.  Bins are used where they hold 4
// tiles or more and the grouped order stays under 2^31 slots (the scatter
// marks spilled keys with the top bit).
static uint32_t bin_capacity(uint64_t n, uint32_t nparts, uint64_t slice_units, uint64_t total_units,
                             uint32_t flags) {
  if ((flags & DTL_PROBE_PART_HISTOGRAM) || n >= (1ull << 30) || total_units == 0) return 0;
  const double share = static_cast<double>(n) * static_cast<double>(std::min(slice_units, total_units)) /
                       static_cast<double>(total_units) / kPartLanes;
  if (share < 4.0 * kTileKeys) return 0;
  const uint64_t want = static_cast<uint64_t>(share * 1.03) + 257;
  const uint64_t cap = (want + kTileKeys - 1) / kTileKeys * kTileKeys;
  if (cap * nparts * kPartLanes >= (1ull << 31)) return 0;
  return static_cast<uint32_t>(cap);
}

static bool aligned_to(const void* p, uint32_t a) {
  return (reinterpret_cast<uintptr_t>(p) % a) == 0;
}

}  // namespace dtlh

using namespace dtlh;

extern "C" {

int dtl_version(void) { return 101; }

int dtl_scratch_release(dtl_stream_t stream) {
  return release_scratch(static_cast<hipStream_t>(stream));
}

uint64_t dtl_scratch_bytes(dtl_stream_t stream, int* captured) {
  return scratch_bytes(static_cast<hipStream_t>(stream), captured);
}

const char* dtl_last_error(void) { return g_last_error; }

const char* dtl_status_string(int s) {
  switch (s) {
    case DTL_OK: return "ok";
    case DTL_EINVAL: return "invalid argument";
    case DTL_ERUNTIME: return "runtime error";
    case DTL_ELOGIC: return "the block count must not exceed 2^31";
    case DTL_ENOMEM: return "out of device memory";
    case DTL_EHIP: return "HIP runtime error";
    case DTL_ENODEVICE: return "no usable GPU";
  }
  return "unknown status";
}

int dtl_bbf_supported(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k) {
  if (z > 1) s = w;
  return bbf_find(word_bits, w, s, z > 1 ? z : 1, k) != nullptr;
}

int dtl_bbf_config_count(void) { return static_cast<int>(bbf_registry().size()); }

int dtl_bbf_split_lanes(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k) {
  if (z > 1) s = w;
  const BbfEntry* e = bbf_find(word_bits, w, s, z > 1 ? z : 1, k);
  return e ? static_cast<int>(e->split_lanes) : 0;
}

int dtl_bbf_desc_init(dtl_bbf_desc* d, uint32_t word_bits, uint64_t m, uint32_t k, uint32_t w,
                      uint32_t s, uint32_t z) {
  if (!d) return DTL_EINVAL;
  std::memset(d, 0, sizeof(*d));
  if (word_bits != 32 && word_bits != 64) return DTL_ERUNTIME;  // filter.hpp:64
  const uint32_t word_bytes = word_bits / 8;
  uint32_t s_eff = 0, k_eff = 0;
  if (z <= 1) {
    // blocked_bloomfilter.hpp:169-229: unsupported values throw, others are
    // clamped and the mismatch is reported after construction.
    if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return DTL_EINVAL;
    if (s != 1 && s != 2 && s != 4 && s != 8 && s != 16) return DTL_EINVAL;
    s_eff = (s == 1) ? 1u : std::min(w * word_bytes, s);
    if (k < 1 || k > 16) return DTL_EINVAL;
    if (k == 1) k_eff = std::max(1u, s_eff);
    else if (k == 16) k_eff = std::max(16u, s_eff);
    else k_eff = std::max((k % s_eff == 0) ? k : 1u, s_eff);
    z = 1;
  } else {
    // zoned_blocked_bloomfilter.hpp:102-215
    if (w <= z) return DTL_EINVAL;
    if (w != 4 && w != 8 && w != 16) return DTL_EINVAL;
    if (z != 2 && z != 4 && z != 8) return DTL_EINVAL;
    z = std::min(w / 2, z);
    if (k < 2 || k > 16 || (k & 1u)) return DTL_EINVAL;
    k_eff = (k == 16) ? std::max(16u, z) : std::max((k % z == 0) ? k : 2u, z);
    s = s_eff = w;
  }
  Addressing a;
  const int rc = make_addressing(m, static_cast<uint64_t>(word_bits) * w, &a);
  if (rc != DTL_OK) return rc;
  if (k_eff != k || s_eff != s) return DTL_EINVAL;
  d->word_bits = word_bits;
  d->w = w;
  d->s = s_eff;
  d->z = z;
  d->k = k_eff;
  d->addr_mode = a.mode;
  d->block_cnt = a.block_cnt;
  d->block_shift = a.shift;
  d->magic = a.magic;
  d->magic_shift = a.mshift;
  d->m = m;
  d->m_actual = static_cast<uint64_t>(a.block_cnt) * w * word_bits;
  d->word_cnt = static_cast<uint64_t>(a.block_cnt) * w;
  d->size_bytes = d->m_actual / 8;
  return DTL_OK;
}

int dtl_cf_desc_init(dtl_cf_desc* d, uint64_t m, uint32_t tag_bits, uint32_t tags_per_bucket) {
  if (!d) return DTL_EINVAL;
  std::memset(d, 0, sizeof(*d));
  // cuckoofilter.hpp:184-206
  if (tag_bits != 4 && tag_bits != 8 && tag_bits != 12 && tag_bits != 16 && tag_bits != 32)
    return DTL_EINVAL;
  if (tags_per_bucket != 1 && tags_per_bucket != 2 && tags_per_bucket != 4) return DTL_EINVAL;
  const uint32_t bpb = (tag_bits * tags_per_bucket + 7) / 8;  // cuckoofilter_table.hpp:89
  Addressing a;
  const int rc = make_addressing(m, static_cast<uint64_t>(bpb) * 8, &a);
  if (rc != DTL_OK) return rc;
  d->tag_bits = tag_bits;
  d->tags_per_bucket = tags_per_bucket;
  d->bytes_per_bucket = bpb;
  d->addr_mode = a.mode;
  d->bucket_cnt = a.block_cnt;
  d->bucket_shift = a.shift;
  d->bucket_mask = a.mask;
  d->magic = a.magic;
  d->magic_shift = a.mshift;
  d->m = m;
  d->m_actual = static_cast<uint64_t>(bpb) * a.block_cnt * 8;
  d->table_words = (static_cast<uint64_t>(bpb) * a.block_cnt + 3) / 4;  // cuckoofilter_table.hpp:245-249
  d->victim_offset = d->table_words * 4;
  d->buffer_bytes = d->victim_offset + 16;
  return DTL_OK;
}

int bbf_build(const dtl_bbf_desc* d, void* d_words, const uint32_t* d_keys, uint64_t n,
              dtl_stream_t stream) {
  if (!desc_ok(d) || !d_words || (n && !d_keys)) return DTL_EINVAL;
  if (!aligned_to(d_words, std::min(16u, block_bytes(d)))) return DTL_EINVAL;
  const BbfEntry* e = bbf_find(d->word_bits, d->w, d->s, d->z, d->k);
  if (!e) return DTL_EINVAL;
  if (n == 0) return DTL_OK;
  clear_stale_error();
  DeviceInfo di;
  int rc = current_device_info(&di, nullptr);
  if (rc != DTL_OK) return rc;
  BbfBuildLaunch L;
  L.words = d_words;
  L.keys = d_keys;
  L.n = n;
  fill_addr(d, &L.addr_magic, &L.addr_mshift, &L.addr_divisor, &L.addr_post);
  L.cu_count = di.cu_count;
  L.stream = static_cast<hipStream_t>(stream);
  return hip_status(e->build(L));
}

int bbf_probe_ex(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
                 uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count,
                 const dtl_probe_params* params, dtl_stream_t stream) {
  if (!desc_ok(d) || !d_words || !d_count || n > 0xFFFFFFFFull) return DTL_EINVAL;
  if (n && (!d_keys || !d_pos)) return DTL_EINVAL;
  if (!aligned_to(d_words, std::min(16u, block_bytes(d)))) return DTL_EINVAL;
  const BbfEntry* e = bbf_find(d->word_bits, d->w, d->s, d->z, d->k);
  if (!e) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  clear_stale_error();
  DeviceInfo di;
  int dev = 0;
  int rc = current_device_info(&di, &dev);
  if (rc != DTL_OK) return rc;
  if (n == 0) return hip_status(write_zero_count(d_count, st));
  const uint32_t ntiles = static_cast<uint32_t>((n + kTileKeys - 1) / kTileKeys);
  const size_t bits_bytes = static_cast<size_t>(ntiles) * kTileWords * 8;
  const size_t cnt_bytes = (static_cast<size_t>(ntiles) + kScanTmpWords) * 4;
  const uint32_t flags = params ? params->flags : 0u;
  uint32_t lds_max = di.max_lds - 2048;  // room for the kernel's own shared arrays
  if (params && params->lds_max_bytes) lds_max = std::min(lds_max, params->lds_max_bytes);
  const bool use_lds = !(flags & DTL_PROBE_FORCE_GLOBAL) && d->size_bytes <= lds_max;
  // Lane split pays where blocks come from L2 / Infinity Cache (1.33x at
  // 4 MiB, and in the partitioned probe); a direct probe of a filter beyond
  // the Infinity Cache is bound by HBM line fills, where one lane per key is
  // ~1% faster (measured on MI355X).
  int split = (flags & DTL_PROBE_LANE_SPLIT) ? 1 : (flags & DTL_PROBE_ONE_LANE) ? 0 : 1;
  if (d->size_bytes > 0xFFFFFFFFull) split = 0;  // the lane-split kernel addresses with 32-bit offsets

  // partitioned probe: slices of 2^slice_log2 bytes (at most kPartMax of them)
  uint32_t slice_shift = 0, nparts = 0;
  if (!use_lds && !(flags & DTL_PROBE_DIRECT) && e->probe_part) {
    const bool want = (flags & DTL_PROBE_PARTITIONED) ||
                      partition_pays(d->size_bytes, n, false);
    // slices of 512 KiB, 2 MiB from 64 MiB on (measured on MI355X, u64 w8 s8
    // k8, 2^28 keys, G probes/s for 512 KiB / 1 MiB / 2 MiB slices: 64 MiB
    // 115 / 120 / 120, 128 MiB 109 / 116 / 118, 256 MiB 104 / 110 / 114;
    // below 64 MiB the three are within 2%)
    const uint32_t slice_log2 = (params && params->slice_log2) ? params->slice_log2
                                : d->size_bytes >= (64ull << 20) ? 21u : 19u;
    const uint32_t bb_log2 = floor_log2(block_bytes(d));
    slice_shift = slice_log2 > bb_log2 ? slice_log2 - bb_log2 : 0u;
    while (slice_shift < 31 && ((static_cast<uint64_t>(d->block_cnt) + (1ull << slice_shift) - 1) >> slice_shift) > kPartMax)
      ++slice_shift;
    nparts = static_cast<uint32_t>((static_cast<uint64_t>(d->block_cnt) + (1ull << slice_shift) - 1) >> slice_shift);
    if (!want || nparts < 2) nparts = 0;
  }
  if (nparts) {
    const uint32_t bucket_shift = hit_bucket_shift(ntiles);
    const uint32_t bucket_tiles = (1u << bucket_shift) / kTileKeys;
    const uint32_t nbuckets = (ntiles + bucket_tiles - 1) / bucket_tiles;
    const uint32_t hit_groups = std::min<uint32_t>(kHitGroupsMax, 2u * static_cast<uint32_t>(std::max(di.cu_count, 1)));
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t cnt_r = up(cnt_bytes);
    const size_t nbins = static_cast<size_t>(nparts) * kPartLanes;
    // bins of fixed capacity (no histogram pass), see bin_capacity; keys past
    // a full bin go to other bins
    const uint32_t bin_cap = bin_capacity(n, nparts, 1ull << slice_shift, d->block_cnt, flags);
    const uint32_t ptiles = bin_cap ? static_cast<uint32_t>(nbins * bin_cap / kTileKeys) : ntiles;
    const size_t pbits_bytes = static_cast<size_t>(ptiles) * kTileWords * 8;
    const size_t pcnt_r = up((static_cast<size_t>(ptiles) + kScanTmpWords) * 4);
    const size_t part_bytes = up((nbins * 3 + 1) * 4);
    const size_t bucket_bytes = up(8 + (static_cast<size_t>(nbuckets) * 3 + 1) * 4);
    const size_t matrix_bytes = up(static_cast<size_t>(hit_groups) * nbuckets * 4);
    const size_t pair_bytes = up(static_cast<size_t>(ptiles) * kTileKeys * 4);
    void* scratch = nullptr;
    rc = get_scratch(dev, st, bits_bytes + cnt_r + pbits_bytes + pcnt_r + part_bytes + bucket_bytes + matrix_bytes +
                                  2 * pair_bytes,
                     &scratch);
    if (rc != DTL_OK) return rc;
    char* sp = static_cast<char*>(scratch);
    BbfPartLaunch P{};
    P.words = d_words;
    P.keys = d_keys;
    P.n = static_cast<uint32_t>(n);
    fill_addr(d, &P.addr_magic, &P.addr_mshift, &P.addr_divisor, &P.addr_post);
    P.slice_shift = slice_shift;
    P.nparts = nparts;
    P.scatter_threads = (params && params->block_threads) ? params->block_threads : 1024u;
    P.bits = reinterpret_cast<uint64_t*>(sp);
    sp += bits_bytes;
    P.tile_cnt = reinterpret_cast<uint32_t*>(sp);
    sp += cnt_r;
    P.pbits = reinterpret_cast<uint64_t*>(sp);
    sp += pbits_bytes;
    P.ptile_cnt = reinterpret_cast<uint32_t*>(sp);
    sp += pcnt_r;
    P.counts = reinterpret_cast<uint32_t*>(sp);
    P.offsets = P.counts + nbins;
    P.cursor = P.offsets + nbins + 1;
    sp += part_bytes;
    P.bin_cap = bin_cap;
    P.ptiles = ptiles;
    P.d_hits = reinterpret_cast<uint64_t*>(sp);
    P.bucket_shift = bucket_shift;
    P.nbuckets = nbuckets;
    P.hit_groups = hit_groups;
    P.bcounts = reinterpret_cast<uint32_t*>(sp + 8);
    P.boffsets = P.bcounts + nbuckets;
    P.bcursor = P.boffsets + nbuckets + 1;
    sp += bucket_bytes;
    P.hcounts = reinterpret_cast<uint32_t*>(sp);
    sp += matrix_bytes;
    P.keys_p = reinterpret_cast<uint32_t*>(sp);
    P.pos_p = reinterpret_cast<uint32_t*>(sp + pair_bytes);
    P.ntiles = ntiles;
    P.split = split;
    P.cu_count = di.cu_count;
    P.stream = st;
    if (params && params->ev_start &&
        hipEventRecord(static_cast<hipEvent_t>(params->ev_start), st) != hipSuccess)
      return DTL_EHIP;
    // a key matches at most once: the hit sort writes the ascending
    // positions and the count itself (no key-order bitmap, no compaction)
    P.out_pos = d_pos;
    P.match_offset = match_offset;
    P.d_count_out = d_count;
    hipError_t he = partition_keys(P);
    if (he == hipSuccess) he = e->probe_part(P);
    if (he == hipSuccess) he = part_reorder_matches(P);
    if (he != hipSuccess) return hip_status(he);
    if (params && params->ev_stop &&
        hipEventRecord(static_cast<hipEvent_t>(params->ev_stop), st) != hipSuccess)
      return DTL_EHIP;
    return DTL_OK;
  }

  uint32_t threads = (params && params->block_threads) ? params->block_threads : 0;
  if (threads == 0) threads = d->size_bytes > 40 * 1024 ? 1024 : 256;
  if (!(flags & (DTL_PROBE_LANE_SPLIT | DTL_PROBE_ONE_LANE)) && d->size_bytes >= (256ull << 20)) split = 0;
  void* scratch = nullptr;
  rc = get_scratch(dev, st, bits_bytes + cnt_bytes, &scratch);
  if (rc != DTL_OK) return rc;
  BbfProbeLaunch L;
  L.words = d_words;
  L.keys = d_keys;
  L.n = static_cast<uint32_t>(n);
  fill_addr(d, &L.addr_magic, &L.addr_mshift, &L.addr_divisor, &L.addr_post);
  L.bits = static_cast<uint64_t*>(scratch);
  L.tile_cnt = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + bits_bytes);
  L.ntiles = ntiles;
  L.filter_bytes = static_cast<uint32_t>(std::min<uint64_t>(d->size_bytes, 0xFFFFFFFFull));
  L.use_lds = use_lds;
  L.lds_threads = threads >= 1024 ? 1024 : 256;
  L.lds_grid = params ? params->grid_blocks : 0;
  L.split = split;
  L.cu_count = di.cu_count;
  L.stream = st;
  if (params && params->ev_start &&
      hipEventRecord(static_cast<hipEvent_t>(params->ev_start), st) != hipSuccess)
    return DTL_EHIP;
  hipError_t he = e->probe(L);
  if (he != hipSuccess) return hip_status(he);
  if (params && params->ev_stop &&
      hipEventRecord(static_cast<hipEvent_t>(params->ev_stop), st) != hipSuccess)
    return DTL_EHIP;
  return hip_status(compact_single(L.bits, L.tile_cnt, ntiles, L.n, match_offset, d_pos, d_count, st));
}

int bbf_probe(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
              uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream) {
  return bbf_probe_ex(d, d_words, d_keys, n, match_offset, d_pos, d_count, nullptr, stream);
}

int cf_probe(const dtl_cf_desc* d, const void* d_buf, const uint32_t* d_keys, uint64_t n,
             uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream) {
  return cf_probe_ex(d, d_buf, d_keys, n, match_offset, d_pos, d_count, nullptr, stream);
}

int cf_build(const dtl_cf_desc* d, void* d_buf, const uint32_t* d_keys, uint64_t n, uint32_t seed,
             uint32_t max_kicks, uint64_t* d_failed, dtl_stream_t stream) {
  if (!d || !d_buf || !d_failed || d->bucket_cnt < 2 || (n && !d_keys)) return DTL_EINVAL;
  if (!aligned_to(d_buf, 16) || !aligned_to(d_keys, 4) || !aligned_to(d_failed, 8)) return DTL_EINVAL;
  if (max_kicks > (1u << 20)) return DTL_EINVAL;
  const CfEntry* e = cf_find(d->tag_bits, d->tags_per_bucket);
  if (!e || !e->build) return DTL_EINVAL;
  if (n == 0) return DTL_OK;
  clear_stale_error();
  DeviceInfo di;
  const int rc = current_device_info(&di, nullptr);
  if (rc != DTL_OK) return rc;
  CfBuildLaunch L;
  L.buf = d_buf;
  L.keys = d_keys;
  L.n = n;
  L.addr_mode = d->addr_mode;
  L.shift = d->bucket_shift;
  L.bucket_mask = d->bucket_mask;
  L.magic = d->magic;
  L.mshift = d->magic_shift;
  L.bucket_cnt = d->bucket_cnt;
  L.victim_offset = d->victim_offset;
  L.max_kicks = max_kicks ? max_kicks : 500u;  // max_relocation_count, cuckoofilter_logic.hpp:122
  L.seed = seed;
  L.failed = d_failed;
  L.cu_count = di.cu_count;
  L.stream = static_cast<hipStream_t>(stream);
  return hip_status(e->build(L));
}

int cf_probe_ex(const dtl_cf_desc* d, const void* d_buf, const uint32_t* d_keys, uint64_t n,
                uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, const dtl_probe_params* params,
                dtl_stream_t stream) {
  if (!d || !d_buf || !d_count || n > 0xFFFFFFFFull || d->bucket_cnt < 2) return DTL_EINVAL;
  if (n && (!d_keys || !d_pos)) return DTL_EINVAL;
  if (!aligned_to(d_buf, 16)) return DTL_EINVAL;
  const CfEntry* e = cf_find(d->tag_bits, d->tags_per_bucket);
  if (!e) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  clear_stale_error();
  DeviceInfo di;
  int dev = 0;
  int rc = current_device_info(&di, &dev);
  if (rc != DTL_OK) return rc;
  if (n == 0) return hip_status(write_zero_count(d_count, st));
  const uint32_t ntiles = static_cast<uint32_t>((n + kTileKeys - 1) / kTileKeys);
  const size_t bits_bytes = static_cast<size_t>(ntiles) * kTileWords * 8;
  const size_t cnt_bytes = (static_cast<size_t>(ntiles) + kScanTmpWords) * 4;
  const uint32_t flags = params ? params->flags : 0u;
  auto ev = [&](void* evp) {
    return !evp || hipEventRecord(static_cast<hipEvent_t>(evp), st) == hipSuccess;
  };

  // partitioned probe: two passes over keys grouped by bucket slice
  const uint64_t table_bytes = d->victim_offset;
  const bool want = !(flags & DTL_PROBE_DIRECT) &&
                    ((flags & DTL_PROBE_PARTITIONED) || partition_pays(table_bytes, n, true));
  uint32_t slice_shift = 0, nparts = 0;
  if (want) {
    // 2 MiB slices from 64 MiB on, as for Bloom filters (16-bit / 4-slot
    // table, 2^28 keys, 512 KiB / 1 MiB / 2 MiB: 64 MiB 56.6 / 57.3 / 57.3,
    // 128 MiB 53.6 / 55.8 / 56.4, 512 MiB 44.5 / 48.5 / 51.3 G probes/s)
    const uint32_t slice_log2 = (params && params->slice_log2) ? params->slice_log2
                                : table_bytes >= (64ull << 20) ? 21u : 19u;
    const uint32_t bb_log2 = floor_log2(d->bytes_per_bucket);
    slice_shift = slice_log2 > bb_log2 ? slice_log2 - bb_log2 : 0u;
    // buckets 0 .. bucket_cnt: a MAGIC alternate index can equal bucket_cnt (it reads the victim bytes)
    const uint64_t nb = static_cast<uint64_t>(d->bucket_cnt) + 1;
    while (slice_shift < 31 && ((nb + (1ull << slice_shift) - 1) >> slice_shift) > kPartMax) ++slice_shift;
    nparts = static_cast<uint32_t>((nb + (1ull << slice_shift) - 1) >> slice_shift);
    if (nparts < 2) nparts = 0;
  }
  if (nparts) {
    const uint32_t bucket_shift = hit_bucket_shift(ntiles);
    const uint32_t bucket_tiles = (1u << bucket_shift) / kTileKeys;
    const uint32_t nbuckets = (ntiles + bucket_tiles - 1) / bucket_tiles;
    const uint32_t hit_groups = std::min<uint32_t>(kHitGroupsMax, 2u * static_cast<uint32_t>(std::max(di.cu_count, 1)));
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t cnt_r = up(cnt_bytes);
    const size_t nbins = static_cast<size_t>(nparts) * kPartLanes;
    // fixed-capacity bins as for Bloom filters (no histogram pass), see
    // bin_capacity; DTL_PROBE_PART_HISTOGRAM keeps the histograms
    const uint32_t bin_cap = bin_capacity(n, nparts, 1ull << slice_shift, d->bucket_cnt, flags);
    const uint32_t ptiles = bin_cap ? static_cast<uint32_t>(nbins * bin_cap / kTileKeys) : ntiles;
    const size_t slots = static_cast<size_t>(ptiles) * kTileKeys;
    const size_t pbits_bytes = static_cast<size_t>(ptiles) * kTileWords * 8;
    const size_t pcnt_r = up((static_cast<size_t>(ptiles) + kScanTmpWords) * 4);
    const size_t part_bytes = up((nbins * 4 + 1) * 4);  // counts of both passes, offsets, cursors
    const size_t bucket_bytes = up(8 + (static_cast<size_t>(nbuckets) * 3 + 1) * 4);
    const size_t matrix_bytes = up(static_cast<size_t>(hit_groups) * nbuckets * 4);
    const size_t key_bytes = up(slots * 4);
    // a key can match in both passes: match positions take up to 2n words
    const size_t hit_bytes = up(static_cast<size_t>(n) * 8);
    const size_t ppos_bytes = std::max(hit_bytes, up(slots * 4));  // grouped positions, then bucketed matches
    void* scratch = nullptr;
    rc = get_scratch(dev, st, bits_bytes + cnt_r + pbits_bytes + pcnt_r + part_bytes + bucket_bytes + 256 +
                                  matrix_bytes + key_bytes + ppos_bytes + hit_bytes, &scratch);
    if (rc != DTL_OK) return rc;
    char* sp = static_cast<char*>(scratch);
    BbfPartLaunch P{};
    P.words = d_buf;
    P.keys = d_keys;
    P.n = static_cast<uint32_t>(n);
    if (d->addr_mode == DTL_ADDR_POW2) {
      P.addr_post = d->bucket_shift;
    } else {
      P.addr_magic = d->magic;
      P.addr_mshift = d->magic_shift;
      P.addr_divisor = d->bucket_cnt;
    }
    P.slice_shift = slice_shift;
    P.nparts = nparts;
    P.tag_mask = static_cast<uint32_t>((1ull << d->tag_bits) - 1ull);
    P.cf_mode = d->addr_mode;
    P.cf_bucket_mask = d->bucket_mask;
    P.cf_bucket_cnt = d->bucket_cnt;
    P.cf_victim_offset = d->victim_offset;
    P.scatter_threads = (params && params->block_threads) ? params->block_threads : 1024u;
    P.bits = reinterpret_cast<uint64_t*>(sp);
    sp += bits_bytes;
    P.tile_cnt = reinterpret_cast<uint32_t*>(sp);
    sp += cnt_r;
    P.pbits = reinterpret_cast<uint64_t*>(sp);
    sp += pbits_bytes;
    P.ptile_cnt = reinterpret_cast<uint32_t*>(sp);
    sp += pcnt_r;
    uint32_t* counts1 = reinterpret_cast<uint32_t*>(sp);
    uint32_t* counts2 = counts1 + nbins;
    P.offsets = counts2 + nbins;
    P.cursor = P.offsets + nbins + 1;
    sp += part_bytes;
    P.d_hits = reinterpret_cast<uint64_t*>(sp);
    P.bucket_shift = bucket_shift;
    P.nbuckets = nbuckets;
    P.hit_groups = hit_groups;
    P.bcounts = reinterpret_cast<uint32_t*>(sp + 8);
    P.boffsets = P.bcounts + nbuckets;
    P.bcursor = P.boffsets + nbuckets + 1;
    sp += bucket_bytes;
    uint64_t* d_h1 = reinterpret_cast<uint64_t*>(sp);
    uint64_t* d_h2 = d_h1 + 1;
    sp += 256;
    P.hcounts = reinterpret_cast<uint32_t*>(sp);
    sp += matrix_bytes;
    P.keys_p = reinterpret_cast<uint32_t*>(sp);
    sp += key_bytes;
    P.pos_p = reinterpret_cast<uint32_t*>(sp);  // grouped positions, later the bucketed matches (<= 2n)
    sp += ppos_bytes;
    uint32_t* hp = reinterpret_cast<uint32_t*>(sp);  // matches of both passes (<= 2n)
    P.ntiles = ntiles;
    P.bin_cap = bin_cap;
    P.ptiles = ptiles;
    P.split = 0;
    P.cu_count = di.cu_count;
    P.stream = st;
    if (!ev(params ? params->ev_start : nullptr)) return DTL_EHIP;
    hipError_t he = hipSuccess;
    // both passes' partition counts from one read of the keys (pass 1), when
    // the two bin arrays fit in a workgroup's LDS
    const bool one_hist = nparts <= 1024;
    // With capacity bins, pass 2 leaves out the keys pass 1 found once they
    // are more than n/32 (a device-side test, no host synchronisation): pass
    // 1's matches are sorted into the key-order bitmap first, pass 2 scatters
    // only keys whose bit is clear, and its matches are ORed into that
    // bitmap.  Below the threshold the early sort's kernels return at once
    // and both passes' matches are sorted together.
    const bool skip2 = bin_cap != 0 && !(flags & DTL_PROBE_CF_PASS2_ALL);
    const uint64_t gate_thr = n / 32;
    for (uint32_t pass = 1; pass <= 2 && he == hipSuccess; ++pass) {
      P.part_kind = pass - 1;
      P.cf_which = pass;
      P.counts = pass == 1 ? counts1 : counts2;
      P.counts_next = (pass == 1 && one_hist) ? counts2 : nullptr;
      P.counts_ready = (pass == 2 && one_hist) ? 1 : 0;
      P.skip_bits = (pass == 2 && skip2) ? P.bits : nullptr;
      P.gate_count = (pass == 2 && skip2) ? d_h1 : nullptr;
      P.gate_thr = gate_thr;
      he = partition_keys(P);
      if (he == hipSuccess) he = e->probe_part(P);
      if (he == hipSuccess) he = part_gather_matches(P, hp, pass == 1 ? nullptr : d_h1, pass == 1 ? d_h1 : d_h2);
      if (he == hipSuccess && pass == 1 && skip2) he = part_sort_matches(P, hp, d_h1, kHitSortEarly, d_h1, gate_thr);
    }
    if (he == hipSuccess) he = add_counts(d_h1, d_h2, P.d_hits, st);
    if (he == hipSuccess)
      he = skip2 ? part_sort_matches(P, hp, P.d_hits, kHitSortRest, d_h1, gate_thr) : part_sort_matches(P, hp, P.d_hits);
    if (he != hipSuccess) return hip_status(he);
    if (!ev(params ? params->ev_stop : nullptr)) return DTL_EHIP;
    return hip_status(compact_single(P.bits, P.tile_cnt, ntiles, P.n, match_offset, d_pos, d_count, st));
  }

  void* scratch = nullptr;
  rc = get_scratch(dev, st, bits_bytes + cnt_bytes, &scratch);
  if (rc != DTL_OK) return rc;
  CfProbeLaunch L;
  L.buf = d_buf;
  L.keys = d_keys;
  L.n = static_cast<uint32_t>(n);
  L.addr_mode = d->addr_mode;
  L.shift = d->bucket_shift;
  L.bucket_mask = d->bucket_mask;
  L.magic = d->magic;
  L.mshift = d->magic_shift;
  L.bucket_cnt = d->bucket_cnt;
  L.victim_offset = d->victim_offset;
  L.bits = static_cast<uint64_t*>(scratch);
  L.tile_cnt = reinterpret_cast<uint32_t*>(static_cast<char*>(scratch) + bits_bytes);
  L.ntiles = ntiles;
  L.stream = st;
  if (!ev(params ? params->ev_start : nullptr)) return DTL_EHIP;
  hipError_t he = e->probe(L);
  if (he != hipSuccess) return hip_status(he);
  if (!ev(params ? params->ev_stop : nullptr)) return DTL_EHIP;
  return hip_status(compact_single(L.bits, L.tile_cnt, ntiles, L.n, match_offset, d_pos, d_count, st));
}

}  // extern "C"
