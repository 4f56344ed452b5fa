// dtl_noalloc.hip -- bbf_probe_noalloc: the blocked Bloom filter probe for
// callers that cannot allocate device memory (code under hipGraph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// It runs bbf_probe's direct path (LDS-staged or read from memory, never
// the partitioned probe, whose scratch is ~8 bytes per key) with the match
// bitmap and tile counts in one file-scope device array, `scratch`, of
// kNoAllocScratchBytes (33.1 MB: room for kNoAllocMaxKeys = 2^28 keys), which
// every GPU the library runs on holds until the process ends.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a host-side
// stand-in, not a device address): on its first call on a device, with that
// device current, the entry point asks the runtime for the array's device
// address with hipGetSymbolAddress and keeps it per device.
//
// Nothing in the scratch must start at zero: the probe kernels write every
// bitmap word and every tile count of the call's tiles before the scan and
// the expansion read them.  Every call on one GPU uses the same array, so
// calls that overlap on one GPU (different streams, or host threads without
// ordering) race; calls ordered on one stream do not.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dtl_hip.h"
#include "dtl_internal.hpp"

namespace dtlh {

constexpr uint64_t kNoAllocMaxKeys = 1ull << 28;
constexpr uint32_t kNoAllocMaxTiles = static_cast<uint32_t>(kNoAllocMaxKeys / kTileKeys);
constexpr size_t kNoAllocBitsBytes = static_cast<size_t>(kNoAllocMaxTiles) * kTileWords * 8;
constexpr size_t kNoAllocCntBytes = (static_cast<size_t>(kNoAllocMaxTiles) + kScanTmpWords) * 4;
constexpr size_t kNoAllocScratchBytes = kNoAllocBitsBytes + kNoAllocCntBytes;
static_assert(kNoAllocBitsBytes % 256 == 0, "tile counts start 256-byte aligned");

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on device `dev` (current), looked up once.
static int scratch_address(int dev, void** out) {
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (dev < 0 || dev >= kMaxDevices) return DTL_ENODEVICE;
  std::lock_guard<std::mutex> g(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_last_error(e != hipSuccess ? e : hipErrorInvalidSymbol);
      return DTL_EHIP;
    }
    addr[dev] = p;
  }
  *out = addr[dev];
  return DTL_OK;
}

}  // namespace dtlh

using namespace dtlh;

extern "C" {

uint64_t dtl_noalloc_max_keys(void) { return kNoAllocMaxKeys; }

int bbf_probe_noalloc(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
                      uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream) {
  if (!d || !d_words || !d_count || d->block_cnt < 2 || d->word_cnt != static_cast<uint64_t>(d->block_cnt) * d->w ||
      (d->word_bits != 32 && d->word_bits != 64))
    return DTL_EINVAL;
  if (n && (!d_keys || !d_pos)) return DTL_EINVAL;
  if (n > kNoAllocMaxKeys) {
    set_last_error_text("bbf_probe_noalloc: more than 2^28 keys (the size of its fixed scratch)");
    return DTL_EINVAL;
  }
  const uint32_t block_bytes = d->w * d->word_bits / 8;
  if (reinterpret_cast<uintptr_t>(d_words) % (block_bytes < 16u ? block_bytes : 16u)) return DTL_EINVAL;
  const BbfEntry* e = bbf_find(d->word_bits, d->w, d->s, d->z, d->k);
  if (!e) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  int dev = 0, cu_count = 0;
  uint32_t max_lds = 0;
  int rc = current_device(&dev, &cu_count, &max_lds);
  if (rc != DTL_OK) return rc;
  void* base = nullptr;
  rc = scratch_address(dev, &base);
  if (rc != DTL_OK) return rc;
  if (n == 0) {
    const hipError_t he = write_zero_count(d_count, st);
    if (he != hipSuccess) set_last_error(he);
    return he == hipSuccess ? DTL_OK : DTL_EHIP;
  }
  const uint32_t ntiles = static_cast<uint32_t>((n + kTileKeys - 1) / kTileKeys);
  BbfProbeLaunch L;
  L.words = d_words;
  L.keys = d_keys;
  L.n = static_cast<uint32_t>(n);
  if (d->addr_mode == DTL_ADDR_POW2) {
    L.addr_magic = L.addr_mshift = L.addr_divisor = 0;
    L.addr_post = d->block_shift;
  } else {
    L.addr_magic = d->magic;
    L.addr_mshift = d->magic_shift;
    L.addr_divisor = d->block_cnt;
    L.addr_post = 0;
  }
  L.bits = static_cast<uint64_t*>(base);
  L.tile_cnt = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + kNoAllocBitsBytes);
  L.ntiles = ntiles;
  L.filter_bytes = static_cast<uint32_t>(d->size_bytes < 0xFFFFFFFFull ? d->size_bytes : 0xFFFFFFFFull);
  // the same choices as bbf_probe's direct path
  L.use_lds = d->size_bytes <= max_lds - 1024;
  L.lds_threads = d->size_bytes > 40 * 1024 ? 1024 : 256;
  L.lds_grid = 0;
  L.split = d->size_bytes < (256ull << 20) ? 1 : 0;
  L.cu_count = cu_count;
  L.stream = st;
  hipError_t he = e->probe(L);
  if (he == hipSuccess) he = compact_single(L.bits, L.tile_cnt, ntiles, L.n, match_offset, d_pos, d_count, st);
  if (he != hipSuccess) {
    set_last_error(he);
    return DTL_EHIP;
  }
  return DTL_OK;
}

}  // extern "C"
