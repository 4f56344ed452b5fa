// dtl_internal.hpp -- host-side plumbing shared by the libdtl_hip sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "dtl_sched.hpp"

namespace dtlh {

constexpr uint32_t kTileKeys = 2048;             // keys per tile of every probe kernel
constexpr uint32_t kTileWords = kTileKeys / 64;  // 64-bit match-bitmap words per tile

// Arguments of one blocked-Bloom probe launch (phase 1: bitmap + tile counts).
struct BbfProbeLaunch {
  const void* words;
  const uint32_t* keys;
  uint32_t n;
  uint32_t addr_magic, addr_mshift, addr_divisor, addr_post;
  uint64_t* bits;
  uint32_t* tile_cnt;
  uint32_t ntiles;
  uint32_t filter_bytes;
  int use_lds;
  uint32_t lds_threads;
  uint32_t lds_grid;      // 0 = from occupancy
  int split;              // lane-split global kernel where the configuration has one
  int cu_count;
  hipStream_t stream;
};

struct BbfBuildLaunch {
  void* words;
  const uint32_t* keys;
  uint64_t n;
  uint32_t addr_magic, addr_mshift, addr_divisor, addr_post;
  int cu_count;
  hipStream_t stream;
};

// One batched probe task as the device sees it.
struct BatchTaskDev {
  const void* words;
  const uint32_t* keys;
  uint32_t* pos;
  uint64_t* count;
  uint32_t n;
  uint32_t tile_begin;    // first global tile of the task
  uint32_t tile_count;
  uint32_t match_offset;
  uint32_t addr_magic, addr_mshift, addr_divisor, addr_post;
  uint32_t filter_bytes;  // staged in LDS when <= the call's LDS threshold
};

struct BbfBatchLaunch {
  const BatchTaskDev* tasks;   // device array
  const uint32_t* tile_task;   // device: global tile -> task index
  uint32_t n_tasks;
  uint32_t total_tiles;
  uint64_t* bits;
  uint32_t* tile_cnt;
  uint32_t* work_counter;      // device, zero at launch start
  uint32_t lds_max_bytes;
  uint32_t threads;
  uint32_t grid;
  uint32_t chunk;              // tiles per grab
  uint32_t half_in_flight;     // 1: half as many keys (pieces) per lane requested at once
  int cu_count;
  hipStream_t stream;
};

// Partitioned probe of one large filter (dtl_partition.hip), blocked Bloom
// or cuckoo.
struct BbfPartLaunch {
  const void* words;
  const uint32_t* keys;
  uint32_t n;
  uint32_t addr_magic, addr_mshift, addr_divisor, addr_post;
  uint32_t slice_shift;   // partition = block (or bucket) index >> slice_shift
  // partition function: 0 = block / first cuckoo bucket of the key;
  // 1 = alternate cuckoo bucket (cf_* fields, tag_mask)
  uint32_t part_kind;
  uint32_t tag_mask, cf_mode, cf_bucket_mask, cf_bucket_cnt;
  uint32_t cf_which;         // cuckoo probe pass: 1 = first bucket + victim, 2 = alternate bucket
  uint64_t cf_victim_offset;
  uint32_t nparts;
  // 0: partitions sized by a histogram pass (offsets from its scan).
  // Otherwise every (partition, lane) bin has this fixed capacity (a
  // multiple of kTileKeys): no histogram pass; a key whose bin is full takes
  // a slot of another bin (probe results do not depend on where a key sits,
  // only locality does), and cursor[] ends as each bin's fill.
  uint32_t bin_cap;
  uint32_t ptiles;        // tiles of the grouped order (ntiles, or bins * bin_cap / kTileKeys)
  uint32_t scatter_threads;  // 256, 512 or 1024 (tile = 16 keys per thread)
  uint32_t* counts;       // nparts * kPartLanes bins
  // cuckoo probe: pass 1 also fills the alternate-bucket bins (counts_next,
  // right after counts) from the same read of the keys; pass 2 then finds
  // its bins filled (counts_ready)
  uint32_t* counts_next;
  int counts_ready;
  uint32_t* offsets;      // bins + 1
  uint32_t* cursor;       // bins
  uint32_t* keys_p;       // keys grouped by partition, ptiles * kTileKeys slots (later: hit positions)
  uint32_t* pos_p;        // their original positions (later: hit positions by bucket)
  uint64_t* pbits;        // match bitmap in partitioned order, ptiles * kTileWords words
  uint32_t* ptile_cnt;    // ptiles (+ kScanTmpWords)
  uint64_t* d_hits;       // device: number of matches
  uint32_t bucket_shift;  // positions per bucket of the hit sort: 2^bucket_shift
  uint32_t nbuckets;      // position buckets (<= kHitBucketsMax)
  uint32_t hit_groups;    // workgroups of the hit sort (<= kHitGroupsMax)
  uint32_t* hcounts;      // hit_groups * nbuckets: per-workgroup bucket counts, then offsets
  uint32_t* bcounts;      // nbuckets
  uint32_t* boffsets;     // nbuckets + 1
  uint32_t* bcursor;      // nbuckets
  uint64_t* bits;         // match bitmap in key order, ntiles * kTileWords words
  uint32_t* tile_cnt;     // ntiles
  uint32_t ntiles;
  int split;              // lane-split probe kernel where the configuration has one
  int cu_count;
  hipStream_t stream;
  // Cuckoo pass 2 with capacity bins: when the device count *gate_count of
  // pass-1 matches exceeds gate_thr, the scatter leaves out every key whose
  // bit is set in skip_bits (key order: pass 1 found it already).
  const uint64_t* skip_bits;
  const uint64_t* gate_count;
  uint64_t gate_thr;
  // Blocked Bloom probe (no key can match twice): when set, the hit sort
  // writes the ascending positions + match_offset here itself and the match
  // count to *d_count_out, instead of the key-order bitmap and tile counts.
  uint32_t* out_pos;
  uint32_t match_offset;
  uint64_t* d_count_out;
};

// Stages of the hit sort (part_sort_matches) around a device-side gate
// *gate > thr, read by the kernels so that no host synchronisation is needed:
// kHitSortAll sorts hp[0 .. *d_hits) into the key-order bitmap;
// kHitSortEarly does the same only when the gate is open (else every kernel
// returns at once); kHitSortRest sorts hp[*gate .. *d_hits) and ORs it into
// the bitmap the early stage wrote when the gate is open, and sorts all of
// hp[0 .. *d_hits) as kHitSortAll when it is closed.
enum HitSortStage : uint32_t { kHitSortAll = 0, kHitSortEarly = 1, kHitSortRest = 2 };

// Hit sort of the partitioned probe: matches are bucketed by position,
// 2^18 positions per bucket or more, so that there are at most
// kHitBucketsMax buckets (an LDS histogram of them per workgroup).  Fewer,
// larger buckets give each workgroup fewer partial-line write streams in the
// placement (2^28 probes on MI355X, 2^18 against 2^16: 50% hits +2.2% Bloom
// and +1.9% cuckoo, 1% hits +1.0% / +0.4%; 2^20 no better than 2^18).
constexpr uint32_t kHitBucketsMax = 8192;
constexpr uint32_t kHitGroupsMax = 512;
constexpr uint32_t kHitShiftMin = 18;
inline uint32_t hit_bucket_shift(uint32_t ntiles) {
  uint32_t s = kHitShiftMin;
  while ((static_cast<uint64_t>(ntiles) * kTileKeys + (1ull << s) - 1) >> s > kHitBucketsMax) ++s;
  return s;
}

// Size classes of a plan's filters.  Each (configuration, class) pair is its
// own persistent launch, because the LDS a launch reserves per workgroup
// decides its occupancy:
//   0: up to kSmallFilterBytes   staged in LDS, 256-thread workgroups, several per CU;
//   1: up to the LDS capacity    staged in LDS, 1024-thread workgroups (32-bit
//                                words; 256 for 64-bit words), one per CU;
//   2: larger                    read from L2 / Infinity Cache / HBM, lane-split
//                                where the configuration allows.
// Before the split, one launch per configuration reserved 32 KiB per
// workgroup (the tuned compromise), so filters of 32-158 KiB were gathered
// from L2 like the large ones.
constexpr uint32_t kSmallFilterBytes = 32 * 1024;

inline uint32_t size_class(uint64_t filter_bytes, uint32_t lds_cap) {
  return filter_bytes <= kSmallFilterBytes ? 0u : filter_bytes <= lds_cap ? 1u : 2u;
}

struct BbfEntry {
  uint32_t word_bits, w, s, z, k;
  hipError_t (*probe)(const BbfProbeLaunch&);
  hipError_t (*build)(const BbfBuildLaunch&);
  hipError_t (*probe_batched)(const BbfBatchLaunch&);  // may be null
  hipError_t (*probe_part)(const BbfPartLaunch&);      // probe phase of the partitioned path
  uint32_t (*batched_grid)(const BbfBatchLaunch&);     // CUs x occupancy of a batched launch shape
  uint32_t split_lanes;                                // lanes per key of the lane-split kernel, 0 = none
};

constexpr uint32_t kPartMax = 2048;   // partitions of the partitioned probe
constexpr uint32_t kPartLanes = 8;    // sub-ranges per partition (one per XCD)

// Partitioned probe phases that do not depend on the configuration:
// histogram + scan + scatter of the keys, and the reordering of the matches
// (partitioned-order bitmap -> positions -> key-order bitmap + tile counts).
hipError_t partition_keys(const BbfPartLaunch& L);
hipError_t part_reorder_matches(const BbfPartLaunch& L);
// Largest hit-sort bucket whose positions k_hit_bits can write in order (in
// tiles of kTileKeys): 2^19 positions, the bucket size of 2^32 keys.
constexpr uint32_t kHitEmitTilesMax = 256;
// The two halves of part_reorder_matches: positions of the grouped-order
// matches appended to hp at *d_base (0 if null), their number to *d_count;
// then the hit sort of hp[0 .. *d_hits) into the key-order bitmap.
hipError_t part_gather_matches(const BbfPartLaunch& L, uint32_t* hp, const uint64_t* d_base, uint64_t* d_count);
hipError_t part_sort_matches(const BbfPartLaunch& L, const uint32_t* hp, const uint64_t* d_hits,
                             HitSortStage stage = kHitSortAll, const uint64_t* gate = nullptr, uint64_t thr = 0);
hipError_t add_counts(const uint64_t* a, const uint64_t* b, uint64_t* out, hipStream_t stream);

struct CfProbeLaunch {
  const void* buf;
  const uint32_t* keys;
  uint32_t n;
  uint32_t addr_mode, shift, bucket_mask, magic, mshift, bucket_cnt;
  uint64_t victim_offset;
  uint64_t* bits;
  uint32_t* tile_cnt;
  uint32_t ntiles;
  hipStream_t stream;
};

struct CfBuildLaunch {
  void* buf;
  const uint32_t* keys;
  uint64_t n;
  uint32_t addr_mode, shift, bucket_mask, magic, mshift, bucket_cnt;
  uint64_t victim_offset;
  uint32_t max_kicks, seed;
  uint64_t* failed;
  int cu_count;
  hipStream_t stream;
};

struct CfEntry {
  uint32_t tag_bits, tags_per_bucket;
  hipError_t (*probe)(const CfProbeLaunch&);
  hipError_t (*probe_part)(const BbfPartLaunch&);  // one pass over grouped keys
  hipError_t (*build)(const CfBuildLaunch&);       // null: buckets straddle 4-byte words
};

// Registries (filled once, on first use).
const std::vector<BbfEntry>& bbf_registry();
const BbfEntry* bbf_find(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k);
const CfEntry* cf_find(uint32_t tag_bits, uint32_t tags_per_bucket);

// Per-(word type, w) registration functions, one translation unit each.
#define DTL_DECLARE_REG(name) void name(std::vector<BbfEntry>& out);
DTL_DECLARE_REG(dtl_reg_u32_w1)
DTL_DECLARE_REG(dtl_reg_u32_w2)
DTL_DECLARE_REG(dtl_reg_u32_w4)
DTL_DECLARE_REG(dtl_reg_u32_w8)
DTL_DECLARE_REG(dtl_reg_u32_w16)
DTL_DECLARE_REG(dtl_reg_u64_w1)
DTL_DECLARE_REG(dtl_reg_u64_w2)
DTL_DECLARE_REG(dtl_reg_u64_w4)
DTL_DECLARE_REG(dtl_reg_u64_w8)
DTL_DECLARE_REG(dtl_reg_u64_w16)
DTL_DECLARE_REG(dtl_reg_z32_w4)
DTL_DECLARE_REG(dtl_reg_z32_w8)
DTL_DECLARE_REG(dtl_reg_z32_w16)
DTL_DECLARE_REG(dtl_reg_z64_w4)
DTL_DECLARE_REG(dtl_reg_z64_w8)
DTL_DECLARE_REG(dtl_reg_z64_w16)
#undef DTL_DECLARE_REG

// Compaction (dtl_compact.hip): exclusive scan of the tile counts (in place)
// and expansion of the bitmap into positions.  compact_single and
// scan_tile_counts use tile_cnt[ntiles .. ntiles + kScanTmpWords) as scratch,
// so every single-list tile-count buffer holds ntiles + kScanTmpWords words.
constexpr uint32_t kScanTmpWords = 1024;
hipError_t compact_single(const uint64_t* bits, uint32_t* tile_cnt, uint32_t ntiles, uint32_t n,
                          uint32_t match_offset, uint32_t* pos, uint64_t* d_count,
                          hipStream_t stream);
hipError_t compact_tasks(const uint64_t* bits, uint32_t* tile_cnt, const BatchTaskDev* tasks,
                         const uint32_t* tile_task, uint32_t n_tasks, uint32_t total_tiles,
                         hipStream_t stream);
hipError_t write_zero_count(uint64_t* d_count, hipStream_t stream);
// Exclusive scan of tile counts in place, total to *d_total.
hipError_t scan_tile_counts(uint32_t* tile_cnt, uint32_t ntiles, uint64_t* d_total, hipStream_t stream);
// For every set bit i of the bitmap (tile offsets from scan_tile_counts):
// out[base + rank] = src[i], in bit order (base = *d_out_base, or 0 if null).
hipError_t expand_gather(const uint64_t* bits, const uint32_t* offs, uint32_t ntiles, const uint32_t* src,
                         uint32_t* out, const uint64_t* d_out_base, hipStream_t stream);
hipError_t zero_u32(uint32_t* p, hipStream_t stream);

// Raises a kernel's dynamic-LDS limit when `bytes` exceeds the default
// 64 KiB (gfx950 workgroups may use up to 160 KiB).
inline hipError_t allow_dynamic_lds(const void* fn, uint32_t bytes) {
  if (bytes <= 64u * 1024u) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

// Writes n distinct keys (seeded bijection of [0, 2^31), bit 31 = high_bit).
hipError_t gen_keys(uint32_t* out, uint64_t n, uint32_t seed, uint32_t high_bit, hipStream_t st);

// Records the text of a failed HIP call (or a message) for dtl_last_error().
void set_last_error(hipError_t e);
void set_last_error_text(const char* msg);

// Device info of the current device (CU count, LDS bytes per workgroup).
int current_device(int* dev, int* cu_count, uint32_t* max_lds);

}  // namespace dtlh
