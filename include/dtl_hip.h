/*
 * dtl_hip.h -- C ABI of libdtl_hip, the gfx950 (MI355X) filter probe path.
 *
 * It replaces the probe/build entry points of the reference's filter classes
 * (peterboncz/bloomfilter-bsd, "DTL"):
 *
 *   bbf_build    dtl::filter::filter_base::batch_insert      filter_base.hpp:25-26
 *                (blocked_bloomfilter_logic::batch_insert     blocked_bloomfilter_logic.hpp:111-117)
 *   bbf_probe    dtl::filter::filter_base::batch_contains    filter_base.hpp:41-44
 *                (internal::dispatch<..>::batch_contains       blocked_bloomfilter_batch_dispatch.hpp:28-69)
 *   cf_probe     dtl::cf::batch_contains                      cf.cpp:49-54
 *                (cuckoofilter_logic::batch_contains           cuckoofilter_logic.hpp:390-479)
 *   cf_build_host dtl::cf::batch_insert (host build)          cf.cpp:39-42
 *                (cuckoofilter_logic::insert                   cuckoofilter_logic.hpp:291-337)
 *   cf_build     the same batch_insert on the GPU (concurrent inserts; same
 *                probe semantics, table layout differs from a sequential build)
 *   bbf_probe_batched  no single reference counterpart: many filter_base::
 *                batch_contains calls (one per filter) in one launch.
 *
 * Conventions
 *   - All buffers passed to bbf_* / cf_probe are device pointers (HBM) owned
 *     by the caller; calls are ordered on `stream` (a hipStream_t; NULL is the
 *     default stream) and return once the work is enqueued.
 *   - Filter words are laid out exactly as the reference lays them out
 *     (block-major, little-endian, no header), so a byte compare against a
 *     reference-built filter is meaningful.
 *   - Match positions: the first *d_count entries of d_pos receive the
 *     ascending positions i + match_offset of every key i that may be in the
 *     filter (the reference's selection vector).  d_pos must hold n entries.
 *     *d_count is a device uint64_t written by the call.
 *   - No exceptions cross the ABI: every function returns a dtl_status.
 *   - bbf_probe / cf_probe keep library-owned scratch per (device, stream),
 *     grown on demand.  Once a call on a stream has been captured into a
 *     hipGraph, that stream's scratch is pinned (the graph holds its
 *     address): a later call on the stream that needs more scratch returns
 *     DTL_ERUNTIME, enqueues nothing and frees nothing, so warm a stream up
 *     with the largest call before capturing it.  Replay a graph on the
 *     stream it was captured from (its replays and the stream's eager calls
 *     then share the scratch in stream order), and release the scratch with
 *     dtl_scratch_release only after every such graph is destroyed.
 */
#ifndef DTL_HIP_H
#define DTL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* dtl_stream_t; /* hipStream_t */

typedef enum dtl_status {
  DTL_OK = 0,
  DTL_EINVAL = -1,      /* std::invalid_argument in the reference */
  DTL_ERUNTIME = -2,    /* std::runtime_error */
  DTL_ELOGIC = -3,      /* std::logic_error (block count out of range) */
  DTL_ENOMEM = -4,      /* device allocation failed */
  DTL_EHIP = -5,        /* a HIP runtime call failed */
  DTL_ENODEVICE = -6    /* no usable GPU */
} dtl_status;

enum { DTL_ADDR_POW2 = 0, DTL_ADDR_MAGIC = 1 };

/* Blocked Bloom filter descriptor, computed on the host by dtl_bbf_desc_init
 * from the same parameters the reference's constructors take
 * (blocked_bloomfilter.hpp:96-116; zoned_blocked_bloomfilter.hpp:98-120). */
typedef struct dtl_bbf_desc {
  uint32_t word_bits;   /* 32 or 64 */
  uint32_t w;           /* words per block */
  uint32_t s;           /* sectors per block (zoned: == w) */
  uint32_t z;           /* zones per block (1: not zoned) */
  uint32_t k;           /* bits set per key */
  uint32_t addr_mode;   /* DTL_ADDR_POW2 when the requested m is a power of two */
  uint32_t block_cnt;
  uint32_t block_shift; /* POW2: 32 - log2(block_cnt) */
  uint32_t magic;       /* MAGIC: block = h0 - ((mulhi(h0, magic) >> magic_shift) * block_cnt) */
  uint32_t magic_shift;
  uint64_t m;           /* requested length in bits */
  uint64_t m_actual;    /* block_cnt * w * word_bits */
  uint64_t word_cnt;    /* number of words of word_bits bits */
  uint64_t size_bytes;  /* m_actual / 8: bytes the caller allocates (zeroed) */
} dtl_bbf_desc;

/* Cuckoo filter descriptor (cuckoofilter.hpp:118-127, cuckoofilter_logic.hpp:153-158).
 * Device buffer layout: packed buckets (table_words u32 words), then the
 * victim cache {u32 index; u32 tag; u8 used} at victim_offset, then padding:
 * allocate buffer_bytes, zeroed.  (The reference places the victim cache
 * after the table but sizes its buffers without it.) */
typedef struct dtl_cf_desc {
  uint32_t tag_bits;         /* 4, 8, 12, 16 or 32 */
  uint32_t tags_per_bucket;  /* 1, 2 or 4 */
  uint32_t bytes_per_bucket;
  uint32_t addr_mode;
  uint32_t bucket_cnt;
  uint32_t bucket_shift;     /* POW2: 32 - log2(bucket_cnt) */
  uint32_t bucket_mask;      /* POW2: bucket_cnt - 1 */
  uint32_t magic;
  uint32_t magic_shift;
  uint32_t reserved;
  uint64_t m;
  uint64_t m_actual;         /* bytes_per_bucket * bucket_cnt * 8 */
  uint64_t table_words;
  uint64_t victim_offset;
  uint64_t buffer_bytes;
} dtl_cf_desc;

/* Tuning knobs of one probe launch (0 = library default). */
typedef struct dtl_probe_params {
  uint32_t lds_max_bytes;   /* filters up to this size are staged in LDS (default: device LDS limit) */
  uint32_t block_threads;   /* threads per workgroup of the persistent kernels (256 or 1024);
                               partitioned probe: of the key scatter (256, 512 or 1024) */
  uint32_t grid_blocks;     /* workgroups of the persistent kernels (default: CUs x occupancy) */
  uint32_t chunk_tiles;     /* batched: tiles of 2048 keys a workgroup takes per grab (default 1) */
  uint32_t flags;           /* DTL_PROBE_* below */
  void* ev_start;           /* optional hipEvent_t recorded right before the probe phase */
  void* ev_stop;            /* optional hipEvent_t recorded right after it (before compaction) */
  uint32_t slice_log2;      /* partitioned probe: log2 of the filter bytes per partition
                               (default 19; 21 from 64 MiB) */
  uint32_t reserved;
} dtl_probe_params;

/* Probe strategy.  By default filters that fit in LDS are staged there;
 * filters of at least 32 MiB probed with at least 2^23 keys (16 MiB: 2^24
 * keys, 8 MiB: 2^25) use the partitioned probe (keys grouped by
 * filter slice so that each slice is read into L2 once); everything else
 * reads blocks straight from memory.  Blocks
 * of 32 B or more whose bits stay inside 16-byte pieces (s >= w; sectors or
 * zones of at most 16 B) are read by several lanes per key, one 16-byte piece
 * each ("lane split").  The result is the same list in every case. */
enum {
  DTL_PROBE_FORCE_GLOBAL = 1u,   /* never stage in LDS */
  DTL_PROBE_PARTITIONED = 2u,    /* always use the partitioned probe (filters larger than LDS) */
  DTL_PROBE_DIRECT = 4u,         /* never use the partitioned probe */
  DTL_PROBE_LANE_SPLIT = 8u,     /* always read blocks lane-split where the configuration allows
                                    (and the filter is under 4 GiB) */
  DTL_PROBE_ONE_LANE = 16u,      /* never: one lane loads a key's whole block */
  DTL_PROBE_PART_HISTOGRAM = 32u, /* partitioned probe of a blocked Bloom filter: size the partitions with
                                    a histogram pass over the keys instead of fixed-capacity bins */
  DTL_PROBE_HALF_IN_FLIGHT = 64u, /* batched probe: each lane requests half as many keys (or 16-byte
                                    pieces of lane-split blocks) before testing them: 4 instead of 8
                                    keys, 8 instead of 16 pieces (fewer registers, more waves) */
  DTL_PROBE_CF_PASS2_ALL = 128u   /* partitioned cuckoo probe with fixed-capacity bins: the second pass
                                    regroups every key, also when the first pass found more than n/32
                                    of them (by default it then leaves those keys out) */
};

/* ---- descriptors -------------------------------------------------------- */
/* Validates (word_bits, k, w, s, z) like the reference's runtime dispatch and
 * sizes the filter for m requested bits; z = 1 selects the sectorized filter
 * (dtl::blocked_bloomfilter), z > 1 the zoned one (dtl::zoned_blocked_bloomfilter,
 * s is then ignored). */
int dtl_bbf_desc_init(dtl_bbf_desc* d, uint32_t word_bits, uint64_t m, uint32_t k, uint32_t w,
                      uint32_t s, uint32_t z);
int dtl_cf_desc_init(dtl_cf_desc* d, uint64_t m, uint32_t tag_bits, uint32_t tags_per_bucket);

/* ---- blocked Bloom filter ---------------------------------------------- */
/* Inserts n keys with atomicOr (order-independent, so the words equal the
 * reference's sequential insert bit for bit). */
int bbf_build(const dtl_bbf_desc* d, void* d_words, const uint32_t* d_keys, uint64_t n,
              dtl_stream_t stream);

/* Probes n keys (n < 2^32) and writes the ascending match positions. */
int bbf_probe(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
              uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream);

/* bbf_probe without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same list as
 * bbf_probe; it never takes the partitioned probe (large filters probed with
 * large batches run at the direct kernel's rate).  Its match bitmap and tile
 * counts live in one fixed device array of 33.1 MB that each GPU the library
 * runs on holds until the process ends; n above dtl_noalloc_max_keys()
 * (2^28) returns DTL_EINVAL.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams, or unordered host
 * threads) race; calls ordered on one stream do not.  It calls no hipMalloc
 * or hipFree. */
int bbf_probe_noalloc(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
                      uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream);
/* Largest n bbf_probe_noalloc accepts (2^28). */
uint64_t dtl_noalloc_max_keys(void);

/* bbf_probe with explicit launch parameters (NULL = defaults). */
int bbf_probe_ex(const dtl_bbf_desc* d, const void* d_words, const uint32_t* d_keys, uint64_t n,
                 uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count,
                 const dtl_probe_params* params, dtl_stream_t stream);

/* ---- cuckoo filter ----------------------------------------------------- */
int cf_probe(const dtl_cf_desc* d, const void* d_buf, const uint32_t* d_keys, uint64_t n,
             uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, dtl_stream_t stream);
/* cf_probe with explicit parameters (NULL = defaults; same list either way).
 * Tables of 16 MiB or more probed with at least 2^23 keys (8 MiB: 2^25)
 * use the partitioned probe: keys grouped by the slice of their first bucket, that
 * bucket and the victim cache tested from L2, then regrouped by the slice of
 * their alternate bucket and that bucket tested; ~20 bytes of scratch per key.
 * Where the first pass finds more than n/32 keys (high hit rates), the second
 * regroups only the keys it did not find (decided on the device).
 * Honoured: flags DTL_PROBE_DIRECT / DTL_PROBE_PARTITIONED / DTL_PROBE_CF_PASS2_ALL, slice_log2,
 * block_threads (scatter workgroup size), ev_start / ev_stop.  No counterpart
 * of its own in the reference (cf.cpp:49-54 has no launch parameters). */
int cf_probe_ex(const dtl_cf_desc* d, const void* d_buf, const uint32_t* d_keys, uint64_t n,
                uint32_t match_offset, uint32_t* d_pos, uint64_t* d_count, const dtl_probe_params* params,
                dtl_stream_t stream);

/* Host build: inserts n keys into a zeroed host buffer of buffer_bytes,
 * kicking victims with mt19937(seed) (the reference seeds a per-thread
 * mt19937 from std::random_device, thread.hpp:71-79).  *inserted receives the
 * number of keys for which insert returned true; returns DTL_OK even when
 * the table overflowed (check *inserted). */
int cf_build_host(const dtl_cf_desc* d, void* h_buf, const uint32_t* h_keys, uint64_t n,
                  uint32_t seed, uint64_t* inserted);

/* GPU build: inserts n device keys into the device table d_buf (buffer_bytes,
 * zeroed before the first call; later calls add keys).  Replaces
 * dtl::cf::batch_insert (cf.cpp:39-42; cuckoofilter_logic.hpp:291-337),
 * which inserts one key after another: here all keys walk their kick chains
 * at once, each slot change being one compare-and-swap, with the reference's
 * rules otherwise (first empty slot of i1, then of i2, then a kick of a
 * pseudo-random slot -- seeded by `seed` -- and a walk of at most max_kicks
 * steps, 0 = 500, ending in the victim cache).  The table is a valid cuckoo
 * table for the keys (every key's tag in one of its two buckets or the victim
 * cache) whenever *d_failed is unchanged: d_failed is a device uint64_t to
 * which the call adds the number of walks that found the victim cache taken
 * (their tag is dropped and some inserted key may then be missed).  Which
 * slot holds which tag depends on the interleaving of the inserts, so bytes
 * differ from cf_build_host's; cf_probe answers the same for every key that
 * was inserted.  Unlike the reference, keys are still placed in the table
 * after the victim cache is used.  DTL_EINVAL for 12-bit tags with 2 or 4
 * slots (buckets straddle 4-byte words): build those with cf_build_host. */
int cf_build(const dtl_cf_desc* d, void* d_buf, const uint32_t* d_keys, uint64_t n, uint32_t seed,
             uint32_t max_kicks, uint64_t* d_failed, dtl_stream_t stream);

/* ---- batched probe ------------------------------------------------------ */
/* One probe task: a filter with its own probe batch and output list. */
typedef struct dtl_bbf_task {
  dtl_bbf_desc desc;        /* (word_bits, w, s, z, k) must match across a call */
  const void* d_words;
  const uint32_t* d_keys;
  uint64_t n;
  uint32_t match_offset;
  uint32_t reserved;
  uint32_t* d_pos;          /* capacity n */
  uint64_t* d_count;        /* device */
} dtl_bbf_task;

/* A batch plan owns the device-side task table and scratch for a fixed set
 * of tasks; running it enqueues, per group, four kernels (clear, persistent
 * probe over the group's tasks, per-task scan, expand) with no allocation or
 * memset, so a run can be captured into a hipGraph and replayed.  Tasks may
 * mix configurations (word_bits, w, s, z, k) and filter sizes: they are
 * grouped by configuration and size class in order of first appearance.
 * The first group runs on `stream`, the others round robin on up to three
 * side streams of the plan (default priority, created on the first run), forked from
 * `stream` by an event and joined back into it before the call returns, so
 * the run is ordered on `stream` as a whole (under capture the fork and join
 * are graph edges).  Task buffers are referenced, not copied: their contents
 * may change between runs; runs of one plan must not overlap. */
typedef struct dtl_batch_plan dtl_batch_plan;

int bbf_batch_plan_create(const dtl_bbf_task* tasks, uint32_t n_tasks,
                          const dtl_probe_params* params, dtl_batch_plan** plan);
int bbf_batch_plan_run(dtl_batch_plan* plan, dtl_stream_t stream);
void bbf_batch_plan_destroy(dtl_batch_plan* plan);
/* Number of configuration groups (persistent launches per run). */
uint32_t bbf_batch_plan_group_count(const dtl_batch_plan* plan);
/* Launch shape of group `group` (as set at creation or by the autotuner), its
 * task count, probe keys and largest filter.  No reference counterpart (plan
 * introspection). */
int bbf_batch_plan_group_shape(const dtl_batch_plan* plan, uint32_t group, dtl_probe_params* shape,
                               uint32_t* n_tasks, uint64_t* keys, uint32_t* max_filter_bytes);

/* Autotunes each group's launch shape (LDS staging threshold, workgroup
 * size, keys per lane in flight, tiles per grab, grid) by timing `reps` runs
 * of each candidate on the group's own tasks; the fastest shapes are kept
 * for later runs and the one of the group with the most keys is copied to
 * *best. */
int bbf_batch_plan_tune(dtl_batch_plan* plan, dtl_stream_t stream, uint32_t reps,
                        dtl_probe_params* best);

/* Create + run + synchronize + destroy.  Each task's list equals what
 * bbf_probe would produce for it. */
int bbf_probe_batched(const dtl_bbf_task* tasks, uint32_t n_tasks, const dtl_probe_params* params,
                      dtl_stream_t stream);

/* ---- learned configuration selection ------------------------------------------ */
/* No single reference counterpart: replaces the calibration sweep + skyline
 * lookup that picks (word size, w, s, z, k, m) from (n, tw)
 * (model/calibration.cpp:251-311, skyline_matrix.hpp:118-155,
 * optimizer.hpp:25-140) with GPU-timed calibration, measured FPRs and an FP32
 * 8-64-64-1 MLP; the calibrated table is the fallback and the yardstick of
 * the MLP's regret. */
typedef struct dtl_calib_params {
  uint32_t log2_m_min;      /* smallest calibrated filter, log2 bits (default 13) */
  uint32_t log2_m_max;      /* largest, log2 bits (default 33 = 1 GiB) */
  uint32_t probe_keys_log2; /* log2 of the absent keys the timing runs draw from, 14..30 (default 27);
                               a run's tasks take windows of up to a quarter of them */
  uint32_t reps;            /* timed passes over every configuration, median kept (default 3) */
  uint32_t fpr_keys_log2;   /* absent keys per FPR measurement (default 24) */
  uint32_t flags;
} dtl_calib_params;

typedef struct dtl_bbf_choice {
  uint32_t word_bits, w, s, z, k, reserved;
  uint64_t m;               /* filter length in bits */
  double cost_ns;           /* calibrated-table overhead per probe: t_lookup + tw * fpr */
  double predicted;         /* the chooser's own estimate (MLP or table) */
} dtl_bbf_choice;

typedef struct dtl_cost_model dtl_cost_model;

int dtl_model_calibrate(const dtl_calib_params* params, dtl_cost_model** model, dtl_stream_t stream);
int dtl_model_train(dtl_cost_model* model, double train_fraction, uint32_t samples, uint32_t epochs,
                    uint32_t seed, dtl_stream_t stream);
int dtl_model_choose(const dtl_cost_model* model, uint64_t n, double tw_ns, int use_mlp,
                     dtl_bbf_choice* choice, dtl_stream_t stream);
double dtl_model_cost(const dtl_cost_model* model, const dtl_bbf_choice* choice, uint64_t n, double tw_ns);
double dtl_model_lookup_ns(const dtl_cost_model* model, int config, double m);
double dtl_model_fpr(const dtl_cost_model* model, int config, double bits_per_key);
int dtl_model_config_count(const dtl_cost_model* model);
int dtl_model_get_config(const dtl_cost_model* model, int config, dtl_bbf_choice* out);
double dtl_model_calibration_seconds(const dtl_cost_model* model);
int dtl_model_save(const dtl_cost_model* model, const char* path);
int dtl_model_load(const char* path, dtl_cost_model** model);
void dtl_model_destroy(dtl_cost_model* model);

/* The MLP itself: FP32, dtl_mlp_input_count() inputs (8), two hidden layers
 * of 64 (ReLU), one output; parameters W1[64][IN] b1[64] W2[64][64] b2[64]
 * W3[64] b3 with IN = dtl_mlp_input_count(). */
typedef struct dtl_mlp dtl_mlp;
int dtl_mlp_input_count(void);
int dtl_mlp_param_count(void);
int dtl_mlp_create(uint32_t seed, dtl_mlp** mlp);
void dtl_mlp_destroy(dtl_mlp* mlp);
/* Adam on 0.5 * mean(w * (f(x) - t)^2); d_x: n rows of dtl_mlp_input_count()
 * floats (row-major, n x 8), d_t: n targets, d_w: n sample weights or NULL
 * (all 1); device memory.  dtl_mlp_predict and dtl_mlp_gradient read d_x the
 * same way; a shorter buffer is read past its end. */
int dtl_mlp_train(dtl_mlp* mlp, const float* d_x, const float* d_t, const float* d_w, uint32_t n,
                  uint32_t epochs, uint32_t batch, float lr, uint32_t seed, dtl_stream_t stream);
/* As dtl_mlp_train; with d_best (n floats, device) each row's weight gains
 * dyn * exp(-max(f(x) - d_best[i], 0) / temp), held constant in the gradient:
 * rows predicted near or below the best target of their query weigh more
 * (temp > 0, in target units). */
int dtl_mlp_train_ex(dtl_mlp* mlp, const float* d_x, const float* d_t, const float* d_w, const float* d_best,
                     float dyn, float temp, uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                     dtl_stream_t stream);
int dtl_mlp_predict(const dtl_mlp* mlp, const float* d_x, uint32_t n, float* d_y, dtl_stream_t stream);
int dtl_mlp_gradient(const dtl_mlp* mlp, const float* d_x, const float* d_t, const float* d_w, uint32_t n,
                     float* d_grad, dtl_stream_t stream);
int dtl_mlp_get_params(const dtl_mlp* mlp, float* h_params);
int dtl_mlp_set_params(dtl_mlp* mlp, const float* h_params);

/* ---- measurement ----------------------------------------------------------- */
/* Random-sector read microbenchmark (no reference counterpart; used to
 * measure the random-access bandwidth the probe kernels are compared with).
 * Reads n_reads pseudo-random aligned sectors (64 B, or 32 B for modes 4-5)
 * of a buffer of buf_bytes (a power of two); the access pattern `mode` is
 * described in dtl_gather.hip.  Modes 8 and 9 stage the buffer (64 B ..
 * 128 KiB, 128 B .. for mode 9) in every workgroup's LDS and read random
 * 4-byte words (8) or random 64 B blocks (9) of it.  d_sink is a device
 * uint64_t scratch word. */
int dtl_gather_bench(const void* d_buf, uint64_t buf_bytes, uint32_t mode, uint64_t n_reads,
                     uint64_t* d_sink, dtl_stream_t stream);

/* ---- scratch ------------------------------------------------------------- */
/* Frees the probe scratch of (current device, stream) after synchronizing the
 * stream, and clears its pin (see Conventions).  Call it only once no graph
 * that captured probe calls on the stream will be replayed.  No reference
 * counterpart (the reference's probes allocate nothing). */
int dtl_scratch_release(dtl_stream_t stream);
/* Bytes of probe scratch held for (current device, stream); *captured (may
 * be NULL) receives 1 if the scratch is pinned by a captured call. */
uint64_t dtl_scratch_bytes(dtl_stream_t stream, int* captured);

/* ---- misc ---------------------------------------------------------------- */
const char* dtl_status_string(int status);
/* Text of the last failed HIP call of this thread inside the library. */
const char* dtl_last_error(void);
/* 1 if a kernel for this configuration is compiled into the library. */
int dtl_bbf_supported(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k);
/* Number of compiled blocked Bloom filter configurations. */
int dtl_bbf_config_count(void);
/* Lanes per key (2, 4 or 8) of the lane-split probe kernel of a configuration
 * (DTL_PROBE_LANE_SPLIT), 0 if it has none (blocks under 32 B, or a sectorized
 * filter whose sectors cross 16-byte pieces; zoned filters with zones of 32 B
 * or more have one, each lane testing the zone words in its own piece) or is
 * not compiled.  No counterpart in
 * the reference (its SIMD kernels have no such choice). */
int dtl_bbf_split_lanes(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k);
int dtl_version(void);

#ifdef __cplusplus
}
#endif
#endif /* DTL_HIP_H */
