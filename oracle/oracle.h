/*
 * oracle.h -- CPU restatement of the DTL filter probe path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may link or call it; the product library
 * (libdtl_hip) never does.
 *
 * Every function restates the behaviour of the reference header named in its
 * comment (paths relative to the reference's source tree).  The restatement
 * follows the reference's scalar code paths, because the reference builds its
 * filters with those (blocked_bloomfilter_logic.hpp:94-117).
 *
 * Parity pinning (what checks each part independently of this file):
 *   hashes, POW2 addressing   closed forms; fixed shift tables for the
 *                             benchmark configurations (tests/test_oracle_*.py)
 *   MAGIC addressing          exhaustive/brute-force checks of
 *                             fast_mod == % (mirrors test/div_test.cpp:11-137)
 *                             and of the "cheap magic" search
 *   block logic (sgew/sltw/   known-answer test src/dtl/filter/filter_test.cpp:16-40,
 *   zoned, k=8 quirk)         fixed shift tables, invariants (no false
 *                             negatives, <= k bits inside the key's block,
 *                             OR-idempotence), empirical FPR vs formula
 *   cuckoo filter             closed forms, alt(alt(i)) == i, no false
 *                             negatives, SWAR vs slot compare
 *                             (test/cuckoofilter_test.cpp:10-66), mt19937
 *                             known answer (C++ standard, 10000th draw)
 *   FPR model (fpr_*)         the Poisson pmf against the exact l^i e^-l / i!;
 *                             blocked and sectorized mixtures against their
 *                             closed forms through the Poisson generating
 *                             function; the self-collision adjustment against
 *                             enumerated expected distinct bits; the zoned
 *                             inner term against its closed form where the
 *                             truncation is immaterial (tests/test_oracle_fpr.py).
 *                             The reference's Boost.Math pdf itself is absent.
 */
#ifndef DTL_ORACLE_H
#define DTL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- hashing (hash_family.hpp:22-122) ---------------------------------- */
uint32_t orc_hash(uint32_t fn, uint32_t key);

/* ---- magic-number modulo (div.hpp:29-92, libdivide.h:719-768) ----------- */
typedef struct {
  uint32_t magic;
  uint32_t shift;
  uint32_t divisor;
} orc_divisor;

int orc_next_cheap_magic(uint32_t n, orc_divisor* out); /* 0 = ok */
uint32_t orc_fast_mod(uint32_t h, orc_divisor d);
uint32_t orc_fast_div(uint32_t h, orc_divisor d);

/* ---- addressing (block_addressing_logic.hpp:40-235) -------------------- */
enum { ORC_POW2 = 0, ORC_MAGIC = 1 };

typedef struct {
  uint32_t mode;          /* ORC_POW2 / ORC_MAGIC */
  uint32_t block_cnt;
  uint32_t block_cnt_log2;/* POW2: log2(block_cnt); MAGIC: log2(next_pow2) */
  uint32_t block_cnt_mask;/* next_pow2(block_cnt) - 1 */
  orc_divisor div;        /* MAGIC only */
} orc_addr;

int orc_addr_init(orc_addr* a, uint32_t mode, uint64_t desired_block_cnt);
uint32_t orc_addr_block_idx(const orc_addr* a, uint32_t hash_value);

/* ---- blocked Bloom filter ---------------------------------------------- */
typedef struct {
  uint32_t word_bits;     /* 32 or 64 */
  uint32_t w;             /* words per block */
  uint32_t s;             /* sectors per block (zoned: == w) */
  uint32_t z;             /* zones (1 = not zoned) */
  uint32_t k;
  uint64_t m;             /* requested length in bits */
  uint64_t m_actual;      /* block_cnt * block bits */
  uint64_t word_cnt;      /* number of word_bits-sized words */
  orc_addr addr;
} orc_bbf;

/* error codes mirror the reference's exceptions */
enum {
  ORC_OK = 0,
  ORC_EINVAL = -1,   /* std::invalid_argument */
  ORC_ELOGIC = -3,   /* std::logic_error (block count) */
};

/* Validates exactly like blocked_bloomfilter.hpp:96-116 / zoned_blocked_bloomfilter.hpp:98-120. */
int orc_bbf_init(orc_bbf* f, uint32_t word_bits, uint64_t m, uint32_t k,
                 uint32_t w, uint32_t s, uint32_t z);

void orc_bbf_insert(const orc_bbf* f, void* words, uint32_t key);
void orc_bbf_batch_insert(const orc_bbf* f, void* words, const uint32_t* keys, uint64_t n);
int orc_bbf_contains(const orc_bbf* f, const void* words, uint32_t key);
/* Scalar batch_contains (blocked_bloomfilter_batch_dispatch.hpp:125-154). */
uint64_t orc_bbf_batch_contains(const orc_bbf* f, const void* words,
                                const uint32_t* keys, uint64_t n,
                                uint32_t* match_pos, uint32_t match_offset);

/* The bit positions one key touches: word index within the filter and bit
 * index within that word, in the order the reference consumes hash bits.
 * Returns the number of entries (always k).  Used by tests. */
int orc_bbf_key_bits(const orc_bbf* f, uint32_t key, uint64_t* word_idx, uint32_t* bit_idx);

/* The SIMD code path of the reference (contains_vec) writes the hash state
 * back after a k=8 word block, the scalar path does not.  These emulate the
 * SIMD semantics and the mixed scalar-head / SIMD-body / scalar-tail dispatch
 * (blocked_bloomfilter_batch_dispatch.hpp:28-69) for a key pointer whose
 * address is `key_addr_mod` modulo the vector alignment. */
int orc_bbf_contains_simd(const orc_bbf* f, const void* words, uint32_t key);
uint64_t orc_bbf_batch_contains_mixed(const orc_bbf* f, const void* words,
                                      const uint32_t* keys, uint64_t n,
                                      uint32_t* match_pos, uint32_t match_offset,
                                      uint32_t vector_len, uint32_t align_bytes,
                                      uint32_t key_addr_mod);
/* 1 if the config hits the k=8 word-block write-back quirk for some key. */
uint64_t orc_bbf_batch_key_bits(const orc_bbf* f, const uint32_t* keys, uint64_t n, uint64_t* word_idx,
                                uint32_t* bit_idx);
void orc_bbf_batch_blocks(const orc_bbf* f, const uint32_t* keys, uint64_t n, uint32_t* blocks);
uint64_t orc_bbf_batch_contains_blocks(const orc_bbf* f, const void* blocks, const uint32_t* keys, uint64_t n,
                                       uint32_t* match_pos, uint32_t match_offset);
int orc_bbf_has_k8_quirk(const orc_bbf* f);

/* ---- cuckoo filter (cuckoofilter_logic.hpp, cuckoofilter_table.hpp) ---- */
typedef struct {
  uint32_t tag_bits;
  uint32_t tags_per_bucket;
  uint32_t bytes_per_bucket;
  uint64_t m;              /* requested bits */
  uint64_t m_actual;       /* bytes_per_bucket * bucket_cnt * 8 (no victim) */
  uint64_t table_words;    /* u32 words covering the buckets */
  uint64_t victim_offset;  /* byte offset of the victim cache */
  uint64_t buffer_bytes;   /* table words + victim (12 B) + 4 B pad */
  orc_addr addr;
} orc_cf;

typedef struct { uint32_t mt[624]; uint32_t idx; } orc_mt19937;
void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);

int orc_cf_init(orc_cf* f, uint64_t m, uint32_t tag_bits, uint32_t tags_per_bucket);
/* returns 1 on success, 0 once the victim cache is in use */
int orc_cf_insert(const orc_cf* f, void* buf, uint32_t key, orc_mt19937* rng);
int orc_cf_contains(const orc_cf* f, const void* buf, uint32_t key);
uint64_t orc_cf_batch_contains(const orc_cf* f, const void* buf,
                               const uint32_t* keys, uint64_t n,
                               uint32_t* match_pos, uint32_t match_offset);
uint32_t orc_cf_tag(const orc_cf* f, uint32_t key);
uint32_t orc_cf_i1(const orc_cf* f, uint32_t key);
uint32_t orc_cf_alt(const orc_cf* f, uint32_t bucket, uint32_t tag);
uint32_t orc_cf_read_tag(const orc_cf* f, const void* buf, uint64_t bucket, uint32_t slot);
/* the SWAR "has value" tests of bitsutil.hpp:33-47, exposed for tests */
uint64_t orc_swar_hasvalue(uint32_t bits, uint64_t x, uint32_t v);
uint64_t orc_cf_count_occupied(const orc_cf* f, const void* buf);
void orc_cf_key_candidates(const orc_cf* f, const uint32_t* keys, uint64_t n, uint32_t* i1, uint32_t* i2,
                           uint32_t* tag);
uint64_t orc_cf_entries(const orc_cf* f, const void* buf, uint32_t* bucket, uint32_t* tag);

/* ---- FPR model (blocked_bloomfilter/math.hpp, fpr.hpp) ------------------ */
double orc_poisson_pmf(double lambda, uint64_t i);
double orc_fpr_bloom(uint64_t m, uint64_t n, double k);
double orc_fpr_p_cache(uint64_t s, uint64_t S, uint64_t B, double k, uint64_t i);
double orc_fpr_blocked(uint64_t m, uint64_t n, double k, uint64_t B, int self_collisions, double eps);
double orc_fpr_sectorized(uint64_t m, uint64_t n, double k, uint64_t B, uint64_t S, int self_collisions, double eps);
double orc_fpr_zoned(uint64_t m, uint64_t n, double k, uint64_t B, uint64_t S, double z, int self_collisions, double eps);
double orc_fpr_config(uint64_t m, uint64_t n, uint32_t k, uint32_t word_bytes, uint32_t w, uint32_t s, uint32_t z);
double orc_fpr_cuckoo(uint32_t tags_per_bucket, uint32_t tag_bits, double load);

#ifdef __cplusplus
}
#endif
#endif
