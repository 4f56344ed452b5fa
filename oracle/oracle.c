/*
 * oracle.c -- CPU restatement of the DTL filter probe path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path (see
 * oracle.h for the list of what pins each function).  The structure follows
 * the reference's recursive templates one function per template, evaluated
 * at run time; file:line citations point into the reference source tree.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* small integer helpers (math.hpp:13-46)                                  */
/* ---------------------------------------------------------------------- */
static int is_pow2_u64(uint64_t x) { /* math.hpp:13-15 */
  return x == 1 ? 1 : (x && ((x & (x - 1)) == 0));
}
static uint64_t next_pow2_u64(uint64_t v) { /* math.hpp:17-19, v >= 2 */
  return 1ull << (64 - __builtin_clzll(v - 1));
}
static uint32_t log2_u32(uint32_t n) { return 31u - (uint32_t)__builtin_clz(n); }

/* ---------------------------------------------------------------------- */
/* hash family: h_i(key) = key * C_i mod 2^32 (hash_family.hpp:22-122)      */
/* ---------------------------------------------------------------------- */
static const uint32_t kMul[17] = {
    596572387u,  370248451u,  2654435769u, 1799596469u, 0x9e3779b1u,
    2284105051u, 1203114875u, 1150766481u, 2729912477u, 1884591559u,
    770785867u,  2667333959u, 1550580529u, 0xcc9e2d51u, 0x1b873593u,
    0x85ebca6bu, 0xc2b2ae35u};

uint32_t orc_hash(uint32_t fn, uint32_t key) { return key * kMul[fn]; }

/* ---------------------------------------------------------------------- */
/* cheap magic divisors (div.hpp:29-57 with libdivide.h:719-768)           */
/* ---------------------------------------------------------------------- */
int orc_next_cheap_magic(uint32_t n, orc_divisor* out) {
  const uint32_t d_max = 0xFFFFFFFFu - 2u; /* div.hpp:18 */
  if (n < 2) n = 2;
  for (uint32_t d = n; d <= d_max; d++) {
    if ((d & (d - 1)) == 0) continue; /* powers of two take libdivide's shift path */
    const uint32_t L = log2_u32(d);
    const unsigned __int128 num = ((unsigned __int128)1) << (32 + L);
    const uint64_t q = (uint64_t)(num / d);
    const uint64_t rem = (uint64_t)(num % d);
    const uint64_t e = d - rem;
    if (e < (1ull << L)) { /* "algorithm 1": no add step needed */
      out->magic = (uint32_t)(q + 1);
      out->shift = L;
      out->divisor = d;
      return 0;
    }
  }
  return -1;
}

uint32_t orc_fast_div(uint32_t h, orc_divisor d) { /* div.hpp:67-80 */
  const uint32_t hi = (uint32_t)(((uint64_t)h * (uint64_t)d.magic) >> 32);
  return hi >> d.shift;
}

uint32_t orc_fast_mod(uint32_t h, orc_divisor d) { /* div.hpp:88-92 */
  return h - orc_fast_div(h, d) * d.divisor;
}

/* ---------------------------------------------------------------------- */
/* block addressing (block_addressing_logic.hpp:40-235)                    */
/* ---------------------------------------------------------------------- */
int orc_addr_init(orc_addr* a, uint32_t mode, uint64_t desired) {
  const uint64_t max_block_cnt = 1ull << 31;
  memset(a, 0, sizeof(*a));
  a->mode = mode;
  if (desired > max_block_cnt) return ORC_ELOGIC;
  if (desired < 2) return ORC_EINVAL; /* next_power_of_two(1) is undefined (math.hpp:17-19) */
  if (mode == ORC_POW2) { /* :163-184 */
    const uint64_t bc = next_pow2_u64(desired);
    if (bc < desired || next_pow2_u64(bc) > max_block_cnt) return ORC_ELOGIC;
    a->block_cnt = (uint32_t)bc;
    a->block_cnt_log2 = log2_u32((uint32_t)bc);
    a->block_cnt_mask = (uint32_t)(bc - 1);
  } else { /* :61-83 */
    orc_divisor d;
    if (orc_next_cheap_magic((uint32_t)desired, &d) != 0) return ORC_ELOGIC;
    const uint64_t bc = d.divisor;
    if (bc < desired || next_pow2_u64(bc) > max_block_cnt) return ORC_ELOGIC;
    a->block_cnt = (uint32_t)bc;
    a->block_cnt_log2 = log2_u32((uint32_t)next_pow2_u64(bc));
    a->block_cnt_mask = (uint32_t)(next_pow2_u64(bc) - 1);
    orc_next_cheap_magic((uint32_t)bc, &a->div); /* the divisor is cheap itself */
  }
  return ORC_OK;
}

uint32_t orc_addr_block_idx(const orc_addr* a, uint32_t h) {
  if (a->mode == ORC_POW2) return h >> (32u - a->block_cnt_log2); /* :205-210 */
  return orc_fast_mod(h, a->div);                                 /* :104-110 */
}

/* ---------------------------------------------------------------------- */
/* block logic                                                             */
/* ---------------------------------------------------------------------- */
/* Static recursion state of the templates: the next hash function index and
 * the number of unconsumed hash bits. */
typedef struct {
  uint32_t fn;
  uint32_t rem;
} sched_t;

typedef struct {
  uint32_t n;
  uint32_t word[64];  /* word offset inside the block */
  uint32_t bit[64];   /* bit inside the word */
} pairs_t;

static void push(pairs_t* p, uint32_t word, uint32_t bit) {
  p->word[p->n] = word;
  p->bit[p->n] = bit;
  p->n++;
}

/* word_block<s_, k_> (blocked_bloomfilter_block_logic_sgew.hpp:43-92) and its
 * k=8 specialization (..._sgew_specialization.hpp:35-167).  Sets k_ bits in a
 * word of word_bits bits split into s_ sectors.  `hv` is the caller's hash
 * value variable: the generic template updates it in place (:77); the k=8
 * specialization computes its hashes in locals and never writes it back
 * (..._specialization.hpp:67-167), unless `simd` selects the vectorized
 * overload, which recurses into the generic template (:171-200). */
static void word_block(uint32_t key, uint32_t word_bits, uint32_t s_, uint32_t k_,
                       sched_t* st, uint32_t* hv, uint32_t word, pairs_t* out, int simd) {
  const uint32_t sector_bits = word_bits / s_;
  const uint32_t bpk = log2_u32(sector_bits);
  const uint32_t k_per_sector = k_ / s_;
  const int write_back = simd || (k_ != 8);
  uint32_t h = *hv;
  for (uint32_t i = 0; i < k_; i++) {
    if (st->rem < bpk) { /* rehash with the next function */
      h = orc_hash(st->fn, key);
      st->fn++;
      st->rem = 32;
    }
    const uint32_t shift = st->rem - bpk;
    const uint32_t sector_idx = i / k_per_sector;
    const uint32_t bit = ((h >> shift) & (sector_bits - 1)) + sector_idx * sector_bits;
    push(out, word, bit);
    st->rem -= bpk;
  }
  if (write_back) *hv = h;
}

/* multiword_block (sgew, s >= w): words in order, k/w bits each
 * (blocked_bloomfilter_block_logic_sgew.hpp:226-445). */
static void multiword_block(const orc_bbf* f, uint32_t key, pairs_t* out, int simd) {
  sched_t st = {1u, 0u}; /* block_hash_fn_idx = 1 (blocked_bloomfilter.hpp:46) */
  uint32_t hv = 0;       /* :336 */
  for (uint32_t j = 0; j < f->w; j++)
    word_block(key, f->word_bits, f->s / f->w, f->k / f->w, &st, &hv, j, out, simd);
}

/* multiword_sector (sltw): each bit picks a word of the sector and a bit of
 * that word from one hash chunk (blocked_bloomfilter_block_logic_sltw.hpp:47-218). */
static void multiword_sector(uint32_t key, uint32_t word_bits, uint32_t words, uint32_t k_,
                             uint32_t sector_base, sched_t* st, uint32_t* hv, pairs_t* out) {
  const uint32_t wl = log2_u32(words);
  const uint32_t bl = log2_u32(word_bits);
  const uint32_t bpk = wl + bl;
  for (uint32_t i = 0; i < k_; i++) {
    if (st->rem < bpk) {
      *hv = orc_hash(st->fn, key);
      st->fn++;
      st->rem = 32;
    }
    const uint32_t word = (*hv >> (st->rem - wl)) & ((1u << wl) - 1u);
    const uint32_t bit = (*hv >> (st->rem - wl - bl)) & ((1u << bl) - 1u);
    push(out, sector_base + word, bit);
    st->rem -= bpk;
  }
}

/* multisector_block (sltw, s < w) (blocked_bloomfilter_block_logic_sltw.hpp:415-591). */
static void multisector_block(const orc_bbf* f, uint32_t key, pairs_t* out) {
  sched_t st = {1u, 0u};
  uint32_t hv = 0;
  const uint32_t wps = f->w / f->s;
  for (uint32_t i = 0; i < f->s; i++)
    multiword_sector(key, f->word_bits, wps, f->k / f->s, i * wps, &st, &hv, out);
}

/* multizone_block (blocked_bloomfilter_block_logic_zoned.hpp:50-285): per zone,
 * one word chosen from log2(w/z) hash bits, then k/z bits in it through
 * word_block<s=1>. */
static void multizone_block(const orc_bbf* f, uint32_t key, pairs_t* out, int simd) {
  sched_t st = {1u, 0u};
  uint32_t hv = 0;
  const uint32_t wpz = f->w / f->z;
  const uint32_t zb = log2_u32(wpz);
  for (uint32_t zi = 0; zi < f->z; zi++) {
    if (st.rem < zb) { /* :85-87, :116 */
      hv = orc_hash(st.fn, key);
      st.fn++;
      st.rem = 32;
    }
    const uint32_t word = zi * wpz + ((hv >> (st.rem - zb)) & (wpz - 1u));
    st.rem -= zb;
    word_block(key, f->word_bits, 1u, f->k / f->z, &st, &hv, word, out, simd);
  }
}

static void key_pairs(const orc_bbf* f, uint32_t key, pairs_t* out, int simd) {
  out->n = 0;
  if (f->z > 1) multizone_block(f, key, out, simd);
  else if (f->s >= f->w) multiword_block(f, key, out, simd);     /* block_logic.hpp:28-32 */
  else multisector_block(f, key, out);
}

static uint64_t block_word_base(const orc_bbf* f, uint32_t key) {
  /* blocked_bloomfilter_logic.hpp:96-98: word index = block_idx << log2(w) */
  const uint32_t h0 = orc_hash(0, key);
  const uint32_t b = orc_addr_block_idx(&f->addr, h0);
  return (uint64_t)b * f->w;
}

static int test_bit(const orc_bbf* f, const void* words, uint64_t wi, uint32_t bit) {
  if (f->word_bits == 32) return (((const uint32_t*)words)[wi] >> bit) & 1u;
  return (int)((((const uint64_t*)words)[wi] >> bit) & 1u);
}

static void set_bit(const orc_bbf* f, void* words, uint64_t wi, uint32_t bit) {
  if (f->word_bits == 32) ((uint32_t*)words)[wi] |= 1u << bit;
  else ((uint64_t*)words)[wi] |= 1ull << bit;
}

/* ---------------------------------------------------------------------- */
/* configuration validation + sizing                                       */
/* ---------------------------------------------------------------------- */
static uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

int orc_bbf_init(orc_bbf* f, uint32_t word_bits, uint64_t m, uint32_t k,
                 uint32_t w, uint32_t s, uint32_t z) {
  memset(f, 0, sizeof(*f));
  if (word_bits != 32 && word_bits != 64) return ORC_EINVAL; /* filter.hpp:64 */
  const uint32_t word_bytes = word_bits / 8;
  uint32_t s_eff, k_eff;
  if (z <= 1) {
    /* blocked_bloomfilter.hpp:169-229 */
    if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return ORC_EINVAL;
    if (s != 1 && s != 2 && s != 4 && s != 8 && s != 16) return ORC_EINVAL;
    s_eff = (s == 1) ? 1 : umin(w * word_bytes, s);
    if (k < 1 || k > 16) return ORC_EINVAL;
    if (k == 1) k_eff = umax(1, s_eff);
    else if (k == 16) k_eff = umax(16, s_eff);
    else k_eff = umax((k % s_eff == 0) ? k : 1u, s_eff);
    z = 1;
  } else {
    /* zoned_blocked_bloomfilter.hpp:102-215 */
    if (w <= z) return ORC_EINVAL;
    if (w != 4 && w != 8 && w != 16) return ORC_EINVAL;
    if (z != 2 && z != 4 && z != 8) return ORC_EINVAL;
    z = umin(w / 2, z);
    if (k < 2 || k > 16 || (k & 1)) return ORC_EINVAL;
    if (k == 16) k_eff = umax(16, z);
    else k_eff = umax((k % z == 0) ? k : 2u, z);
    s = s_eff = w;
  }
  f->word_bits = word_bits;
  f->w = w;
  f->s = s_eff;
  f->z = z;
  f->k = k_eff;
  f->m = m;
  /* construct: blocked_bloomfilter_logic.hpp:80-82, addressing mode from the
   * requested m (blocked_bloomfilter.hpp:235-237) */
  const uint64_t B = (uint64_t)word_bits * w;
  const uint32_t mode = is_pow2_u64(m) ? ORC_POW2 : ORC_MAGIC;
  int rc = orc_addr_init(&f->addr, mode, (m + (B - 1)) / B);
  if (rc != ORC_OK) return rc;
  f->m_actual = (uint64_t)f->addr.block_cnt * B;
  f->word_cnt = (uint64_t)f->addr.block_cnt * w;
  if (k_eff != k || s_eff != s) return ORC_EINVAL; /* :107-114 */
  return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* insert / contains / batch_contains                                      */
/* ---------------------------------------------------------------------- */
void orc_bbf_insert(const orc_bbf* f, void* words, uint32_t key) { /* logic.hpp:94-104 */
  pairs_t p;
  key_pairs(f, key, &p, 0);
  const uint64_t base = block_word_base(f, key);
  for (uint32_t i = 0; i < p.n; i++) set_bit(f, words, base + p.word[i], p.bit[i]);
}

void orc_bbf_batch_insert(const orc_bbf* f, void* words, const uint32_t* keys, uint64_t n) {
  for (uint64_t i = 0; i < n; i++) orc_bbf_insert(f, words, keys[i]); /* :111-117 */
}

static int contains_impl(const orc_bbf* f, const void* words, uint32_t key, int simd) {
  pairs_t p;
  key_pairs(f, key, &p, simd);
  const uint64_t base = block_word_base(f, key);
  int found = 1;
  for (uint32_t i = 0; i < p.n; i++) found &= test_bit(f, words, base + p.word[i], p.bit[i]);
  return found;
}

int orc_bbf_contains(const orc_bbf* f, const void* words, uint32_t key) {
  return contains_impl(f, words, key, 0); /* logic.hpp:124-136 */
}

int orc_bbf_contains_simd(const orc_bbf* f, const void* words, uint32_t key) {
  return contains_impl(f, words, key, 1);
}

uint64_t orc_bbf_batch_contains(const orc_bbf* f, const void* words, const uint32_t* keys,
                                uint64_t n, uint32_t* match_pos, uint32_t match_offset) {
  /* batch_dispatch.hpp:125-154: ascending positions read_pos + match_offset */
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; i++)
    if (contains_impl(f, words, keys[i], 0)) match_pos[c++] = (uint32_t)i + match_offset;
  return c;
}

uint64_t orc_bbf_batch_contains_mixed(const orc_bbf* f, const void* words, const uint32_t* keys,
                                      uint64_t n, uint32_t* match_pos, uint32_t match_offset,
                                      uint32_t vector_len, uint32_t align_bytes,
                                      uint32_t key_addr_mod) {
  /* batch_dispatch.hpp:37-67: scalar head (even when aligned), SIMD body,
   * scalar tail */
  uint64_t head = (align_bytes - (key_addr_mod % align_bytes)) / 4u;
  if (head > n) head = n;
  const uint64_t body = ((n - head) / vector_len) * vector_len;
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; i++) {
    const int simd = (i >= head && i < head + body);
    if (contains_impl(f, words, keys[i], simd)) match_pos[c++] = (uint32_t)i + match_offset;
  }
  return c;
}

int orc_bbf_key_bits(const orc_bbf* f, uint32_t key, uint64_t* word_idx, uint32_t* bit_idx) {
  pairs_t p;
  key_pairs(f, key, &p, 0);
  const uint64_t base = block_word_base(f, key);
  for (uint32_t i = 0; i < p.n; i++) {
    word_idx[i] = base + p.word[i];
    bit_idx[i] = p.bit[i];
  }
  return (int)p.n;
}

/* Large-filter checking without a dense copy of the filter on the host.
 * Every function below runs the same key_pairs / block addressing as
 * orc_bbf_insert and orc_bbf_contains. */

/* (word index, bit) pairs of every key of a batch, in key order; returns
 * their number (at most 64 per key).  OR-ing 1 << bit into word index over
 * all pairs gives exactly the words orc_bbf_batch_insert sets. */
uint64_t orc_bbf_batch_key_bits(const orc_bbf* f, const uint32_t* keys, uint64_t n, uint64_t* word_idx,
                                uint32_t* bit_idx) {
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; i++) c += (uint64_t)orc_bbf_key_bits(f, keys[i], word_idx + c, bit_idx + c);
  return c;
}

/* Block index of every key (block_addressing_logic.hpp:104-110, :205-210). */
void orc_bbf_batch_blocks(const orc_bbf* f, const uint32_t* keys, uint64_t n, uint32_t* blocks) {
  for (uint64_t i = 0; i < n; i++) blocks[i] = orc_addr_block_idx(&f->addr, orc_hash(0, keys[i]));
}

/* batch_contains where `blocks` holds, for key i, the w words of the block
 * orc_bbf_batch_blocks names for it (n x w words): the same list as
 * orc_bbf_batch_contains on the whole filter. */
uint64_t orc_bbf_batch_contains_blocks(const orc_bbf* f, const void* blocks, const uint32_t* keys, uint64_t n,
                                       uint32_t* match_pos, uint32_t match_offset) {
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; i++) {
    pairs_t p;
    key_pairs(f, keys[i], &p, 0);
    int found = 1;
    for (uint32_t j = 0; j < p.n; j++) found &= test_bit(f, blocks, i * f->w + p.word[j], p.bit[j]);
    if (found) match_pos[c++] = (uint32_t)i + match_offset;
  }
  return c;
}

int orc_bbf_has_k8_quirk(const orc_bbf* f) {
  /* The scalar and SIMD schedules differ for some key iff they differ for a
   * key whose hashes are all non-zero in every consumed chunk; checking a few
   * keys suffices because the schedule is key-independent. */
  static const uint32_t probe_keys[4] = {0x12345677u, 0x9abcdef1u, 0x0badf00du, 0x7fffffffu};
  for (int t = 0; t < 4; t++) {
    pairs_t a, b;
    key_pairs(f, probe_keys[t], &a, 0);
    key_pairs(f, probe_keys[t], &b, 1);
    for (uint32_t i = 0; i < a.n; i++)
      if (a.word[i] != b.word[i] || a.bit[i] != b.bit[i]) return 1;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* mt19937 (the engine behind dtl::this_thread::rand32, thread.hpp:71-79)  */
/* Matsumoto & Nishimura 1998, as standardised in C++11 [rand.eng.mers].   */
/* ---------------------------------------------------------------------- */
void orc_mt_seed(orc_mt19937* g, uint32_t seed) {
  g->mt[0] = seed;
  for (uint32_t i = 1; i < 624; i++)
    g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + i;
  g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937* g) {
  if (g->idx >= 624) {
    for (uint32_t i = 0; i < 624; i++) {
      const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      g->mt[i] = v;
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* ---------------------------------------------------------------------- */
/* cuckoo filter                                                           */
/* ---------------------------------------------------------------------- */
int orc_cf_init(orc_cf* f, uint64_t m, uint32_t tag_bits, uint32_t tpb) {
  memset(f, 0, sizeof(*f));
  /* cuckoofilter.hpp:184-206 */
  if (tag_bits != 4 && tag_bits != 8 && tag_bits != 12 && tag_bits != 16 && tag_bits != 32)
    return ORC_EINVAL;
  if (tpb != 1 && tpb != 2 && tpb != 4) return ORC_EINVAL;
  f->tag_bits = tag_bits;
  f->tags_per_bucket = tpb;
  f->bytes_per_bucket = (tag_bits * tpb + 7) / 8; /* cuckoofilter_table.hpp:89 */
  f->m = m;
  const uint64_t bucket_bits = (uint64_t)f->bytes_per_bucket * 8;
  const uint32_t mode = is_pow2_u64(m) ? ORC_POW2 : ORC_MAGIC; /* cuckoofilter.hpp:211-213 */
  int rc = orc_addr_init(&f->addr, mode, (m + bucket_bits - 1) / bucket_bits); /* logic.hpp:154-158 */
  if (rc != ORC_OK) return rc;
  const uint64_t nb = f->addr.block_cnt;
  f->table_words = (f->bytes_per_bucket * nb + 3) / 4; /* table.hpp:245-249 */
  f->victim_offset = f->table_words * 4;               /* logic.hpp:157 */
  f->buffer_bytes = f->victim_offset + 16;             /* victim (12 B) + pad */
  f->m_actual = (uint64_t)f->bytes_per_bucket * nb * 8; /* cuckoofilter.hpp:259 */
  return ORC_OK;
}

static uint32_t tag_mask(const orc_cf* f) {
  return (uint32_t)((1ull << f->tag_bits) - 1);
}

uint32_t orc_cf_tag(const orc_cf* f, uint32_t key) { /* logic.hpp:179-186, 202-207 */
  uint32_t t = orc_hash(1, key) & tag_mask(f);
  t += (t == 0);
  return t;
}

uint32_t orc_cf_i1(const orc_cf* f, uint32_t key) {
  return orc_addr_block_idx(&f->addr, orc_hash(0, key));
}

uint32_t orc_cf_alt(const orc_cf* f, uint32_t i, uint32_t tag) {
  if (f->addr.mode == ORC_POW2) /* logic.hpp:214-223 */
    return (i ^ (tag * 0x5bd1e995u)) & f->addr.block_cnt_mask;
  /* MAGIC: logic.hpp:241-252 */
  const uint32_t c = f->addr.block_cnt;
  const uint32_t s_mod_m = orc_fast_mod(tag * 0x5bd1e995u, f->addr.div);
  const uint32_t x = c - s_mod_m;
  const uint32_t mask = (x < i) ? 0xFFFFFFFFu : 0u;
  return (mask & c) + (x - i);
}

static uint64_t bucket_value_lo(const orc_cf* f, const uint8_t* buf, uint64_t bucket) {
  /* little-endian value of the bucket's first <= 8 bytes.  The reference
   * loads 8 bytes at the bucket start (cuckoofilter_table_scalar.hpp:61-62);
   * the bytes past the bucket cannot change the SWAR result because the
   * borrows of the subtraction only propagate upwards. */
  uint8_t tmp[8] = {0};
  memcpy(tmp, buf + bucket * f->bytes_per_bucket, f->bytes_per_bucket < 8 ? f->bytes_per_bucket : 8);
  uint64_t lo;
  memcpy(&lo, tmp, 8);
  return lo;
}

uint32_t orc_cf_read_tag(const orc_cf* f, const void* buf, uint64_t i, uint32_t j) {
  /* cuckoofilter_table.hpp:111-136 */
  const uint8_t* p = (const uint8_t*)buf + i * f->bytes_per_bucket;
  uint32_t tag = 0;
  switch (f->tag_bits) {
    case 4: tag = (uint32_t)(p[j >> 1] >> ((j & 1) << 2)); break;
    case 8: tag = p[j]; break;
    case 12: { uint16_t v; memcpy(&v, p + j + (j >> 1), 2); tag = (uint32_t)(v >> ((j & 1) << 2)); } break;
    case 16: { uint16_t v; memcpy(&v, p + (j << 1), 2); tag = v; } break;
    case 32: { uint32_t v; memcpy(&v, p + 4 * j, 4); tag = v; } break;
  }
  return tag & tag_mask(f);
}

static void write_tag(const orc_cf* f, void* buf, uint64_t i, uint32_t j, uint32_t t) {
  /* cuckoofilter_table.hpp:143-178 */
  uint8_t* p = (uint8_t*)buf + i * f->bytes_per_bucket;
  const uint32_t tag = t & tag_mask(f);
  switch (f->tag_bits) {
    case 4:
      p += (j >> 1);
      if ((j & 1) == 0) { *p &= 0xf0; *p |= (uint8_t)tag; }
      else { *p &= 0x0f; *p |= (uint8_t)(tag << 4); }
      break;
    case 8: p[j] = (uint8_t)tag; break;
    case 12: {
      uint16_t v;
      p += j + (j >> 1);
      memcpy(&v, p, 2);
      if ((j & 1) == 0) { v &= 0xf000; v |= (uint16_t)tag; }
      else { v &= 0x000f; v |= (uint16_t)(tag << 4); }
      memcpy(p, &v, 2);
    } break;
    case 16: { uint16_t v = (uint16_t)tag; memcpy(p + 2 * j, &v, 2); } break;
    case 32: memcpy(p + 4 * j, &tag, 4); break;
  }
}

uint64_t orc_swar_hasvalue(uint32_t bits, uint64_t x, uint32_t n) {
  /* bitsutil.hpp:33-47 */
  switch (bits) {
    case 4: { uint64_t v = x ^ (0x1111ull * n); return ((v - 0x1111ull) & ~v & 0x8888ull); }
    case 8: { uint64_t v = x ^ (0x01010101ull * n); return ((v - 0x01010101ull) & ~v & 0x80808080ull); }
    case 12: { uint64_t v = x ^ (0x001001001001ull * n); return ((v - 0x001001001001ull) & ~v & 0x800800800800ull); }
    case 16: { uint64_t v = x ^ (0x0001000100010001ull * n); return ((v - 0x0001000100010001ull) & ~v & 0x8000800080008000ull); }
  }
  return 0;
}

static int find_tag_in_buckets(const orc_cf* f, const void* buf, uint32_t i1, uint32_t i2, uint32_t tag) {
  /* cuckoofilter_table_scalar.hpp:53-81 */
  if (f->tags_per_bucket == 4 && f->tag_bits != 32) {
    const uint64_t v1 = bucket_value_lo(f, (const uint8_t*)buf, i1);
    const uint64_t v2 = bucket_value_lo(f, (const uint8_t*)buf, i2);
    return (orc_swar_hasvalue(f->tag_bits, v1, tag) | orc_swar_hasvalue(f->tag_bits, v2, tag)) != 0;
  }
  for (uint32_t j = 0; j < f->tags_per_bucket; j++)
    if (orc_cf_read_tag(f, buf, i1, j) == tag || orc_cf_read_tag(f, buf, i2, j) == tag) return 1;
  return 0;
}

typedef struct { uint32_t index; uint32_t tag; uint8_t used; } victim_view;

static victim_view read_victim(const orc_cf* f, const void* buf) {
  victim_view v;
  const uint8_t* p = (const uint8_t*)buf + f->victim_offset;
  memcpy(&v.index, p, 4);
  memcpy(&v.tag, p + 4, 4);
  v.used = p[8];
  return v;
}

static int insert_tag_to_bucket(const orc_cf* f, void* buf, uint32_t i, uint32_t tag,
                                int kickout, uint32_t* oldtag, orc_mt19937* rng) {
  /* cuckoofilter_table.hpp:185-202 */
  for (uint32_t j = 0; j < f->tags_per_bucket; j++) {
    if (orc_cf_read_tag(f, buf, i, j) == 0) {
      write_tag(f, buf, i, j, tag);
      return 1;
    }
  }
  if (kickout) {
    const uint32_t r = orc_mt_next(rng) % f->tags_per_bucket;
    *oldtag = orc_cf_read_tag(f, buf, i, r);
    write_tag(f, buf, i, r, tag);
  }
  return 0;
}

int orc_cf_insert(const orc_cf* f, void* buf, uint32_t key, orc_mt19937* rng) {
  /* cuckoofilter_logic.hpp:291-322 */
  uint8_t* vp = (uint8_t*)buf + f->victim_offset;
  if (vp[8]) return 0;
  const uint32_t i = orc_cf_i1(f, key);
  const uint32_t tag = orc_cf_tag(f, key);
  uint32_t curindex = i, curtag = tag, oldtag;
  for (uint32_t count = 0; count < 500; count++) {
    const int kickout = count > 0;
    oldtag = 0;
    if (insert_tag_to_bucket(f, buf, curindex, curtag, kickout, &oldtag, rng)) return 1;
    if (kickout) curtag = oldtag;
    curindex = orc_cf_alt(f, curindex, curtag);
  }
  memcpy(vp, &curindex, 4);
  memcpy(vp + 4, &curtag, 4);
  vp[8] = 1;
  return 1;
}

int orc_cf_contains(const orc_cf* f, const void* buf, uint32_t key) {
  /* cuckoofilter_logic.hpp:346-362 */
  const uint32_t i1 = orc_cf_i1(f, key);
  const uint32_t tag = orc_cf_tag(f, key);
  const uint32_t i2 = orc_cf_alt(f, i1, tag);
  const victim_view v = read_victim(f, buf);
  const int found = v.used && (tag == v.tag) && (i1 == v.index || i2 == v.index);
  return found | find_tag_in_buckets(f, buf, i1, i2, tag);
}

uint64_t orc_cf_batch_contains(const orc_cf* f, const void* buf, const uint32_t* keys, uint64_t n,
                               uint32_t* match_pos, uint32_t match_offset) {
  /* cuckoofilter_logic.hpp:406-434 */
  uint64_t c = 0;
  for (uint64_t i = 0; i < n; i++)
    if (orc_cf_contains(f, buf, keys[i])) match_pos[c++] = (uint32_t)i + match_offset;
  return c;
}

uint64_t orc_cf_count_occupied(const orc_cf* f, const void* buf) {
  /* cuckoofilter_logic.hpp:495-499 */
  uint64_t c = 0;
  for (uint64_t i = 0; i < f->addr.block_cnt; i++)
    for (uint32_t j = 0; j < f->tags_per_bucket; j++) c += orc_cf_read_tag(f, buf, i, j) != 0;
  return c + read_victim(f, buf).used;
}

/* Inputs of a placement check of a table built in any insert order: the two
 * candidate buckets and the tag of every key (cuckoofilter_logic.hpp:
 * 179-186, 202-252), and every occupied slot of the table. */
void orc_cf_key_candidates(const orc_cf* f, const uint32_t* keys, uint64_t n, uint32_t* i1, uint32_t* i2,
                           uint32_t* tag) {
  for (uint64_t k = 0; k < n; k++) {
    tag[k] = orc_cf_tag(f, keys[k]);
    i1[k] = orc_cf_i1(f, keys[k]);
    i2[k] = orc_cf_alt(f, i1[k], tag[k]);
  }
}

uint64_t orc_cf_entries(const orc_cf* f, const void* buf, uint32_t* bucket, uint32_t* tag) {
  uint64_t c = 0;
  for (uint64_t i = 0; i < f->addr.block_cnt; i++)
    for (uint32_t j = 0; j < f->tags_per_bucket; j++) {
      const uint32_t t = orc_cf_read_tag(f, buf, i, j);
      if (t) {
        bucket[c] = (uint32_t)i;
        tag[c] = t;
        c++;
      }
    }
  return c;
}

/* ---------------------------------------------------------------------- */
/* FPR model (blocked_bloomfilter/math.hpp:17-145, fpr.hpp:11-60)           */
/* The Poisson pmf is evaluated as exp(i ln(l) - l - lgamma(i+1)) instead of */
/* Boost.Math's pdf; tests/test_oracle_fpr.py pins it against the exact     */
/* l^i e^-l / i! and every formula below against closed forms.             */
/* ---------------------------------------------------------------------- */
static double poisson_pmf(double lambda, uint64_t i) {
  if (i == 0) return exp(-lambda);
  return exp((double)i * log(lambda) - lambda - lgamma((double)i + 1.0));
}

double orc_fpr_bloom(uint64_t m, uint64_t n, double k) { /* math.hpp:17-22 */
  return pow(1.0 - pow(1.0 - (1.0 / (double)m), k * (double)n), k);
}

double orc_fpr_blocked(uint64_t m, uint64_t n, double k, uint64_t B, int self_collisions, double eps) {
  double f = 0, c = (double)m / (double)n, lambda = (double)B / c;
  double k_act = k;
  if (self_collisions) k_act = (double)B * (1.0 - pow(1.0 - 1.0 / (double)B, k_act));
  double d_sum = 0.0;
  for (uint64_t i = 0; (d_sum + eps) < 1.0; i++) {
    const double d = poisson_pmf(lambda, i);
    d_sum += d;
    f += d * orc_fpr_bloom(B, i, k_act);
  }
  return f;
}

double orc_fpr_sectorized(uint64_t m, uint64_t n, double k, uint64_t B, uint64_t S, int self_collisions, double eps) {
  double f = 0, c = (double)m / (double)n, lambda = (double)B / c, s = (double)B / (double)S;
  double d_sum = 0.0;
  double k_per_s = k / s;
  if (self_collisions) k_per_s = (double)S * (1.0 - pow(1.0 - 1.0 / (double)S, k_per_s));
  for (uint64_t i = 0; (d_sum + eps) < 1.0; i++) {
    const double d = poisson_pmf(lambda, i);
    d_sum += d;
    f += d * pow(orc_fpr_bloom(S, i, k_per_s), s);
  }
  return f;
}

double orc_poisson_pmf(double lambda, uint64_t i) { return poisson_pmf(lambda, i); }

static double p_cache(uint64_t s, uint64_t S, uint64_t B, double k, uint64_t i) {
  const double k_per_s = k / (double)s;
  double sum = 0.0;
  for (uint64_t j = 1; j <= i; j++) {
    const double ev = ((double)i * (double)s * (double)S) / (double)B;
    sum += poisson_pmf(ev, j) * orc_fpr_bloom(S, j, k_per_s);
  }
  return pow(sum, (double)s);
}

double orc_fpr_p_cache(uint64_t s, uint64_t S, uint64_t B, double k, uint64_t i) { /* math.hpp:99-113 */
  return p_cache(s, S, B, k, i);
}

double orc_fpr_zoned(uint64_t m, uint64_t n, double k, uint64_t B, uint64_t S, double z, int self_collisions, double eps) {
  /* math.hpp:117-145; the self-collision k_per_s is computed but unused there */
  (void)self_collisions;
  double f = 0, c = (double)m / (double)n, lambda = (double)B / c;
  double d_sum = 0.0;
  for (uint64_t i = 0; (d_sum + eps) < 1.0; i++) {
    const double d = poisson_pmf(lambda, i);
    d_sum += d;
    f += d * p_cache((uint64_t)z, S, B, k, i);
  }
  return f;
}

double orc_fpr_config(uint64_t m, uint64_t n, uint32_t k, uint32_t word_bytes, uint32_t w, uint32_t s, uint32_t z) {
  /* fpr.hpp:11-60 with self_collisions = true and epsilon = 1e-5 */
  const uint64_t B = (uint64_t)word_bytes * 8 * w;
  const uint64_t S = B / s;
  const double eps = 0.00001;
  if (S == B && z == 1) return orc_fpr_blocked(m, n, k, B, 1, eps);
  if (S < B && z == 1) return orc_fpr_sectorized(m, n, k, B, S, 1, eps);
  if (S < B && z > 1) return orc_fpr_zoned(m, n, k, B, S, z, 1, eps);
  return -1.0;
}

double orc_fpr_cuckoo(uint32_t tpb, uint32_t tag_bits, double load) { /* math.hpp:198-205 */
  return 1.0 - pow(1.0 - 1.0 / (pow(2.0, tag_bits) - 1.0), 2.0 * tpb * load);
}
