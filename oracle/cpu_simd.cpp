// cpu_simd.cpp -- multithreaded AVX-512 / AVX2 restatement of the reference's
// batch_contains, used as the CPU baseline that bench.py times beside the
// GPU, plus a threaded build used as a checker (TEST/BENCH INFRASTRUCTURE
// ONLY; the product never links it).
//
// It follows the reference's SIMD probe structure
// (blocked_bloomfilter_batch_dispatch.hpp:28-69, cuckoofilter_logic.hpp:440-479):
// keys go in batches of 1024 (batchwise.hpp:9-21); a SIMD body gathers the
// words of a vector of keys (vector_helper.hpp:16-161), tests every bit and
// turns the match mask into ascending positions -- AVX-512 with a
// compress-store (intrin_avx512.hpp:40-46), AVX2 with a lookup table of
// positions per 8-bit (or 4-bit) mask, stored whole, and a popcount
// (intrin_avx2.hpp:63-75, 138-147); a scalar tail takes the rest.  Workers
// are pinned, one per physical core of the process's CPU set (the reference
// pins worker i to CPU i, thread.hpp:100-109, 193-207, and runs one worker
// per core, platform.cpp:22); each probes a contiguous slice into its own
// output, like the reference's benchmark (benchmark.cpp:123-164).
//
// Hand-vectorized configurations (every other one runs the scalar oracle per
// slice): the four benchmark filters --
//   u32 w=1 s=1 k=4 POW2        bits of h1 at shifts 27, 22, 17, 12 (5 bits)
//   u64 w=8 s=8 k=8 POW2        word j: h1 at 26, 20, 14, 8, 2, then h2 at 26, 20, 14 (6 bits)
//   u32 w=16 z=2 k=6 MAGIC      zone 0: word bits(h1,29,3), bits h1 at 24, 19, 14;
//                               zone 1: word 8 + bits(h1,11,3), bits h1 at 6, 1, h2 at 27
//   cuckoo 16-bit tags, 4 slots, POW2 (SWAR compare of both buckets)
// Every body is checked against the scalar oracle by tests/test_cpu_baseline.py.
#include <immintrin.h>
#include <pthread.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

constexpr uint32_t kC0 = 596572387u, kC1 = 370248451u, kC2 = 2654435769u;
constexpr size_t kBatch = 1024;

enum Isa { kScalar = 0, kAvx2 = 1, kAvx512 = 2 };

bool have_avx512() {
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
         __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512dq");
}

bool have_avx2() { return __builtin_cpu_supports("avx2"); }

// ---------------------------------------------------------------------------
// AVX2 mask -> positions: for every 8-bit mask the ascending lane numbers of
// its set bits (16-bit, padded with zeros), built once.
// ---------------------------------------------------------------------------
struct PosLut {
  alignas(16) uint16_t pos8[256][8];
  uint32_t pos4[16][4];
  PosLut() {
    for (int m = 0; m < 256; ++m) {
      int c = 0;
      for (int b = 0; b < 8; ++b)
        if (m >> b & 1) pos8[m][c++] = static_cast<uint16_t>(b);
      for (; c < 8; ++c) pos8[m][c] = 0;
    }
    for (int m = 0; m < 16; ++m) {
      int c = 0;
      for (int b = 0; b < 4; ++b)
        if (m >> b & 1) pos4[m][c++] = static_cast<uint32_t>(b);
      for (; c < 4; ++c) pos4[m][c] = 0;
    }
  }
};
const PosLut kLut;

__attribute__((target("avx2"))) inline uint32_t* store_positions8(uint32_t* w, uint32_t mask, uint32_t base) {
  const __m128i p16 = _mm_load_si128(reinterpret_cast<const __m128i*>(kLut.pos8[mask]));
  const __m256i p = _mm256_add_epi32(_mm256_cvtepu16_epi32(p16), _mm256_set1_epi32(static_cast<int>(base)));
  _mm256_storeu_si256(reinterpret_cast<__m256i*>(w), p);
  return w + __builtin_popcount(mask);
}

__attribute__((target("avx2"))) inline uint32_t* store_positions4(uint32_t* w, uint32_t mask, uint32_t base) {
  const __m128i p = _mm_add_epi32(_mm_loadu_si128(reinterpret_cast<const __m128i*>(kLut.pos4[mask])),
                                  _mm_set1_epi32(static_cast<int>(base)));
  _mm_storeu_si128(reinterpret_cast<__m128i*>(w), p);
  return w + __builtin_popcount(mask);
}

// ---------------------------------------------------------------------------
// u32 w=1 s=1 k=4 (register-blocked), POW2
// ---------------------------------------------------------------------------
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t w1k4_avx512(const uint32_t* words, uint32_t shift, const uint32_t* keys, uint64_t n, uint32_t* out,
                     uint32_t offset) {
  uint32_t* w = out;
  const __m512i c0 = _mm512_set1_epi32(static_cast<int>(kC0)), c1 = _mm512_set1_epi32(static_cast<int>(kC1));
  const __m512i m31 = _mm512_set1_epi32(31), one = _mm512_set1_epi32(1);
  const __m512i lane = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 16 <= n; i += 16) {
    const __m512i k = _mm512_loadu_si512(keys + i);
    const __m512i blk = _mm512_srl_epi32(_mm512_mullo_epi32(k, c0), sh);
    const __m512i word = _mm512_i32gather_epi32(blk, words, 4);
    const __m512i h1 = _mm512_mullo_epi32(k, c1);
    __m512i mask = _mm512_sllv_epi32(one, _mm512_srli_epi32(h1, 27));
    mask = _mm512_or_si512(mask, _mm512_sllv_epi32(one, _mm512_and_si512(_mm512_srli_epi32(h1, 22), m31)));
    mask = _mm512_or_si512(mask, _mm512_sllv_epi32(one, _mm512_and_si512(_mm512_srli_epi32(h1, 17), m31)));
    mask = _mm512_or_si512(mask, _mm512_sllv_epi32(one, _mm512_and_si512(_mm512_srli_epi32(h1, 12), m31)));
    const __mmask16 ok = _mm512_cmpeq_epi32_mask(_mm512_and_si512(word, mask), mask);
    _mm512_mask_compressstoreu_epi32(w, ok, _mm512_add_epi32(lane, _mm512_set1_epi32(static_cast<int>(i + offset))));
    w += __builtin_popcount(ok);
  }
  return static_cast<uint64_t>(w - out);
}

__attribute__((target("avx2")))
uint64_t w1k4_avx2(const uint32_t* words, uint32_t shift, const uint32_t* keys, uint64_t n, uint32_t* out,
                   uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0)), c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i m31 = _mm256_set1_epi32(31), one = _mm256_set1_epi32(1);
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i blk = _mm256_srl_epi32(_mm256_mullo_epi32(k, c0), sh);
    const __m256i word = _mm256_i32gather_epi32(reinterpret_cast<const int*>(words), blk, 4);
    const __m256i h1 = _mm256_mullo_epi32(k, c1);
    __m256i mask = _mm256_sllv_epi32(one, _mm256_srli_epi32(h1, 27));
    mask = _mm256_or_si256(mask, _mm256_sllv_epi32(one, _mm256_and_si256(_mm256_srli_epi32(h1, 22), m31)));
    mask = _mm256_or_si256(mask, _mm256_sllv_epi32(one, _mm256_and_si256(_mm256_srli_epi32(h1, 17), m31)));
    mask = _mm256_or_si256(mask, _mm256_sllv_epi32(one, _mm256_and_si256(_mm256_srli_epi32(h1, 12), m31)));
    const __m256i eq = _mm256_cmpeq_epi32(_mm256_and_si256(word, mask), mask);
    const uint32_t bits = static_cast<uint32_t>(_mm256_movemask_ps(_mm256_castsi256_ps(eq)));
    w = store_positions8(w, bits, static_cast<uint32_t>(i + offset));
  }
  return static_cast<uint64_t>(w - out);
}

// ---------------------------------------------------------------------------
// u64 w=8 s=8 k=8 (cache-sectorized), POW2
// ---------------------------------------------------------------------------
constexpr int kSh8[8] = {26, 20, 14, 8, 2, 26, 20, 14};  // word j's shift; words 5..7 from h2

__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t w8s8k8_avx512(const uint64_t* words, uint32_t shift, const uint32_t* keys, uint64_t n, uint32_t* out,
                       uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0));
  const __m256i c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i c2 = _mm256_set1_epi32(static_cast<int>(kC2));
  const __m512i one = _mm512_set1_epi64(1), m63 = _mm512_set1_epi64(63);
  const __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i blk = _mm256_srl_epi32(_mm256_mullo_epi32(k, c0), sh);
    const __m512i base = _mm512_slli_epi64(_mm512_cvtepu32_epi64(blk), 3);
    const __m512i h1 = _mm512_cvtepu32_epi64(_mm256_mullo_epi32(k, c1));
    const __m512i h2 = _mm512_cvtepu32_epi64(_mm256_mullo_epi32(k, c2));
    __mmask8 ok = 0xFF;
    for (int j = 0; j < 8; ++j) {
      const __m512i idx = _mm512_add_epi64(base, _mm512_set1_epi64(j));
      const __m512i word = _mm512_i64gather_epi64(idx, reinterpret_cast<const long long*>(words), 8);
      const __m512i bit = _mm512_and_si512(_mm512_srl_epi64(j < 5 ? h1 : h2, _mm_cvtsi32_si128(kSh8[j])), m63);
      const __m512i mask = _mm512_sllv_epi64(one, bit);
      ok &= _mm512_cmpeq_epi64_mask(_mm512_and_si512(word, mask), mask);
    }
    _mm256_mask_compressstoreu_epi32(w, ok, _mm256_add_epi32(lane, _mm256_set1_epi32(static_cast<int>(i + offset))));
    w += __builtin_popcount(ok);
  }
  return static_cast<uint64_t>(w - out);
}

__attribute__((target("avx2")))
uint64_t w8s8k8_avx2(const uint64_t* words, uint32_t shift, const uint32_t* keys, uint64_t n, uint32_t* out,
                     uint32_t offset) {
  uint32_t* w = out;
  const __m128i c0 = _mm_set1_epi32(static_cast<int>(kC0));
  const __m128i c1 = _mm_set1_epi32(static_cast<int>(kC1));
  const __m128i c2 = _mm_set1_epi32(static_cast<int>(kC2));
  const __m256i one = _mm256_set1_epi64x(1), m63 = _mm256_set1_epi64x(63);
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 4 <= n; i += 4) {   // 4 keys per step: one 64-bit lane per key
    const __m128i k = _mm_loadu_si128(reinterpret_cast<const __m128i*>(keys + i));
    const __m128i blk = _mm_srl_epi32(_mm_mullo_epi32(k, c0), sh);
    const __m128i base = _mm_slli_epi32(blk, 3);   // word index < 2^31 for filters up to 16 GiB
    const __m256i h1 = _mm256_cvtepu32_epi64(_mm_mullo_epi32(k, c1));
    const __m256i h2 = _mm256_cvtepu32_epi64(_mm_mullo_epi32(k, c2));
    __m256i ok = _mm256_set1_epi64x(-1);
    for (int j = 0; j < 8; ++j) {
      const __m128i idx = _mm_add_epi32(base, _mm_set1_epi32(j));
      const __m256i word = _mm256_i32gather_epi64(reinterpret_cast<const long long*>(words), idx, 8);
      const __m256i bit = _mm256_and_si256(_mm256_srl_epi64(j < 5 ? h1 : h2, _mm_cvtsi32_si128(kSh8[j])), m63);
      const __m256i mask = _mm256_sllv_epi64(one, bit);
      ok = _mm256_and_si256(ok, _mm256_cmpeq_epi64(_mm256_and_si256(word, mask), mask));
    }
    const uint32_t bits = static_cast<uint32_t>(_mm256_movemask_pd(_mm256_castsi256_pd(ok)));
    w = store_positions4(w, bits, static_cast<uint32_t>(i + offset));
  }
  return static_cast<uint64_t>(w - out);
}

// ---------------------------------------------------------------------------
// u32 w=16 z=2 k=6 (zoned), MAGIC: block = h0 - (mulhi(h0, magic) >> ms) * d
// ---------------------------------------------------------------------------
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
inline __m512i mulhi32_avx512(__m512i a, __m512i b) {
  const __m512i even = _mm512_srli_epi64(_mm512_mul_epu32(a, b), 32);
  const __m512i odd = _mm512_mul_epu32(_mm512_srli_epi64(a, 32), _mm512_srli_epi64(b, 32));
  return _mm512_mask_blend_epi32(0xAAAA, even, odd);   // odd lanes: high half of the odd products
}

__attribute__((target("avx2"))) inline __m256i mulhi32_avx2(__m256i a, __m256i b) {
  const __m256i even = _mm256_srli_epi64(_mm256_mul_epu32(a, b), 32);
  const __m256i odd = _mm256_mul_epu32(_mm256_srli_epi64(a, 32), _mm256_srli_epi64(b, 32));
  return _mm256_blend_epi32(even, odd, 0xAA);
}

// 1 << ((h >> s) & 31) per lane
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
inline __m512i bit512(__m512i one, __m512i m31, __m512i h, int s) {
  return _mm512_sllv_epi32(one, _mm512_and_si512(_mm512_srl_epi32(h, _mm_cvtsi32_si128(s)), m31));
}

__attribute__((target("avx2"))) inline __m256i bit256(__m256i one, __m256i m31, __m256i h, int s) {
  return _mm256_sllv_epi32(one, _mm256_and_si256(_mm256_srl_epi32(h, _mm_cvtsi32_si128(s)), m31));
}

__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t zoned16k6_avx512(const uint32_t* words, uint32_t magic, uint32_t mshift, uint32_t divisor,
                          const uint32_t* keys, uint64_t n, uint32_t* out, uint32_t offset) {
  uint32_t* w = out;
  const __m512i c0 = _mm512_set1_epi32(static_cast<int>(kC0)), c1 = _mm512_set1_epi32(static_cast<int>(kC1));
  const __m512i c2 = _mm512_set1_epi32(static_cast<int>(kC2));
  const __m512i mg = _mm512_set1_epi32(static_cast<int>(magic)), dv = _mm512_set1_epi32(static_cast<int>(divisor));
  const __m512i m31 = _mm512_set1_epi32(31), m7 = _mm512_set1_epi32(7), one = _mm512_set1_epi32(1);
  const __m512i eight = _mm512_set1_epi32(8);
  const __m512i lane = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  const __m128i ms = _mm_cvtsi32_si128(static_cast<int>(mshift));
  uint64_t i = 0;
  for (; i + 16 <= n; i += 16) {
    const __m512i k = _mm512_loadu_si512(keys + i);
    const __m512i h0 = _mm512_mullo_epi32(k, c0);
    const __m512i q = _mm512_srl_epi32(mulhi32_avx512(h0, mg), ms);
    const __m512i blk = _mm512_sub_epi32(h0, _mm512_mullo_epi32(q, dv));
    const __m512i base = _mm512_slli_epi32(blk, 4);   // 16 words per block (filters below 2^28 words)
    const __m512i h1 = _mm512_mullo_epi32(k, c1), h2 = _mm512_mullo_epi32(k, c2);
    const __m512i wi0 = _mm512_add_epi32(base, _mm512_srli_epi32(h1, 29));
    const __m512i wi1 = _mm512_add_epi32(base, _mm512_add_epi32(eight, _mm512_and_si512(_mm512_srli_epi32(h1, 11), m7)));
    const __m512i w0 = _mm512_i32gather_epi32(wi0, words, 4);
    const __m512i w1 = _mm512_i32gather_epi32(wi1, words, 4);
    const __m512i k0 = _mm512_or_si512(_mm512_or_si512(bit512(one, m31, h1, 24), bit512(one, m31, h1, 19)),
                                       bit512(one, m31, h1, 14));
    const __m512i k1 = _mm512_or_si512(_mm512_or_si512(bit512(one, m31, h1, 6), bit512(one, m31, h1, 1)),
                                       bit512(one, m31, h2, 27));
    const __mmask16 ok = _mm512_cmpeq_epi32_mask(_mm512_and_si512(w0, k0), k0) &
                         _mm512_cmpeq_epi32_mask(_mm512_and_si512(w1, k1), k1);
    _mm512_mask_compressstoreu_epi32(w, ok, _mm512_add_epi32(lane, _mm512_set1_epi32(static_cast<int>(i + offset))));
    w += __builtin_popcount(ok);
  }
  return static_cast<uint64_t>(w - out);
}

__attribute__((target("avx2")))
uint64_t zoned16k6_avx2(const uint32_t* words, uint32_t magic, uint32_t mshift, uint32_t divisor,
                        const uint32_t* keys, uint64_t n, uint32_t* out, uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0)), c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i c2 = _mm256_set1_epi32(static_cast<int>(kC2));
  const __m256i mg = _mm256_set1_epi32(static_cast<int>(magic)), dv = _mm256_set1_epi32(static_cast<int>(divisor));
  const __m256i m31 = _mm256_set1_epi32(31), m7 = _mm256_set1_epi32(7), one = _mm256_set1_epi32(1);
  const __m256i eight = _mm256_set1_epi32(8);
  const __m128i ms = _mm_cvtsi32_si128(static_cast<int>(mshift));
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i h0 = _mm256_mullo_epi32(k, c0);
    const __m256i q = _mm256_srl_epi32(mulhi32_avx2(h0, mg), ms);
    const __m256i blk = _mm256_sub_epi32(h0, _mm256_mullo_epi32(q, dv));
    const __m256i base = _mm256_slli_epi32(blk, 4);
    const __m256i h1 = _mm256_mullo_epi32(k, c1), h2 = _mm256_mullo_epi32(k, c2);
    const __m256i wi0 = _mm256_add_epi32(base, _mm256_srli_epi32(h1, 29));
    const __m256i wi1 = _mm256_add_epi32(base, _mm256_add_epi32(eight, _mm256_and_si256(_mm256_srli_epi32(h1, 11), m7)));
    const __m256i w0 = _mm256_i32gather_epi32(reinterpret_cast<const int*>(words), wi0, 4);
    const __m256i w1 = _mm256_i32gather_epi32(reinterpret_cast<const int*>(words), wi1, 4);
    const __m256i k0 = _mm256_or_si256(_mm256_or_si256(bit256(one, m31, h1, 24), bit256(one, m31, h1, 19)),
                                       bit256(one, m31, h1, 14));
    const __m256i k1 = _mm256_or_si256(_mm256_or_si256(bit256(one, m31, h1, 6), bit256(one, m31, h1, 1)),
                                       bit256(one, m31, h2, 27));
    const __m256i eq = _mm256_and_si256(_mm256_cmpeq_epi32(_mm256_and_si256(w0, k0), k0),
                                        _mm256_cmpeq_epi32(_mm256_and_si256(w1, k1), k1));
    const uint32_t bits = static_cast<uint32_t>(_mm256_movemask_ps(_mm256_castsi256_ps(eq)));
    w = store_positions8(w, bits, static_cast<uint32_t>(i + offset));
  }
  return static_cast<uint64_t>(w - out);
}

// ---------------------------------------------------------------------------
// Cuckoo <16,4>, POW2: two 8-byte buckets per key, SWAR 16-bit compare
// (bitsutil.hpp:45-47 hasvalue16), victim cache entry.
// ---------------------------------------------------------------------------
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t cf_16_4_avx512(const uint8_t* buf, uint32_t shift, uint32_t bmask, uint32_t v_index, uint32_t v_tag,
                        int v_used, const uint32_t* keys, uint64_t n, uint32_t* out, uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0)), c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i ffff = _mm256_set1_epi32(0xFFFF), alt = _mm256_set1_epi32(0x5bd1e995);
  const __m256i bm = _mm256_set1_epi32(static_cast<int>(bmask));
  const __m512i rep = _mm512_set1_epi64(0x0001000100010001ll);
  const __m512i hi = _mm512_set1_epi64(static_cast<long long>(0x8000800080008000ull));
  const __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
  const __m256i vi = _mm256_set1_epi32(static_cast<int>(v_index)), vt = _mm256_set1_epi32(static_cast<int>(v_tag));
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i i1 = _mm256_srl_epi32(_mm256_mullo_epi32(k, c0), sh);
    __m256i tag = _mm256_and_si256(_mm256_mullo_epi32(k, c1), ffff);
    const __mmask8 z = _mm256_cmpeq_epi32_mask(tag, _mm256_setzero_si256());
    tag = _mm256_mask_add_epi32(tag, z, tag, _mm256_set1_epi32(1));
    const __m256i i2 = _mm256_and_si256(_mm256_xor_si256(i1, _mm256_mullo_epi32(tag, alt)), bm);
    const __m512i b1 = _mm512_i32gather_epi64(i1, reinterpret_cast<const long long*>(buf), 8);
    const __m512i b2 = _mm512_i32gather_epi64(i2, reinterpret_cast<const long long*>(buf), 8);
    const __m512i t4 = _mm512_mullo_epi64(_mm512_cvtepu32_epi64(tag), rep);
    const __m512i x1 = _mm512_xor_si512(b1, t4), x2 = _mm512_xor_si512(b2, t4);
    const __m512i z1 = _mm512_and_si512(_mm512_andnot_si512(x1, _mm512_sub_epi64(x1, rep)), hi);
    const __m512i z2 = _mm512_and_si512(_mm512_andnot_si512(x2, _mm512_sub_epi64(x2, rep)), hi);
    const __m512i zz = _mm512_or_si512(z1, z2);
    __mmask8 hit = _mm512_test_epi64_mask(zz, zz);
    if (v_used)
      hit |= _mm256_cmpeq_epi32_mask(tag, vt) & (_mm256_cmpeq_epi32_mask(i1, vi) | _mm256_cmpeq_epi32_mask(i2, vi));
    _mm256_mask_compressstoreu_epi32(w, hit, _mm256_add_epi32(lane, _mm256_set1_epi32(static_cast<int>(i + offset))));
    w += __builtin_popcount(hit);
  }
  return static_cast<uint64_t>(w - out);
}

__attribute__((target("avx2")))
uint64_t cf_16_4_avx2(const uint8_t* buf, uint32_t shift, uint32_t bmask, uint32_t v_index, uint32_t v_tag,
                      int v_used, const uint32_t* keys, uint64_t n, uint32_t* out, uint32_t offset) {
  uint32_t* w = out;
  const __m128i c0 = _mm_set1_epi32(static_cast<int>(kC0)), c1 = _mm_set1_epi32(static_cast<int>(kC1));
  const __m128i ffff = _mm_set1_epi32(0xFFFF), alt = _mm_set1_epi32(0x5bd1e995);
  const __m128i bm = _mm_set1_epi32(static_cast<int>(bmask));
  const __m256i rep = _mm256_set1_epi64x(0x0001000100010001ll);
  const __m256i hi = _mm256_set1_epi64x(static_cast<long long>(0x8000800080008000ull));
  const __m128i vi = _mm_set1_epi32(static_cast<int>(v_index)), vt = _mm_set1_epi32(static_cast<int>(v_tag));
  const __m128i sh = _mm_cvtsi32_si128(static_cast<int>(shift));
  uint64_t i = 0;
  for (; i + 4 <= n; i += 4) {   // 4 keys per step: one 64-bit bucket lane per key
    const __m128i k = _mm_loadu_si128(reinterpret_cast<const __m128i*>(keys + i));
    const __m128i i1 = _mm_srl_epi32(_mm_mullo_epi32(k, c0), sh);
    __m128i tag = _mm_and_si128(_mm_mullo_epi32(k, c1), ffff);
    tag = _mm_sub_epi32(tag, _mm_cmpeq_epi32(tag, _mm_setzero_si128()));   // 0 -> 1
    const __m128i i2 = _mm_and_si128(_mm_xor_si128(i1, _mm_mullo_epi32(tag, alt)), bm);
    const __m256i b1 = _mm256_i32gather_epi64(reinterpret_cast<const long long*>(buf), i1, 8);
    const __m256i b2 = _mm256_i32gather_epi64(reinterpret_cast<const long long*>(buf), i2, 8);
    // the tag in all four 16-bit slots of a 64-bit lane
    const __m256i t64 = _mm256_cvtepu32_epi64(tag);
    const __m256i t4 = _mm256_or_si256(_mm256_or_si256(t64, _mm256_slli_epi64(t64, 16)),
                                       _mm256_or_si256(_mm256_slli_epi64(t64, 32), _mm256_slli_epi64(t64, 48)));
    const __m256i x1 = _mm256_xor_si256(b1, t4), x2 = _mm256_xor_si256(b2, t4);
    const __m256i z1 = _mm256_and_si256(_mm256_andnot_si256(x1, _mm256_sub_epi64(x1, rep)), hi);
    const __m256i z2 = _mm256_and_si256(_mm256_andnot_si256(x2, _mm256_sub_epi64(x2, rep)), hi);
    const __m256i nz = _mm256_cmpeq_epi64(_mm256_or_si256(z1, z2), _mm256_setzero_si256());
    uint32_t bits = static_cast<uint32_t>(~_mm256_movemask_pd(_mm256_castsi256_pd(nz))) & 0xFu;
    if (v_used) {
      const __m128i v = _mm_and_si128(_mm_cmpeq_epi32(tag, vt),
                                      _mm_or_si128(_mm_cmpeq_epi32(i1, vi), _mm_cmpeq_epi32(i2, vi)));
      bits |= static_cast<uint32_t>(_mm_movemask_ps(_mm_castsi128_ps(v)));
    }
    w = store_positions4(w, bits, static_cast<uint32_t>(i + offset));
  }
  return static_cast<uint64_t>(w - out);
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
enum BbfKind { kNone, kW1K4, kW8S8K8, kZoned16K6 };

BbfKind bbf_kind(const orc_bbf* f) {
  if (f->word_bits == 32 && f->w == 1 && f->s == 1 && f->z == 1 && f->k == 4 && f->addr.mode == ORC_POW2)
    return kW1K4;
  if (f->word_bits == 64 && f->w == 8 && f->s == 8 && f->z == 1 && f->k == 8 && f->addr.mode == ORC_POW2 &&
      f->addr.block_cnt <= (1u << 28))
    return kW8S8K8;
  if (f->word_bits == 32 && f->w == 16 && f->z == 2 && f->k == 6 && f->addr.mode == ORC_MAGIC &&
      f->addr.block_cnt <= (1u << 27))
    return kZoned16K6;
  return kNone;
}

// ISAs with a body for this configuration on this CPU: bit kAvx2 (1) and
// bit kAvx512 (2).
int bbf_isas(const orc_bbf* f) {
  if (bbf_kind(f) == kNone) return 0;
  return (have_avx2() ? 1 : 0) | (have_avx512() ? 2 : 0);
}

int cf_isas(const orc_cf* f) {
  if (!(f->tag_bits == 16 && f->tags_per_bucket == 4 && f->addr.mode == ORC_POW2)) return 0;
  return (have_avx2() ? 1 : 0) | (have_avx512() ? 2 : 0);
}

// SIMD body of one batch; returns matches and sets *done to the keys covered.
uint64_t bbf_body(const orc_bbf* f, const void* words, const uint32_t* keys, uint64_t len, uint32_t* out,
                  uint32_t offset, int isa, uint64_t* done) {
  const uint32_t shift = 32u - f->addr.block_cnt_log2;
  const uint32_t* w32 = static_cast<const uint32_t*>(words);
  const uint64_t* w64 = static_cast<const uint64_t*>(words);
  const orc_divisor& dv = f->addr.div;
  switch (bbf_kind(f)) {
    case kW1K4:
      *done = len & ~uint64_t(isa == kAvx512 ? 15 : 7);
      return isa == kAvx512 ? w1k4_avx512(w32, shift, keys, *done, out, offset)
                            : w1k4_avx2(w32, shift, keys, *done, out, offset);
    case kW8S8K8:
      *done = len & ~uint64_t(isa == kAvx512 ? 7 : 3);
      return isa == kAvx512 ? w8s8k8_avx512(w64, shift, keys, *done, out, offset)
                            : w8s8k8_avx2(w64, shift, keys, *done, out, offset);
    case kZoned16K6:
      *done = len & ~uint64_t(isa == kAvx512 ? 15 : 7);
      return isa == kAvx512 ? zoned16k6_avx512(w32, dv.magic, dv.shift, dv.divisor, keys, *done, out, offset)
                            : zoned16k6_avx2(w32, dv.magic, dv.shift, dv.divisor, keys, *done, out, offset);
    case kNone:
      break;
  }
  *done = 0;
  return 0;
}

// One slice: 1024-key batches; SIMD body of whole vectors, scalar tail.
uint64_t bbf_slice(const orc_bbf* f, const void* words, const uint32_t* keys, uint64_t n, uint32_t* out,
                   uint32_t offset, int isa) {
  uint64_t c = 0;
  for (uint64_t b = 0; b < n; b += kBatch) {
    const uint64_t len = std::min<uint64_t>(kBatch, n - b);
    uint64_t done = 0;
    if (isa != kScalar) c += bbf_body(f, words, keys + b, len, out + c, static_cast<uint32_t>(offset + b), isa, &done);
    c += orc_bbf_batch_contains(f, words, keys + b + done, len - done, out + c,
                                static_cast<uint32_t>(offset + b + done));
  }
  return c;
}

uint64_t cf_slice(const orc_cf* f, const void* buf, const uint32_t* keys, uint64_t n, uint32_t* out,
                  uint32_t offset, int isa) {
  uint64_t c = 0;
  const uint8_t* vp = static_cast<const uint8_t*>(buf) + f->victim_offset;
  uint32_t v_index, v_tag;
  std::memcpy(&v_index, vp, 4);
  std::memcpy(&v_tag, vp + 4, 4);
  const uint32_t shift = 32u - f->addr.block_cnt_log2;
  for (uint64_t b = 0; b < n; b += kBatch) {
    const uint64_t len = std::min<uint64_t>(kBatch, n - b);
    uint64_t done = 0;
    if (isa == kAvx512) {
      done = len & ~uint64_t(7);
      c += cf_16_4_avx512(static_cast<const uint8_t*>(buf), shift, f->addr.block_cnt_mask, v_index, v_tag,
                          vp[8] != 0, keys + b, done, out + c, static_cast<uint32_t>(offset + b));
    } else if (isa == kAvx2) {
      done = len & ~uint64_t(3);
      c += cf_16_4_avx2(static_cast<const uint8_t*>(buf), shift, f->addr.block_cnt_mask, v_index, v_tag,
                        vp[8] != 0, keys + b, done, out + c, static_cast<uint32_t>(offset + b));
    }
    c += orc_cf_batch_contains(f, buf, keys + b + done, len - done, out + c,
                               static_cast<uint32_t>(offset + b + done));
  }
  return c;
}

// One CPU per physical core of this process's CPU set (the first of each set
// of SMT siblings), in CPU order.
std::vector<int> core_cpus() {
  cpu_set_t mask;
  CPU_ZERO(&mask);
  std::vector<int> out;
  if (sched_getaffinity(0, sizeof(mask), &mask) != 0) return out;
  std::set<std::string> seen;
  for (int c = 0; c < CPU_SETSIZE; ++c) {
    if (!CPU_ISSET(c, &mask)) continue;
    char path[128];
    std::snprintf(path, sizeof(path), "/sys/devices/system/cpu/cpu%d/topology/thread_siblings_list", c);
    std::string sib = std::to_string(c);
    if (FILE* fh = std::fopen(path, "r")) {
      char buf[256] = {0};
      if (std::fgets(buf, sizeof(buf), fh)) sib = buf;
      std::fclose(fh);
    }
    if (seen.insert(sib).second) out.push_back(c);
  }
  return out;
}

int g_pin = 1;

template <class F>
uint64_t run_threads(int threads, uint64_t n, uint32_t* const* outs, uint64_t* counts, F&& fn) {
  threads = std::max(1, threads);
  static const std::vector<int> cpus = core_cpus();
  std::vector<std::thread> pool;
  const uint64_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const uint64_t b = std::min<uint64_t>(n, per * t), e = std::min<uint64_t>(n, b + per);
    pool.emplace_back([&, t, b, e] {
      if (g_pin && !cpus.empty()) {
        cpu_set_t m;
        CPU_ZERO(&m);
        CPU_SET(cpus[t % cpus.size()], &m);
        (void)pthread_setaffinity_np(pthread_self(), sizeof(m), &m);
      }
      counts[t] = fn(b, e - b, outs[t]);
    });
  }
  for (auto& th : pool) th.join();
  uint64_t total = 0;
  for (int t = 0; t < threads; ++t) total += counts[t];
  return total;
}

int pick_isa(int want, int avail) {
  // want: -1 = best available, 0 = scalar, 1 = AVX2, 2 = AVX-512
  if (want < 0) return (avail & 2) ? kAvx512 : (avail & 1) ? kAvx2 : kScalar;
  if (want == kAvx512) return (avail & 2) ? kAvx512 : kScalar;
  if (want == kAvx2) return (avail & 1) ? kAvx2 : kScalar;
  return kScalar;
}

}  // namespace

extern "C" {

// Bit 1: an AVX2 body covers this configuration on this CPU; bit 2: AVX-512.
int cpub_bbf_simd_available(const orc_bbf* f) { return bbf_isas(f); }
int cpub_cf_simd_available(const orc_cf* f) { return cf_isas(f); }

// Physical cores of this process's CPU set (workers are pinned to them).
int cpub_physical_cores(void) { return static_cast<int>(std::max<size_t>(1, core_cpus().size())); }

// Pin workers (1, the default) or leave them to the scheduler (0).
void cpub_set_pinning(int on) { g_pin = on ? 1 : 0; }

// Probes keys[0..n) with `threads` pinned threads using the body `isa`
// (-1 best available, 0 scalar, 1 AVX2, 2 AVX-512; a body this configuration
// or CPU lacks runs scalar); thread t writes the positions of its slice
// (ascending, offset + index) to outs[t] and its count to counts[t].  Returns
// the total count; *seconds is the wall time of the parallel region and
// *isa_used the body that ran.
uint64_t cpub_bbf_probe(const orc_bbf* f, const void* words, const uint32_t* keys, uint64_t n, int threads,
                        int isa, uint32_t* const* outs, uint64_t* counts, double* seconds, int* isa_used) {
  const int use = pick_isa(isa, bbf_isas(f));
  if (isa_used) *isa_used = use;
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t c = run_threads(threads, n, outs, counts, [&](uint64_t b, uint64_t len, uint32_t* out) {
    return bbf_slice(f, words, keys + b, len, out, static_cast<uint32_t>(b), use);
  });
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return c;
}

uint64_t cpub_cf_probe(const orc_cf* f, const void* buf, const uint32_t* keys, uint64_t n, int threads, int isa,
                       uint32_t* const* outs, uint64_t* counts, double* seconds, int* isa_used) {
  const int use = pick_isa(isa, cf_isas(f));
  if (isa_used) *isa_used = use;
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t c = run_threads(threads, n, outs, counts, [&](uint64_t b, uint64_t len, uint32_t* out) {
    return cf_slice(f, buf, keys + b, len, out, static_cast<uint32_t>(b), use);
  });
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return c;
}

// Many filters, each with its own key batch (the semi-join): `threads`
// pinned workers take whole tasks from a shared counter, in the given order,
// and probe each with one thread in 1024-key batches (positions from 0).
// Returns the total count; *seconds is the wall time of the parallel region.
uint64_t cpub_bbf_probe_tasks(uint32_t n_tasks, const orc_bbf* const* filters, const void* const* words,
                              const uint32_t* const* keys, const uint64_t* n, uint32_t* const* outs,
                              uint64_t* counts, int threads, int isa, double* seconds) {
  threads = std::max(1, threads);
  static const std::vector<int> cpus = core_cpus();
  std::atomic<uint32_t> next{0};
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      if (g_pin && !cpus.empty()) {
        cpu_set_t m;
        CPU_ZERO(&m);
        CPU_SET(cpus[t % cpus.size()], &m);
        (void)pthread_setaffinity_np(pthread_self(), sizeof(m), &m);
      }
      for (uint32_t i = next.fetch_add(1); i < n_tasks; i = next.fetch_add(1)) {
        const int use = pick_isa(isa, bbf_isas(filters[i]));
        counts[i] = bbf_slice(filters[i], words[i], keys[i], n[i], outs[i], 0, use);
      }
    });
  }
  for (auto& th : pool) th.join();
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  uint64_t total = 0;
  for (uint32_t i = 0; i < n_tasks; ++i) total += counts[i];
  return total;
}

// Builds the filter from keys[0..n) with `threads` threads (a checker for
// full-size GPU builds, not a baseline): thread t owns the blocks
// [t * B / T, (t + 1) * B / T), scans every key and inserts, with the scalar
// oracle's orc_bbf_insert, those whose block it owns.  A key only touches
// words of its own block, so no two threads write one word, and the result is
// the sequential build's (OR is order-independent).
void cpub_bbf_build(const orc_bbf* f, void* words, const uint32_t* keys, uint64_t n, int threads) {
  threads = std::max(1, threads);
  const uint64_t nb = f->addr.block_cnt;
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t) {
    const uint64_t lo = nb * t / threads, hi = nb * (t + 1) / threads;
    pool.emplace_back([=] {
      for (uint64_t i = 0; i < n; ++i) {
        const uint32_t b = orc_addr_block_idx(&f->addr, orc_hash(0, keys[i]));
        if (b >= lo && b < hi) orc_bbf_insert(f, words, keys[i]);
      }
    });
  }
  for (auto& th : pool) th.join();
}

}  // extern "C"
