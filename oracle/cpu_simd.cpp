// cpu_simd.cpp -- multithreaded AVX-512 restatement of the reference's
// batch_contains, used only as the CPU baseline that bench.py times beside
// the GPU (TEST/BENCH INFRASTRUCTURE ONLY; the product never links it).
//
// It follows the reference's SIMD probe structure
// (blocked_bloomfilter_batch_dispatch.hpp:28-69, cuckoofilter_logic.hpp:440-479):
// keys are processed in batches of 1024 (batchwise.hpp:9-21), a SIMD body of
// 8 keys per AVX-512 step gathers the words (vector_helper.hpp:122-161),
// tests all bits and compress-stores the matching positions in ascending
// order (intrin_avx512.hpp:40-46); each thread probes a contiguous slice into
// its own output, like the reference's benchmark (benchmark.cpp:123-164).
// Configurations without a hand-vectorized body use the scalar oracle per
// slice.  Every path is checked against the scalar oracle by tests/.
#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

constexpr uint32_t kC0 = 596572387u, kC1 = 370248451u, kC2 = 2654435769u;
constexpr size_t kBatch = 1024;

bool have_avx512() {
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
         __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512dq");
}

// Cache-sectorized BBF, u64 words, w=8, s=8, k=8, POW2 (bits per word:
// words 0..4 from h1 at shifts 26,20,14,8,2; words 5..7 from h2 at 26,20,14).
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t bbf_w8s8k8_avx512(const uint64_t* words, uint32_t shift, const uint32_t* keys, uint64_t n,
                           uint32_t* out, uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0));
  const __m256i c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i c2 = _mm256_set1_epi32(static_cast<int>(kC2));
  const __m512i one = _mm512_set1_epi64(1);
  const __m512i m63 = _mm512_set1_epi64(63);
  const __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i h0 = _mm256_mullo_epi32(k, c0);
    const __m256i blk = _mm256_srli_epi32(h0, static_cast<int>(shift));
    const __m512i base = _mm512_slli_epi64(_mm512_cvtepu32_epi64(blk), 3);
    const __m512i h1 = _mm512_cvtepu32_epi64(_mm256_mullo_epi32(k, c1));
    const __m512i h2 = _mm512_cvtepu32_epi64(_mm256_mullo_epi32(k, c2));
    __mmask8 ok = 0xFF;
    static const int sh1[5] = {26, 20, 14, 8, 2};
    static const int sh2[3] = {26, 20, 14};
    for (int j = 0; j < 8; ++j) {
      const __m512i idx = _mm512_add_epi64(base, _mm512_set1_epi64(j));
      const __m512i word = _mm512_i64gather_epi64(idx, reinterpret_cast<const long long*>(words), 8);
      const __m512i hv = j < 5 ? h1 : h2;
      const int sh = j < 5 ? sh1[j] : sh2[j - 5];
      const __m512i bit = _mm512_and_si512(_mm512_srli_epi64(hv, static_cast<unsigned>(sh)), m63);
      const __m512i mask = _mm512_sllv_epi64(one, bit);
      ok &= _mm512_cmpeq_epi64_mask(_mm512_and_si512(word, mask), mask);
    }
    const __m256i pos = _mm256_add_epi32(lane, _mm256_set1_epi32(static_cast<int>(i + offset)));
    _mm256_mask_compressstoreu_epi32(w, ok, pos);
    w += __builtin_popcount(ok);
  }
  return static_cast<uint64_t>(w - out);
}

// Cuckoo <16,4>, POW2: two 8-byte buckets per key, SWAR 16-bit compare.
__attribute__((target("avx512f,avx512bw,avx512vl,avx512dq")))
uint64_t cf_16_4_avx512(const uint8_t* buf, uint32_t shift, uint32_t bmask, uint32_t v_index,
                        uint32_t v_tag, int v_used, const uint32_t* keys, uint64_t n, uint32_t* out,
                        uint32_t offset) {
  uint32_t* w = out;
  const __m256i c0 = _mm256_set1_epi32(static_cast<int>(kC0));
  const __m256i c1 = _mm256_set1_epi32(static_cast<int>(kC1));
  const __m256i ffff = _mm256_set1_epi32(0xFFFF);
  const __m256i alt = _mm256_set1_epi32(0x5bd1e995);
  const __m256i bm = _mm256_set1_epi32(static_cast<int>(bmask));
  const __m512i rep = _mm512_set1_epi64(0x0001000100010001ll);
  const __m512i hi = _mm512_set1_epi64(static_cast<long long>(0x8000800080008000ull));
  const __m256i lane = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
  const __m256i vi = _mm256_set1_epi32(static_cast<int>(v_index));
  const __m256i vt = _mm256_set1_epi32(static_cast<int>(v_tag));
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256i k = _mm256_loadu_si256(reinterpret_cast<const __m256i*>(keys + i));
    const __m256i i1 = _mm256_srli_epi32(_mm256_mullo_epi32(k, c0), static_cast<int>(shift));
    __m256i tag = _mm256_and_si256(_mm256_mullo_epi32(k, c1), ffff);
    const __mmask8 z = _mm256_cmpeq_epi32_mask(tag, _mm256_setzero_si256());
    tag = _mm256_mask_add_epi32(tag, z, tag, _mm256_set1_epi32(1));
    const __m256i i2 = _mm256_and_si256(_mm256_xor_si256(i1, _mm256_mullo_epi32(tag, alt)), bm);
    const __m512i b1 = _mm512_i32gather_epi64(i1, reinterpret_cast<const long long*>(buf), 8);
    const __m512i b2 = _mm512_i32gather_epi64(i2, reinterpret_cast<const long long*>(buf), 8);
    const __m512i t4 = _mm512_mullo_epi64(_mm512_cvtepu32_epi64(tag), rep);
    const __m512i x1 = _mm512_xor_si512(b1, t4);
    const __m512i x2 = _mm512_xor_si512(b2, t4);
    const __m512i z1 = _mm512_and_si512(_mm512_andnot_si512(x1, _mm512_sub_epi64(x1, rep)), hi);
    const __m512i z2 = _mm512_and_si512(_mm512_andnot_si512(x2, _mm512_sub_epi64(x2, rep)), hi);
    __mmask8 hit = _mm512_test_epi64_mask(_mm512_or_si512(z1, z2), _mm512_or_si512(z1, z2));
    if (v_used) {
      const __mmask8 vt_eq = _mm256_cmpeq_epi32_mask(tag, vt);
      const __mmask8 vi_eq = _mm256_cmpeq_epi32_mask(i1, vi) | _mm256_cmpeq_epi32_mask(i2, vi);
      hit |= vt_eq & vi_eq;
    }
    const __m256i pos = _mm256_add_epi32(lane, _mm256_set1_epi32(static_cast<int>(i + offset)));
    _mm256_mask_compressstoreu_epi32(w, hit, pos);
    w += __builtin_popcount(hit);
  }
  return static_cast<uint64_t>(w - out);
}

// One slice: 1024-key batches; SIMD body of whole 8-key groups, scalar tail.
uint64_t bbf_slice(const orc_bbf* f, const void* words, const uint32_t* keys, uint64_t n,
                   uint32_t* out, uint32_t offset, bool simd) {
  uint64_t c = 0;
  for (uint64_t b = 0; b < n; b += kBatch) {
    const uint64_t len = std::min<uint64_t>(kBatch, n - b);
    uint64_t done = 0;
    if (simd) {
      const uint64_t body = len & ~uint64_t(7);
      c += bbf_w8s8k8_avx512(static_cast<const uint64_t*>(words), 32u - f->addr.block_cnt_log2,
                             keys + b, body, out + c, static_cast<uint32_t>(offset + b));
      done = body;
    }
    c += orc_bbf_batch_contains(f, words, keys + b + done, len - done, out + c,
                                static_cast<uint32_t>(offset + b + done));
  }
  return c;
}

uint64_t cf_slice(const orc_cf* f, const void* buf, const uint32_t* keys, uint64_t n, uint32_t* out,
                  uint32_t offset, bool simd) {
  uint64_t c = 0;
  const uint8_t* vp = static_cast<const uint8_t*>(buf) + f->victim_offset;
  uint32_t v_index, v_tag;
  std::memcpy(&v_index, vp, 4);
  std::memcpy(&v_tag, vp + 4, 4);
  for (uint64_t b = 0; b < n; b += kBatch) {
    const uint64_t len = std::min<uint64_t>(kBatch, n - b);
    uint64_t done = 0;
    if (simd) {
      const uint64_t body = len & ~uint64_t(7);
      c += cf_16_4_avx512(static_cast<const uint8_t*>(buf), 32u - f->addr.block_cnt_log2,
                          f->addr.block_cnt_mask, v_index, v_tag, vp[8] != 0, keys + b, body,
                          out + c, static_cast<uint32_t>(offset + b));
      done = body;
    }
    c += orc_cf_batch_contains(f, buf, keys + b + done, len - done, out + c,
                               static_cast<uint32_t>(offset + b + done));
  }
  return c;
}

template <class F>
uint64_t run_threads(int threads, uint64_t n, uint32_t* const* outs, uint64_t* counts, F&& fn) {
  threads = std::max(1, threads);
  std::vector<std::thread> pool;
  const uint64_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const uint64_t b = std::min<uint64_t>(n, per * t), e = std::min<uint64_t>(n, b + per);
    pool.emplace_back([&, t, b, e] { counts[t] = fn(b, e - b, outs[t]); });
  }
  for (auto& th : pool) th.join();
  uint64_t total = 0;
  for (int t = 0; t < threads; ++t) total += counts[t];
  return total;
}

}  // namespace

extern "C" {

// 1 if the hand-vectorized body covers this configuration on this CPU.
int cpub_bbf_simd_available(const orc_bbf* f) {
  return have_avx512() && f->word_bits == 64 && f->w == 8 && f->s == 8 && f->z == 1 && f->k == 8 &&
         f->addr.mode == ORC_POW2;
}

int cpub_cf_simd_available(const orc_cf* f) {
  return have_avx512() && f->tag_bits == 16 && f->tags_per_bucket == 4 && f->addr.mode == ORC_POW2;
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

// Probes keys[0..n) with `threads` threads; thread t writes the positions of
// its slice (ascending, offset + index) to outs[t] and its count to counts[t].
// Returns the total count and the wall time of the parallel region.
uint64_t cpub_bbf_probe(const orc_bbf* f, const void* words, const uint32_t* keys, uint64_t n,
                        int threads, int simd, uint32_t* const* outs, uint64_t* counts,
                        double* seconds) {
  const bool use_simd = simd && cpub_bbf_simd_available(f);
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t c = run_threads(threads, n, outs, counts, [&](uint64_t b, uint64_t len, uint32_t* out) {
    return bbf_slice(f, words, keys + b, len, out, static_cast<uint32_t>(b), use_simd);
  });
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return c;
}

uint64_t cpub_cf_probe(const orc_cf* f, const void* buf, const uint32_t* keys, uint64_t n,
                       int threads, int simd, uint32_t* const* outs, uint64_t* counts,
                       double* seconds) {
  const bool use_simd = simd && cpub_cf_simd_available(f);
  const auto t0 = std::chrono::steady_clock::now();
  const uint64_t c = run_threads(threads, n, outs, counts, [&](uint64_t b, uint64_t len, uint32_t* out) {
    return cf_slice(f, buf, keys + b, len, out, static_cast<uint32_t>(b), use_simd);
  });
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return c;
}

}  // extern "C"
