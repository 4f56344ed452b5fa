// dtl_sched.hpp -- compile-time description of which filter bits a key
// touches, shared by the host library and the gfx950 kernels.
//
// The blocked Bloom filter of the reference consumes its hash functions in a
// fixed, key-independent schedule: hash 0 addresses the block, hashes 1.. are
// cut MSB-first into fixed-width chunks, and a hash is replaced by the next
// one when fewer bits remain than a chunk needs.  Here that schedule is
// evaluated once at compile time per configuration (word size, words per
// block w, sectors s, zones z, k) into a table of chunks; the kernels unroll
// over the table so every shift and mask is an immediate.
//
// Reference templates the walk follows:
//   word_block            blocked_bloomfilter_block_logic_sgew.hpp:43-92
//   k=8 specialization    blocked_bloomfilter_block_logic_sgew_specialization.hpp:35-167
//                         (computes its hashes in locals and does not write the
//                         caller's hash value back; modelled by `stale`)
//   multiword_block       blocked_bloomfilter_block_logic_sgew.hpp:226-376  (s >= w)
//   multiword_sector /
//   multisector_block     blocked_bloomfilter_block_logic_sltw.hpp:47-528   (s < w)
//   multizone_block       blocked_bloomfilter_block_logic_zoned.hpp:50-210
//   first in-block hash   blocked_bloomfilter.hpp:46 (index 1)
#pragma once

#include <cstdint>
#include <type_traits>
#include <utility>

#ifndef DTL_HD
#if defined(__HIPCC__)
#define DTL_HD __host__ __device__
#else
#define DTL_HD
#endif
#endif

namespace dtlh {

// Multipliers of the hash family (hash_family.hpp:22-122): h_i(x) = x * kMul[i].
struct MulTable {
  uint32_t v[17];
};
constexpr MulTable kMul = {{596572387u, 370248451u, 2654435769u, 1799596469u, 0x9e3779b1u,
                            2284105051u, 1203114875u, 1150766481u, 2729912477u, 1884591559u,
                            770785867u, 2667333959u, 1550580529u, 0xcc9e2d51u, 0x1b873593u,
                            0x85ebca6bu, 0xc2b2ae35u}};

// A run of hash bits: ((H[src] >> shift) & ((1 << nbits) - 1)) + base.
// H[0] is the constant 0: the value a hash variable holds before the first
// rehash, which the k=8 specialization can leave behind.
struct Chunk {
  uint8_t src = 0;
  uint8_t shift = 0;
  uint8_t nbits = 0;
  uint8_t base = 0;
};

// Bits are grouped by the word they land in: sgew has one group per word
// (static word index), sltw one group per bit (word chosen by hash bits
// inside its sector), zoned one group per zone (word chosen once per zone).
struct Sched {
  int ngroups = 0;
  int nbits = 0;
  int max_src = 0;
  uint8_t gword[16] = {};   // first candidate word of the group inside the block
  Chunk gdyn[16] = {};      // dynamic word offset (nbits == 0: static word)
  uint8_t bgroup[16] = {};  // group of each bit
  Chunk bbit[16] = {};      // bit index inside the group's word
};

constexpr uint32_t clog2(uint32_t v) {
  uint32_t r = 0;
  while ((1u << (r + 1)) <= v) ++r;
  return r;
}

struct WalkState {
  uint32_t fn;   // next hash function to rehash with
  uint32_t rem;  // unconsumed bits of the current hash
  uint8_t hv;    // hash slot the runtime hash variable holds (0 = zero)
};

constexpr void walk_word_block(Sched& S, WalkState& st, uint32_t word_bits, uint32_t s_,
                               uint32_t k_, int group) {
  const uint32_t sector_bits = word_bits / s_;
  const uint32_t bpk = clog2(sector_bits);
  const uint32_t kps = k_ / s_;
  const uint8_t entry_hv = st.hv;
  for (uint32_t i = 0; i < k_; ++i) {
    if (st.rem < bpk) {
      st.hv = static_cast<uint8_t>(st.fn);
      st.fn += 1;
      st.rem = 32;
    }
    Chunk c;
    c.src = st.hv;
    c.shift = static_cast<uint8_t>(st.rem - bpk);
    c.nbits = static_cast<uint8_t>(bpk);
    c.base = static_cast<uint8_t>((i / kps) * sector_bits);
    S.bgroup[S.nbits] = static_cast<uint8_t>(group);
    S.bbit[S.nbits] = c;
    S.nbits += 1;
    if (c.src > S.max_src) S.max_src = c.src;
    st.rem -= bpk;
  }
  // The k=8 specialization leaves the caller's hash value untouched.
  if (k_ == 8) st.hv = entry_hv;
}

constexpr Sched make_sched(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t z, uint32_t k) {
  Sched S{};
  WalkState st{1u, 0u, 0u};
  if (z > 1) {
    const uint32_t wpz = w / z;
    const uint32_t zb = clog2(wpz);
    for (uint32_t zi = 0; zi < z; ++zi) {
      if (st.rem < zb) {
        st.hv = static_cast<uint8_t>(st.fn);
        st.fn += 1;
        st.rem = 32;
      }
      const int g = S.ngroups++;
      S.gword[g] = static_cast<uint8_t>(zi * wpz);
      Chunk d;
      d.src = st.hv;
      d.shift = static_cast<uint8_t>(st.rem - zb);
      d.nbits = static_cast<uint8_t>(zb);
      S.gdyn[g] = d;
      if (d.src > S.max_src) S.max_src = d.src;
      st.rem -= zb;
      walk_word_block(S, st, word_bits, 1u, k / z, g);
    }
  } else if (s >= w) {
    for (uint32_t j = 0; j < w; ++j) {
      const int g = S.ngroups++;
      S.gword[g] = static_cast<uint8_t>(j);
      walk_word_block(S, st, word_bits, s / w, k / w, g);
    }
  } else {
    const uint32_t wps = w / s;
    const uint32_t wl = clog2(wps);
    const uint32_t bl = clog2(word_bits);
    for (uint32_t sec = 0; sec < s; ++sec) {
      for (uint32_t i = 0; i < k / s; ++i) {
        if (st.rem < wl + bl) {
          st.hv = static_cast<uint8_t>(st.fn);
          st.fn += 1;
          st.rem = 32;
        }
        const int g = S.ngroups++;
        S.gword[g] = static_cast<uint8_t>(sec * wps);
        Chunk d;
        d.src = st.hv;
        d.shift = static_cast<uint8_t>(st.rem - wl);
        d.nbits = static_cast<uint8_t>(wl);
        S.gdyn[g] = d;
        Chunk b;
        b.src = st.hv;
        b.shift = static_cast<uint8_t>(st.rem - wl - bl);
        b.nbits = static_cast<uint8_t>(bl);
        S.bgroup[S.nbits] = static_cast<uint8_t>(g);
        S.bbit[S.nbits] = b;
        S.nbits += 1;
        if (st.hv > S.max_src) S.max_src = st.hv;
        st.rem -= wl + bl;
      }
    }
  }
  return S;
}

// Lane-split form of a schedule.  A block of 32 B or more is read by
// `lanes` = block bytes / 16 lanes, one 16-byte piece each, and every lane
// tests only the bits whose word lies in its piece.  The pieces fall into
// `lanes / span` sets of `span` pieces, and every set holds the same pattern
// of groups (same first word relative to the set, same dynamic word-offset
// width, same bits per group with the same widths and bases); only the hash
// sources and shifts differ from set to set, and a lane folds those into
// per-lane multipliers (split_fold) once per kernel.
//   span 1: every group's candidate words stay inside one piece (s >= w,
//           sectors or zones of at most 16 B);
//   span > 1 (zoned filters with zones of 32 B or more): every zone's
//           candidate words are exactly the span pieces of its set; the lane
//           of sub-piece q tests the zone only when its dynamic word offset d
//           has d / (words per piece) == q, and passes it otherwise, so the
//           AND over the key's lanes is the block test.
// Every lane of a set hashes and selects all of the set's groups.  A zone is
// one group, so zoned filters always take the spanning form (measured on
// MI355X, 2^26 probes: u32 w16 z2 k6 1.15-1.47x, u64 w16 z2 k8 1.42-1.54x
// over one lane per key).  A sector of 32 B or more holds k/s groups; the
// spanning form measured 1.04-1.27x at k/s <= 4 but 8-60% slower at 8 and 16,
// and a split form that came and went with k inside one (word, w, s) family
// is a step the learned cost model smooths over, so sectors keep span 1.
struct SplitSched {
  bool ok = false;
  int lanes = 1;           // 16-byte pieces per block
  int span = 1;            // pieces per set
  int wpp = 0;             // words per piece
  int gpp = 0;             // groups per set
  int bpg = 0;             // bits per group
  uint8_t lword[16] = {};  // first word of local group gl, relative to its set
  uint8_t dbits[16] = {};  // dynamic word-offset width of local group gl
  uint8_t bbits[16] = {};  // width of local bit gl * bpg + i
  uint8_t bbase[16] = {};  // base of local bit gl * bpg + i
};

constexpr SplitSched make_split(const Sched& S, uint32_t word_bytes, uint32_t w, bool zoned) {
  SplitSched P{};
  const uint32_t bytes = w * word_bytes;
  if (bytes < 32 || S.ngroups == 0) return P;
  const int L = static_cast<int>(bytes / 16);
  const int wpp = static_cast<int>(16 / word_bytes);
  const int gw = 1 << S.gdyn[0].nbits;  // candidate words of a group
  const int span = gw > wpp ? gw / wpp : 1;
  if (span > L || L % span != 0) return P;
  const int sets = L / span;
  if (S.ngroups % sets != 0 || S.nbits % S.ngroups != 0) return P;
  const int gpp = S.ngroups / sets, bpg = S.nbits / S.ngroups;
  if (span > 1 && !zoned) return P;
  for (int b = 0; b < S.nbits; ++b)
    if (S.bgroup[b] != b / bpg) return P;
  for (int g = 0; g < S.ngroups; ++g) {
    const int set = g / gpp, gl = g % gpp;
    const int first = S.gword[g], words = 1 << S.gdyn[g].nbits, last = first + words - 1;
    if (span == 1) {
      if (first / wpp != set || last / wpp != set) return P;
    } else if (words != span * wpp || first != set * span * wpp) {
      return P;
    }
    const int lw = first - set * span * wpp;
    if (set == 0) {
      P.lword[gl] = static_cast<uint8_t>(lw);
      P.dbits[gl] = S.gdyn[g].nbits;
    } else if (P.lword[gl] != lw || P.dbits[gl] != S.gdyn[g].nbits) {
      return P;
    }
    for (int i = 0; i < bpg; ++i) {
      const Chunk c = S.bbit[g * bpg + i];
      const int j = gl * bpg + i;
      if (c.nbits == 0) return P;
      if (set == 0) {
        P.bbits[j] = c.nbits;
        P.bbase[j] = c.base;
      } else if (P.bbits[j] != c.nbits || P.bbase[j] != c.base) {
        return P;
      }
    }
  }
  P.ok = true;
  P.lanes = L;
  P.span = span;
  P.wpp = wpp;
  P.gpp = gpp;
  P.bpg = bpg;
  return P;
}

// ((H[src] >> shift) & ((1 << nbits) - 1)) == (key * split_fold(c)) >> (32 - nbits)
// for nbits >= 1: shifting the multiplier left by 32 - nbits - shift moves the
// chunk's bits of the product to its top (H[0] = 0 folds to multiplier 0).
constexpr uint32_t split_fold(Chunk c) {
  return c.src == 0 ? 0u : kMul.v[c.src] << (32u - c.nbits - c.shift);
}

// Validity of a configuration exactly as the reference's runtime dispatch
// accepts it (blocked_bloomfilter.hpp:169-229; zoned_blocked_bloomfilter.hpp:102-215).
constexpr bool valid_sgew_sltw(uint32_t word_bits, uint32_t w, uint32_t s, uint32_t k) {
  if (word_bits != 32 && word_bits != 64) return false;
  if (w != 1 && w != 2 && w != 4 && w != 8 && w != 16) return false;
  if (s != 1 && s != 2 && s != 4 && s != 8 && s != 16) return false;
  if (s > w * (word_bits / 8)) return false;
  if (k < 1 || k > 16 || (k % s) != 0) return false;
  return true;
}

constexpr bool valid_zoned(uint32_t word_bits, uint32_t w, uint32_t z, uint32_t k) {
  if (word_bits != 32 && word_bits != 64) return false;
  if (w != 4 && w != 8 && w != 16) return false;
  if (z != 2 && z != 4 && z != 8) return false;
  if (z >= w) return false;
  if (k < 2 || k > 16 || (k & 1u) || (k % z) != 0) return false;
  return true;
}

template <class W, uint32_t Wc, uint32_t Sc, uint32_t Zc, uint32_t Kc>
struct BbfCfg {
  using word_t = W;
  static constexpr uint32_t word_bits = sizeof(W) * 8;
  static constexpr uint32_t w = Wc;
  static constexpr uint32_t s = Sc;
  static constexpr uint32_t z = Zc;
  static constexpr uint32_t k = Kc;
  static constexpr Sched sched = make_sched(word_bits, Wc, Sc, Zc, Kc);
  static constexpr SplitSched split = make_split(sched, sizeof(W), Wc, Zc > 1);
  static_assert(Zc > 1 ? valid_zoned(word_bits, Wc, Zc, Kc) : valid_sgew_sltw(word_bits, Wc, Sc, Kc),
                "unsupported configuration");
};

template <class F, int... I>
DTL_HD constexpr void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}

// Calls f(std::integral_constant<int, i>) for i in [0, N).
template <int N, class F>
DTL_HD constexpr void static_for(F&& f) {
  static_for_impl(static_cast<F&&>(f), std::make_integer_sequence<int, N>{});
}

}  // namespace dtlh
