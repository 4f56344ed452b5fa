"""Pins the divisor *choice* of MAGIC addressing: next_cheap_magic(n)
(src/dtl/div.hpp:29-57) returns the first d >= max(n, 2) for which libdivide's
u32 generator (thirdparty/libdivide/libdivide.h:719-768, branchfree = 0)
takes neither the power-of-two shift path nor the add path ("algorithm 1").
The block count of every MAGIC filter is that d, so it is a byte-level parity
fact.

The generator is restated below from libdivide.h on its own terms -- the
64/32 division of 2^(32 + floor_log2 d) by d, e = d - rem, the `more` byte
with its ADD_MARKER (0x40) and U32_SHIFT_PATH (0x80) flags -- and the search
from div.hpp; both the oracle (orc_next_cheap_magic) and the product's host
sizing (dtl_bbf_desc_init, dtl_cf_desc_init) must agree with it for every
n < 2^16 + 256, for every filter of the 4096-partition semi-join workload and
for sampled large n.  A search that starts at n + 1, accepts the add path,
or stops one divisor early fails here.  No GPU is needed.
"""
import ctypes

import numpy as np
import pytest

import bench
import dtl_hip as D
import oracle as O

ADD_MARKER = 0x40
U32_SHIFT_PATH = 0x80
SHIFT_MASK = 0x1F
D_MAX = 0xFFFFFFFF - 2  # div.hpp:18


def libdivide_64_div_32_to_32(u1, u0, v):
    """(u1 * 2^32 + u0) / v for u1 < v: 32-bit quotient and remainder."""
    num = (u1 << 32) | u0
    q, r = divmod(num, v)
    assert q < (1 << 32)
    return q, r


def libdivide_u32_gen(d):
    """(magic, more) of libdivide_internal_u32_gen(d, branchfree = 0)."""
    floor_log_2_d = d.bit_length() - 1
    if d & (d - 1) == 0:
        return 0, floor_log_2_d | U32_SHIFT_PATH
    proposed_m, rem = libdivide_64_div_32_to_32(1 << floor_log_2_d, 0, d)
    assert 0 < rem < d
    e = d - rem
    if e < (1 << floor_log_2_d):
        more = floor_log_2_d
    else:
        proposed_m += proposed_m
        twice_rem = rem + rem
        if twice_rem >= d or twice_rem < rem:
            proposed_m += 1
        more = floor_log_2_d | ADD_MARKER
    return (1 + proposed_m) & 0xFFFFFFFF, more


def next_cheap_magic(n, gen=libdivide_u32_gen):
    n = max(n, 2)
    d = n
    while d <= D_MAX:
        magic, more = gen(d)
        if not (more & ADD_MARKER) and not (more & U32_SHIFT_PATH):
            return magic, more & SHIFT_MASK, d
        d += 1
    raise AssertionError("no cheap divisor")


def _cheap_table(top):
    """cheap[d] for d < top, computed once so the exhaustive sweep is fast."""
    cheap = np.zeros(top, dtype=bool)
    magic = np.zeros(top, dtype=np.uint64)
    for d in range(2, top):
        mg, more = libdivide_u32_gen(d)
        if not (more & (ADD_MARKER | U32_SHIFT_PATH)):
            cheap[d] = True
            magic[d] = mg
    return cheap, magic


def product_magic(n, word_bits=32, w=16, z=2, k=6):
    """The product's MAGIC sizing for ceil(m / B) == n blocks."""
    bits = word_bits * w
    d = D.BbfDesc()
    rc = D.load().dtl_bbf_desc_init(ctypes.byref(d), word_bits, n * bits - 1, k, w, w, z)
    assert rc == 0, (n, rc)
    assert d.addr_mode == D.ADDR_MAGIC
    return d.magic, d.magic_shift, d.block_cnt


def test_restatement_reproduces_libdivide_algorithm_1_exactly():
    """The restated generator's no-add magics divide every tested dividend
    exactly (what libdivide's algorithm 1 promises), and add-path divisors
    exist in the range (so the search really skips some)."""
    rng = np.random.default_rng(5)
    skipped = 0
    for d in list(range(3, 3000)) + rng.integers(3, 2**32 - 3, 300).tolist():
        magic, more = libdivide_u32_gen(d)
        if more & U32_SHIFT_PATH:
            continue
        if more & ADD_MARKER:
            skipped += 1
            continue
        sh = more & SHIFT_MASK
        a = np.concatenate([rng.integers(0, 2**32, 2000, dtype=np.uint64),
                            np.array([0, 1, d - 1, d, 2**32 - 1, (2**32 // d) * d - 1], dtype=np.uint64)])
        q = ((a * np.uint64(magic)) >> np.uint64(32)) >> np.uint64(sh)
        assert np.array_equal(q, a // np.uint64(d)), d
    assert skipped > 100


def test_every_small_n_oracle_and_product():
    top = (1 << 16) + 256
    cheap, magic = _cheap_table(top + 64)
    nxt = np.full(top + 64, -1, dtype=np.int64)
    nearest = -1
    for d in range(top + 63, -1, -1):
        if cheap[d]:
            nearest = d
        nxt[d] = nearest
    for n in range(0, top):
        d = int(nxt[max(n, 2)])
        assert d >= max(n, 2)
        # minimality: every smaller candidate is a power of two or takes the add path
        for dp in range(max(n, 2), d):
            mg, more = libdivide_u32_gen(dp)
            assert more & (ADD_MARKER | U32_SHIFT_PATH), (n, dp)
        od = O.cheap_magic(n)
        assert (od.magic, od.shift, od.divisor) == (int(magic[d]), d.bit_length() - 1, d), n
        if n >= 2:
            assert product_magic(n) == (int(magic[d]), d.bit_length() - 1, d), n


def test_semi_join_workload_filters():
    """Every filter of the 4096-partition semi-join (u32 words, w = 16, z = 2,
    k = 6, m = 10 bits per build key): ceil(m / 512) -> the restated search."""
    for nb, _, _ in bench.semi_join_sizes(4096, 11):
        m = 10 * nb
        want = next_cheap_magic(-(-m // 512))
        d = D.BbfDesc()
        assert D.load().dtl_bbf_desc_init(ctypes.byref(d), 32, m, 6, 16, 16, 2) == 0
        assert d.addr_mode == D.ADDR_MAGIC
        assert (d.magic, d.magic_shift, d.block_cnt) == want, nb
        od = O.cheap_magic(-(-m // 512))
        assert (od.magic, od.shift, od.divisor) == want


@pytest.mark.parametrize("seed", [1, 2])
def test_sampled_large_n(seed):
    rng = np.random.default_rng(seed)
    ns = rng.integers(1 << 16, 1 << 31, 400).tolist() + [(1 << 31) - 1, (1 << 30) + 1, 3 << 28]
    checked = 0
    for n in ns:
        want = next_cheap_magic(n)
        od = O.cheap_magic(n)
        assert (od.magic, od.shift, od.divisor) == want, n
        # cuckoo tables share the addressing: 8-byte buckets, m not a power of two
        cd = D.CfDesc()
        if want[2] <= (1 << 31) and n * 64 < (1 << 40):
            rc = D.load().dtl_cf_desc_init(ctypes.byref(cd), n * 64 - 1, 16, 4)
            if rc == 0:
                assert (cd.magic, cd.magic_shift, cd.bucket_cnt) == want, n
                checked += 1
    assert checked > 300
