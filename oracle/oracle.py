"""ctypes binding of the C oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(_HERE, "_build")

_lib = None
_cpu = None


class Divisor(C.Structure):
    _fields_ = [("magic", C.c_uint32), ("shift", C.c_uint32), ("divisor", C.c_uint32)]


class Addr(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("block_cnt", C.c_uint32), ("block_cnt_log2", C.c_uint32),
                ("block_cnt_mask", C.c_uint32), ("div", Divisor)]


class BBF(C.Structure):
    _fields_ = [("word_bits", C.c_uint32), ("w", C.c_uint32), ("s", C.c_uint32), ("z", C.c_uint32),
                ("k", C.c_uint32), ("m", C.c_uint64), ("m_actual", C.c_uint64),
                ("word_cnt", C.c_uint64), ("addr", Addr)]


class CF(C.Structure):
    _fields_ = [("tag_bits", C.c_uint32), ("tags_per_bucket", C.c_uint32),
                ("bytes_per_bucket", C.c_uint32), ("m", C.c_uint64), ("m_actual", C.c_uint64),
                ("table_words", C.c_uint64), ("victim_offset", C.c_uint64),
                ("buffer_bytes", C.c_uint64), ("addr", Addr)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_uint32)]


POW2, MAGIC = 0, 1
OK, EINVAL, ELOGIC = 0, -1, -3

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_BUILD, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run __graft_entry__.build() (missing %s)" % path)
        L = C.CDLL(path)
        L.orc_hash.restype = C.c_uint32
        L.orc_hash.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_next_cheap_magic.restype = C.c_int
        L.orc_next_cheap_magic.argtypes = [C.c_uint32, C.POINTER(Divisor)]
        L.orc_fast_mod.restype = C.c_uint32
        L.orc_fast_mod.argtypes = [C.c_uint32, Divisor]
        L.orc_fast_div.restype = C.c_uint32
        L.orc_fast_div.argtypes = [C.c_uint32, Divisor]
        L.orc_addr_init.restype = C.c_int
        L.orc_addr_init.argtypes = [C.POINTER(Addr), C.c_uint32, C.c_uint64]
        L.orc_addr_block_idx.restype = C.c_uint32
        L.orc_addr_block_idx.argtypes = [C.POINTER(Addr), C.c_uint32]
        L.orc_bbf_init.restype = C.c_int
        L.orc_bbf_init.argtypes = [C.POINTER(BBF), C.c_uint32, C.c_uint64, C.c_uint32,
                                   C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_bbf_insert.restype = None
        L.orc_bbf_insert.argtypes = [C.POINTER(BBF), C.c_void_p, C.c_uint32]
        L.orc_bbf_batch_insert.restype = None
        L.orc_bbf_batch_insert.argtypes = [C.POINTER(BBF), C.c_void_p, _u32p, C.c_uint64]
        L.orc_bbf_contains.restype = C.c_int
        L.orc_bbf_contains.argtypes = [C.POINTER(BBF), C.c_void_p, C.c_uint32]
        L.orc_bbf_contains_simd.restype = C.c_int
        L.orc_bbf_contains_simd.argtypes = [C.POINTER(BBF), C.c_void_p, C.c_uint32]
        L.orc_bbf_batch_contains.restype = C.c_uint64
        L.orc_bbf_batch_contains.argtypes = [C.POINTER(BBF), C.c_void_p, _u32p, C.c_uint64,
                                             _u32p, C.c_uint32]
        L.orc_bbf_batch_contains_mixed.restype = C.c_uint64
        L.orc_bbf_batch_contains_mixed.argtypes = [C.POINTER(BBF), C.c_void_p, _u32p, C.c_uint64,
                                                   _u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                   C.c_uint32]
        L.orc_bbf_key_bits.restype = C.c_int
        L.orc_bbf_key_bits.argtypes = [C.POINTER(BBF), C.c_uint32, _u64p, _u32p]
        L.orc_bbf_batch_key_bits.restype = C.c_uint64
        L.orc_bbf_batch_key_bits.argtypes = [C.POINTER(BBF), _u32p, C.c_uint64, _u64p, _u32p]
        L.orc_bbf_batch_blocks.restype = None
        L.orc_bbf_batch_blocks.argtypes = [C.POINTER(BBF), _u32p, C.c_uint64, _u32p]
        L.orc_bbf_batch_contains_blocks.restype = C.c_uint64
        L.orc_bbf_batch_contains_blocks.argtypes = [C.POINTER(BBF), C.c_void_p, _u32p, C.c_uint64, _u32p,
                                                    C.c_uint32]
        L.orc_bbf_has_k8_quirk.restype = C.c_int
        L.orc_bbf_has_k8_quirk.argtypes = [C.POINTER(BBF)]
        L.orc_mt_seed.restype = None
        L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_mt_next.argtypes = [C.POINTER(MT)]
        L.orc_cf_init.restype = C.c_int
        L.orc_cf_init.argtypes = [C.POINTER(CF), C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_cf_insert.restype = C.c_int
        L.orc_cf_insert.argtypes = [C.POINTER(CF), C.c_void_p, C.c_uint32, C.POINTER(MT)]
        L.orc_cf_contains.restype = C.c_int
        L.orc_cf_contains.argtypes = [C.POINTER(CF), C.c_void_p, C.c_uint32]
        L.orc_cf_batch_contains.restype = C.c_uint64
        L.orc_cf_batch_contains.argtypes = [C.POINTER(CF), C.c_void_p, _u32p, C.c_uint64, _u32p,
                                            C.c_uint32]
        for fn in ("orc_cf_tag", "orc_cf_i1"):
            getattr(L, fn).restype = C.c_uint32
            getattr(L, fn).argtypes = [C.POINTER(CF), C.c_uint32]
        L.orc_cf_alt.restype = C.c_uint32
        L.orc_cf_alt.argtypes = [C.POINTER(CF), C.c_uint32, C.c_uint32]
        L.orc_cf_read_tag.restype = C.c_uint32
        L.orc_cf_read_tag.argtypes = [C.POINTER(CF), C.c_void_p, C.c_uint64, C.c_uint32]
        L.orc_swar_hasvalue.restype = C.c_uint64
        L.orc_swar_hasvalue.argtypes = [C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_cf_count_occupied.restype = C.c_uint64
        L.orc_cf_count_occupied.argtypes = [C.POINTER(CF), C.c_void_p]
        L.orc_cf_key_candidates.restype = None
        L.orc_cf_key_candidates.argtypes = [C.POINTER(CF), _u32p, C.c_uint64, _u32p, _u32p, _u32p]
        L.orc_cf_entries.restype = C.c_uint64
        L.orc_cf_entries.argtypes = [C.POINTER(CF), C.c_void_p, _u32p, _u32p]
        L.orc_fpr_config.restype = C.c_double
        L.orc_fpr_config.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_uint32]
        L.orc_poisson_pmf.restype = C.c_double
        L.orc_poisson_pmf.argtypes = [C.c_double, C.c_uint64]
        L.orc_fpr_blocked.restype = C.c_double
        L.orc_fpr_blocked.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_int, C.c_double]
        L.orc_fpr_sectorized.restype = C.c_double
        L.orc_fpr_sectorized.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64,
                                         C.c_int, C.c_double]
        L.orc_fpr_zoned.restype = C.c_double
        L.orc_fpr_zoned.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64,
                                    C.c_double, C.c_int, C.c_double]
        L.orc_fpr_p_cache.restype = C.c_double
        L.orc_fpr_p_cache.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64]
        L.orc_fpr_cuckoo.restype = C.c_double
        L.orc_fpr_cuckoo.argtypes = [C.c_uint32, C.c_uint32, C.c_double]
        L.orc_fpr_bloom.restype = C.c_double
        L.orc_fpr_bloom.argtypes = [C.c_uint64, C.c_uint64, C.c_double]
        _lib = L
    return _lib


def hash32(fn, key):
    return lib().orc_hash(fn, key)


def cheap_magic(n):
    d = Divisor()
    rc = lib().orc_next_cheap_magic(n, C.byref(d))
    assert rc == 0
    return d


class Bloom:
    """A blocked Bloom filter as the reference's scalar code defines it."""

    def __init__(self, word_bits, m, k, w=1, s=1, z=1):
        self.d = BBF()
        self.rc = lib().orc_bbf_init(C.byref(self.d), word_bits, m, k, w, s, z)
        if self.rc != OK:
            raise ValueError("invalid configuration (rc=%d)" % self.rc)
        self.dtype = np.uint32 if word_bits == 32 else np.uint64

    @property
    def word_cnt(self):
        return self.d.word_cnt

    @property
    def m_actual(self):
        return self.d.m_actual

    @property
    def block_cnt(self):
        return self.d.addr.block_cnt

    @property
    def addr_mode(self):
        return self.d.addr.mode

    def new_words(self):
        return np.zeros(self.d.word_cnt, dtype=self.dtype)

    def insert(self, words, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        lib().orc_bbf_batch_insert(C.byref(self.d), words.ctypes.data, keys, len(keys))

    def contains(self, words, key):
        return bool(lib().orc_bbf_contains(C.byref(self.d), words.ctypes.data, int(key)))

    def contains_simd(self, words, key):
        return bool(lib().orc_bbf_contains_simd(C.byref(self.d), words.ctypes.data, int(key)))

    def batch_contains(self, words, keys, match_offset=0):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.zeros(max(len(keys), 1), dtype=np.uint32)
        c = lib().orc_bbf_batch_contains(C.byref(self.d), words.ctypes.data, keys, len(keys), out,
                                         match_offset)
        return out[:c].copy()

    def batch_contains_mixed(self, words, keys, match_offset, vector_len, align_bytes, addr_mod):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.zeros(max(len(keys), 1), dtype=np.uint32)
        c = lib().orc_bbf_batch_contains_mixed(C.byref(self.d), words.ctypes.data, keys, len(keys),
                                               out, match_offset, vector_len, align_bytes, addr_mod)
        return out[:c].copy()

    def key_bits(self, key):
        wi = np.zeros(64, dtype=np.uint64)
        bi = np.zeros(64, dtype=np.uint32)
        n = lib().orc_bbf_key_bits(C.byref(self.d), int(key), wi, bi)
        return list(zip(wi[:n].tolist(), bi[:n].tolist()))

    def sparse_words(self, keys):
        """(sorted word indexes, their words) of the filter orc_bbf_batch_insert
        builds from `keys`: every other word is zero."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        cap = max(len(keys) * 64, 1)
        wi = np.zeros(cap, dtype=np.uint64)
        bi = np.zeros(cap, dtype=np.uint32)
        c = lib().orc_bbf_batch_key_bits(C.byref(self.d), keys, len(keys), wi, bi)
        wi, bi = wi[:c], bi[:c]
        order = np.argsort(wi, kind="stable")
        wi, bi = wi[order], bi[order]
        masks = np.left_shift(np.ones(c, dtype=np.uint64), bi.astype(np.uint64))
        if c == 0:
            return wi, masks.astype(self.dtype)
        starts = np.concatenate([[0], np.nonzero(wi[1:] != wi[:-1])[0] + 1])
        vals = np.bitwise_or.reduceat(masks, starts)
        return wi[starts], vals.astype(self.dtype)

    def blocks_of(self, keys):
        """Block index of every key."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.zeros(max(len(keys), 1), dtype=np.uint32)
        lib().orc_bbf_batch_blocks(C.byref(self.d), keys, len(keys), out)
        return out[: len(keys)]

    def batch_contains_blocks(self, blocks, keys, match_offset=0):
        """batch_contains where blocks[i] (w words) is key i's block."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        blocks = np.ascontiguousarray(blocks, dtype=self.dtype)
        assert blocks.size == len(keys) * self.d.w
        out = np.zeros(max(len(keys), 1), dtype=np.uint32)
        c = lib().orc_bbf_batch_contains_blocks(C.byref(self.d), blocks.ctypes.data, keys, len(keys), out,
                                                match_offset)
        return out[:c].copy()

    def has_k8_quirk(self):
        return bool(lib().orc_bbf_has_k8_quirk(C.byref(self.d)))


def bbf_init_rc(word_bits, m, k, w=1, s=1, z=1):
    d = BBF()
    return lib().orc_bbf_init(C.byref(d), word_bits, m, k, w, s, z)


class MT19937:
    def __init__(self, seed=5489):
        self.g = MT()
        lib().orc_mt_seed(C.byref(self.g), seed)

    def next(self):
        return lib().orc_mt_next(C.byref(self.g))


class Cuckoo:
    """A cuckoo filter as the reference's scalar code defines it.  The buffer
    holds the packed buckets, then the 12-byte victim cache, then 4 pad bytes."""

    def __init__(self, m, tag_bits=16, tags_per_bucket=4):
        self.d = CF()
        rc = lib().orc_cf_init(C.byref(self.d), m, tag_bits, tags_per_bucket)
        if rc != OK:
            raise ValueError("invalid configuration (rc=%d)" % rc)

    @property
    def bucket_cnt(self):
        return self.d.addr.block_cnt

    @property
    def buffer_bytes(self):
        return self.d.buffer_bytes

    def new_buffer(self):
        return np.zeros((self.d.buffer_bytes + 7) // 8 * 8, dtype=np.uint8)

    def insert(self, buf, keys, rng):
        ok = True
        for k in np.asarray(keys, dtype=np.uint32).tolist():
            ok &= bool(lib().orc_cf_insert(C.byref(self.d), buf.ctypes.data, k, C.byref(rng.g)))
        return ok

    def contains(self, buf, key):
        return bool(lib().orc_cf_contains(C.byref(self.d), buf.ctypes.data, int(key)))

    def batch_contains(self, buf, keys, match_offset=0):
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        out = np.zeros(max(len(keys), 1), dtype=np.uint32)
        c = lib().orc_cf_batch_contains(C.byref(self.d), buf.ctypes.data, keys, len(keys), out,
                                        match_offset)
        return out[:c].copy()

    def tag(self, key):
        return lib().orc_cf_tag(C.byref(self.d), key)

    def i1(self, key):
        return lib().orc_cf_i1(C.byref(self.d), key)

    def alt(self, i, tag):
        return lib().orc_cf_alt(C.byref(self.d), i, tag)

    def read_tag(self, buf, i, j):
        return lib().orc_cf_read_tag(C.byref(self.d), buf.ctypes.data, i, j)

    def count_occupied(self, buf):
        return lib().orc_cf_count_occupied(C.byref(self.d), buf.ctypes.data)

    def victim(self, buf):
        """(index, tag, used) of the victim cache {u32 index; u32 tag; u8 used}."""
        off = self.d.victim_offset
        v = np.frombuffer(buf[off: off + 8].tobytes(), dtype=np.uint32)
        return int(v[0]), int(v[1]), bool(buf[off + 8])

    def placement_valid(self, buf, keys):
        """True iff the table holds exactly one tag per key, each in one of
        that key's two buckets (i1 or alt(i1, tag)) or in the victim cache
        with one of those indexes: a perfect matching between the keys and
        the occupied entries.  Holds for a table built by the reference's
        insert in any key order and with any kick choices, because a kick
        moves a tag between the two buckets alt() pairs for it
        (cuckoofilter_logic.hpp:291-322); it is what a probe relies on.
        Keys must be distinct."""
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import maximum_bipartite_matching
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        n = len(keys)
        i1 = np.zeros(max(n, 1), np.uint32)
        i2 = np.zeros_like(i1)
        tg = np.zeros_like(i1)
        lib().orc_cf_key_candidates(C.byref(self.d), keys, n, i1, i2, tg)
        i1, i2, tg = i1[:n], i2[:n], tg[:n]
        cap = self.bucket_cnt * self.d.tags_per_bucket
        eb = np.zeros(max(cap, 1), np.uint32)
        et = np.zeros_like(eb)
        ne = lib().orc_cf_entries(C.byref(self.d), buf.ctypes.data, eb, et)
        vi, vt, vused = self.victim(buf)
        if ne + int(vused) != n:
            return False
        if n == 0:
            return True
        code = (eb[:ne].astype(np.uint64) << np.uint64(32)) | et[:ne]
        order = np.argsort(code, kind="stable")
        code = code[order]
        rows, cols = [], []
        for bucket in (i1, i2.copy()):
            if bucket is not i1:
                bucket[i2 == i1] = np.uint32(0xFFFFFFFF)  # one bucket only: no second edge set
            c = (bucket.astype(np.uint64) << np.uint64(32)) | tg
            lo = np.searchsorted(code, c, "left")
            hi = np.searchsorted(code, c, "right")
            cnt = (hi - lo).astype(np.int64)
            r = np.repeat(np.arange(n), cnt)
            start = np.repeat(lo - np.cumsum(cnt) + cnt, cnt) + np.arange(cnt.sum())
            rows.append(r)
            cols.append(order[start])
        if vused:
            vk = np.nonzero((tg == vt) & ((i1 == vi) | (i2 == vi)))[0]
            rows.append(vk)
            cols.append(np.full(len(vk), ne, dtype=np.int64))
        r = np.concatenate(rows)
        c = np.concatenate(cols)
        g = csr_matrix((np.ones(len(r), np.int8), (r, c)), shape=(n, ne + int(vused)))
        match = maximum_bipartite_matching(g, perm_type="column")
        return bool((match >= 0).all())


def swar_hasvalue(bits, x, v):
    return lib().orc_swar_hasvalue(bits, x, v)


def fpr_config(m, n, k, word_bytes, w, s, z):
    return lib().orc_fpr_config(m, n, k, word_bytes, w, s, z)


def fpr_cuckoo(tags_per_bucket, tag_bits, load):
    return lib().orc_fpr_cuckoo(tags_per_bucket, tag_bits, load)
