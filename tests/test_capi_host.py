"""CPU-side checks of libdtl_hip: the library loads, exports every function
include/dtl_hip.h declares, and its host-side logic (descriptor sizing and
validation, the cuckoo host build) agrees with the oracle.  No kernel is
launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import dtl_hip as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            name = m.group(1)
            if name not in ("if", "sizeof", "return"):
                names.add(name)
    return names


def header_constants():
    """name -> value of every `DTL_NAME = value` enumerator in include/*.h."""
    vals = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", fn)).read(), flags=re.S)
        for m in re.finditer(r"\b(DTL_[A-Z0-9_]+)\s*=\s*(-?\d+)u?\b", text):
            vals[m.group(1)] = int(m.group(2))
    return vals


def test_python_constants_match_the_header():
    """Every status code, addressing mode and probe flag the Python binding
    names has the header's value, and every probe flag of the header has a
    Python name."""
    hv = header_constants()
    pairs = {"OK": "DTL_OK", "EINVAL": "DTL_EINVAL", "ERUNTIME": "DTL_ERUNTIME", "ELOGIC": "DTL_ELOGIC",
             "ENOMEM": "DTL_ENOMEM", "EHIP": "DTL_EHIP", "ENODEVICE": "DTL_ENODEVICE",
             "ADDR_POW2": "DTL_ADDR_POW2", "ADDR_MAGIC": "DTL_ADDR_MAGIC"}
    flags = [k for k in hv if k.startswith("DTL_PROBE_")]
    assert len(flags) >= 8
    for k in flags:
        pairs[k[len("DTL_"):]] = k
    for py, h in pairs.items():
        assert getattr(D, py) == hv[h], (py, h)
    vals = [hv[k] for k in flags]
    assert len(set(vals)) == len(vals) and all(v & (v - 1) == 0 for v in vals)


def test_library_exports_every_declared_symbol():
    L = D.load()
    names = declared_functions()
    assert {"bbf_build", "bbf_probe", "cf_probe", "bbf_probe_batched"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(D.EXPORTS) <= names


def test_version_and_status_strings():
    L = D.load()
    assert L.dtl_version() >= 100
    assert D.status_string(D.EINVAL) == "invalid argument"


def valid_configs():
    out = []
    for wb in (32, 64):
        for w in (1, 2, 4, 8, 16):
            for s in (1, 2, 4, 8, 16):
                if s > w * wb // 8:
                    continue
                for k in range(1, 17):
                    if k % s == 0:
                        out.append((wb, w, s, 1, k))
        for w in (4, 8, 16):
            for z in (2, 4, 8):
                if z >= w:
                    continue
                for k in range(2, 17, 2):
                    if k % z == 0:
                        out.append((wb, w, w, z, k))
    return out


def test_every_reference_configuration_is_compiled():
    """Every configuration the reference's runtime dispatch accepts
    (blocked_bloomfilter.hpp:169-229, zoned_blocked_bloomfilter.hpp:173-215)
    has a kernel."""
    L = D.load()
    cfgs = valid_configs()
    assert L.dtl_bbf_config_count() == len(cfgs)
    for wb, w, s, z, k in cfgs:
        assert L.dtl_bbf_supported(wb, w, s, z, k), (wb, w, s, z, k)
    assert not L.dtl_bbf_supported(32, 1, 8, 1, 8)
    assert not L.dtl_bbf_supported(32, 4, 4, 4, 8)


def test_lane_split_classification():
    """The lane-split kernel exists for blocks of at least 32 B, with lanes
    per key = block bytes / 16: for zoned filters always (zones of 32 B and
    up span whole pieces, each lane testing the words in its own piece), for
    the others where every word a bit can land in lies inside one 16-byte
    piece: s >= w (one sector per word or finer) or sectors of at most 16 B."""
    L = D.load()
    for wb, w, s, z, k in valid_configs():
        block = w * wb // 8
        if z > 1:
            ok = True
        else:
            ok = s >= w or (w // s) * wb // 8 <= 16
        want = block // 16 if block >= 32 and ok else 0
        assert L.dtl_bbf_split_lanes(wb, w, s, z, k) == want, (wb, w, s, z, k)
    assert L.dtl_bbf_split_lanes(64, 8, 8, 1, 8) == 4      # the cache-sectorized benchmark filter
    assert L.dtl_bbf_split_lanes(64, 8, 2, 1, 8) == 0      # 32 B sectors cross pieces
    assert L.dtl_bbf_split_lanes(32, 16, 16, 2, 6) == 4    # 32 B zones (the semi-join's filter)
    assert L.dtl_bbf_split_lanes(64, 16, 16, 2, 8) == 8    # 64 B zones of a 128 B block
    assert L.dtl_bbf_split_lanes(32, 4, 4, 1, 4) == 0      # 16 B blocks: one lane already
    assert L.dtl_bbf_split_lanes(32, 1, 8, 1, 8) == 0      # not a configuration


@pytest.mark.parametrize("cfg", [
    (32, 1024, 8, 1, 1, 1), (32, 1 << 20, 4, 1, 1, 1), (64, 1 << 17, 8, 8, 8, 1),
    (64, 1 << 33, 8, 8, 8, 1), (32, 10 * 1000, 6, 16, 16, 2), (32, 10 * (1 << 22), 6, 16, 16, 2),
    (64, 123457, 7, 4, 1, 1), (32, 3 * 512, 16, 16, 16, 8), (64, 99999, 16, 2, 16, 1),
])
def test_bbf_descriptor_matches_oracle(cfg):
    wb, m, k, w, s, z = cfg
    d = D.bbf_desc(wb, m, k, w, s, z)
    o = O.Bloom(wb, m, k, w, s, z)
    assert d.block_cnt == o.block_cnt
    assert d.addr_mode == o.addr_mode
    assert d.m_actual == o.m_actual
    assert d.word_cnt == o.word_cnt
    assert d.size_bytes == o.m_actual // 8
    if d.addr_mode == D.ADDR_MAGIC:
        assert d.magic == o.d.addr.div.magic and d.magic_shift == o.d.addr.div.shift
    else:
        assert d.block_shift == 32 - o.d.addr.block_cnt_log2


def test_bbf_validation_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(400):
        wb = int(rng.choice([32, 64]))
        w = int(rng.choice([1, 2, 3, 4, 8, 16, 32]))
        s = int(rng.choice([1, 2, 4, 8, 16, 3]))
        z = int(rng.choice([1, 1, 2, 4, 8, 3]))
        k = int(rng.integers(0, 18))
        m = int(rng.choice([1024, 1 << 20, 12345 * 8, 64, 10]))
        assert D.bbf_desc_rc(wb, m, k, w, s, z) == O.bbf_init_rc(wb, m, k, w, s, z), (wb, m, k, w, s, z)


@pytest.mark.parametrize("m,t,b", [(1 << 20, 16, 4), (1 << 32, 16, 4), (1000 * 64, 16, 4),
                                   (1 << 16, 8, 2), (77777, 12, 4), (1 << 18, 32, 1)])
def test_cf_descriptor_matches_oracle(m, t, b):
    d = D.cf_desc(m, t, b)
    o = O.Cuckoo(m, t, b)
    assert d.bucket_cnt == o.bucket_cnt
    assert d.buffer_bytes == o.buffer_bytes
    assert d.victim_offset == o.d.victim_offset
    assert d.m_actual == o.d.m_actual


@pytest.mark.parametrize("t,b,load", [(16, 4, 0.95), (8, 4, 0.9), (12, 4, 0.9), (4, 2, 0.5),
                                      (32, 2, 0.8), (16, 1, 0.4), (12, 2, 0.8)])
def test_cuckoo_host_build_matches_oracle(t, b, load):
    """The product's host insert (dtl_cuckoo_host.cpp) writes byte-identical
    tables to the oracle's restatement for the same kick-RNG seed, including
    the victim cache after overflow."""
    m = (1 << 14) * ((t * b + 7) // 8) * 8
    d = D.cf_desc(m, t, b)
    o = O.Cuckoo(m, t, b)
    rng = np.random.default_rng(42)
    n = int(load * d.bucket_cnt * b)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    buf, ins = D.cf_build_host(d, keys, seed=1234)
    obuf = o.new_buffer()
    g = O.MT19937(1234)
    oins = sum(1 for k in keys.tolist() if O.lib().orc_cf_insert(O.C.byref(o.d), obuf.ctypes.data, k,
                                                                  O.C.byref(g.g)))
    assert ins == oins
    assert np.array_equal(buf[: d.buffer_bytes], obuf[: d.buffer_bytes])


def test_cuckoo_host_build_overflow_sets_victim():
    d = D.cf_desc(64 * 64, 16, 4)
    keys = (np.arange(1, 5000, dtype=np.uint64) * 2654435761 % 2**32).astype(np.uint32)
    buf, ins = D.cf_build_host(d, keys, seed=7)
    assert ins < len(keys)
    assert buf[d.victim_offset + 8] == 1


def test_gather_bench_rejects_bad_arguments():
    """Argument checks of the roofline microbenchmark run before any GPU call."""
    L = D.load()
    fake = ctypes.c_void_p(4096)
    sink = ctypes.c_void_p(8192)
    assert L.dtl_gather_bench(fake, 1 << 20, 10, 1024, sink, None) == D.EINVAL  # modes 0..9
    assert L.dtl_gather_bench(fake, 1 << 20, 8, 1024, sink, None) == D.EINVAL   # LDS modes: <= 128 KiB
    assert L.dtl_gather_bench(fake, 1 << 20, 9, 1024, sink, None) == D.EINVAL
    assert L.dtl_gather_bench(fake, 32, 8, 1024, sink, None) == D.EINVAL        # LDS mode: >= 64 B
    assert L.dtl_gather_bench(fake, 64, 9, 1024, sink, None) == D.EINVAL        # two 64 B blocks at least
    assert L.dtl_gather_bench(fake, (1 << 20) + 64, 0, 1024, sink, None) == D.EINVAL  # not a power of two
    assert L.dtl_gather_bench(None, 1 << 20, 0, 1024, sink, None) == D.EINVAL


def test_cf_build_rejects_bad_arguments():
    """GPU cuckoo build: argument checks run before any GPU call; 12-bit tags
    with 2 or 4 slots (3- and 6-byte buckets straddle 4-byte words) are not
    built on the GPU."""
    L = D.load()
    fake = ctypes.c_void_p(4096)
    cnt = ctypes.c_void_p(8192)
    for t, b in ((12, 2), (12, 4)):
        d = D.cf_desc(1 << 16, t, b)
        assert L.cf_build(ctypes.byref(d), fake, fake, 10, 0, 0, cnt, None) == D.EINVAL
    d = D.cf_desc(1 << 16, 16, 4)
    assert L.cf_build(ctypes.byref(d), None, fake, 10, 0, 0, cnt, None) == D.EINVAL
    assert L.cf_build(ctypes.byref(d), fake, fake, 10, 0, 0, None, None) == D.EINVAL
    assert L.cf_build(ctypes.byref(d), ctypes.c_void_p(4100), fake, 10, 0, 0, cnt, None) == D.EINVAL  # 16 B align
    assert L.cf_build(ctypes.byref(d), fake, fake, 10, 0, (1 << 20) + 1, cnt, None) == D.EINVAL   # max_kicks cap
    assert L.cf_build(ctypes.byref(d), fake, None, 0, 0, 0, cnt, None) == D.OK                  # nothing to do


def test_python_binding_rejects_host_and_wrong_width_tensors():
    """The binding passes raw pointers to kernels, so it refuses host tensors
    and keys that are not 4-byte words before any GPU call."""
    torch = pytest.importorskip("torch")
    desc = D.bbf_desc(32, 1024, 8, 1, 1)
    host_words = torch.zeros(int(desc.size_bytes), dtype=torch.uint8)
    with pytest.raises(ValueError):
        D.bbf_probe(desc, host_words, torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError):
        D.bbf_build(desc, host_words, torch.zeros(4, dtype=torch.int64))
    cd = D.cf_desc(1 << 16, 16, 4)
    with pytest.raises(ValueError):
        D.cf_probe(cd, torch.zeros(int(cd.buffer_bytes), dtype=torch.uint8), torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError):
        D.cf_build(cd, torch.zeros(int(cd.buffer_bytes), dtype=torch.uint8), torch.zeros(4, dtype=torch.int32))


class _FakeDeviceTensor:
    """Looks like a contiguous GPU tensor to the binding's argument checks;
    its data pointer is never used because every case below must be refused
    before the library is called."""

    def __init__(self, numel, itemsize, dtype=None, device="cuda:0", shape=None):
        import torch
        self.is_cuda = True
        self._numel = numel
        self._itemsize = itemsize
        self.dtype = dtype if dtype is not None else {1: torch.uint8, 4: torch.int32, 8: torch.int64}[itemsize]
        self.device = torch.device(device)
        self.shape = shape if shape is not None else (numel,)

    def is_contiguous(self):
        return True

    def element_size(self):
        return self._itemsize

    def numel(self):
        return self._numel

    def dim(self):
        return len(self.shape)

    def data_ptr(self):
        raise AssertionError("the binding called into the library with an invalid argument")


def test_python_binding_rejects_undersized_and_mismatched_buffers():
    """Undersized filters, tables, position buffers and counts, and operands
    on different devices, raise ValueError without a library call."""
    pytest.importorskip("torch")
    desc = D.bbf_desc(64, 1 << 20, 8, 8, 8, 1)
    F = _FakeDeviceTensor
    keys = F(1000, 4)
    with pytest.raises(ValueError, match="words"):
        D.bbf_probe(desc, F(int(desc.size_bytes) - 1, 1), keys)
    with pytest.raises(ValueError, match="words"):
        D.bbf_build(desc, F(int(desc.size_bytes) // 2, 1), keys)
    words = F(int(desc.size_bytes), 1)
    with pytest.raises(ValueError, match="out"):
        D.bbf_probe(desc, words, keys, out=F(999, 4), count=F(1, 8))
    with pytest.raises(ValueError, match="count"):
        D.bbf_probe(desc, words, keys, out=F(1000, 4), count=F(1, 4))
    with pytest.raises(ValueError, match="keys"):
        D.bbf_probe(desc, words, F(1000, 4, device="cuda:1"), out=F(1000, 4), count=F(1, 8))
    with pytest.raises(ValueError, match="out"):
        D.bbf_probe(desc, words, keys, out=F(1000, 4, device="cuda:1"), count=F(1, 8))
    cd = D.cf_desc(1 << 20, 16, 4)
    with pytest.raises(ValueError, match="table"):
        D.cf_probe(cd, F(int(cd.buffer_bytes) - 8, 1), keys)
    with pytest.raises(ValueError, match="table"):
        D.cf_build(cd, F(int(cd.table_words) * 4, 1), keys)
    with pytest.raises(ValueError, match="out"):
        D.cf_probe(cd, F(int(cd.buffer_bytes), 1), keys, out=F(10, 4), count=F(1, 8))


def test_mlp_input_width_is_queried_and_documented():
    """The MLP reads rows of dtl_mlp_input_count() = 8 floats; the header
    says so, the parameter count follows from it, and the binding refuses
    feature matrices of any other width before a launch."""
    import torch
    L = D.load()
    n_in = L.dtl_mlp_input_count()
    assert n_in == 8
    assert L.dtl_mlp_param_count() == 64 * n_in + 64 + 64 * 64 + 64 + 64 + 1
    header = open(os.path.join(ROOT, "include", "dtl_hip.h")).read()
    assert "n x 8" in header and "W1[64][IN]" in header
    assert "n x 10" not in header and "W1[64][10]" not in header
    mlp = D.MLP.__new__(D.MLP)      # the argument checks only; creating one needs a GPU
    mlp.L, mlp.h, mlp.n_inputs, mlp.n_params = L, None, n_in, L.dtl_mlp_param_count()
    F = _FakeDeviceTensor
    for bad in (F(64 * 10, 4, torch.float32, shape=(64, 10)), F(64 * 7, 4, torch.float32, shape=(64, 7)),
                F(512, 4, torch.float32, shape=(512,)), F(512, 8, torch.float64, shape=(64, 8))):
        with pytest.raises(ValueError, match="x"):
            mlp.predict(bad)
    x = F(512, 4, torch.float32, shape=(64, 8))
    with pytest.raises(ValueError, match="t"):
        mlp.gradient(x, F(63, 4, torch.float32))
    with pytest.raises(ValueError, match="w"):
        mlp.train(x, F(64, 4, torch.float32), 1, 64, 1e-3, w=F(64, 2, torch.float16))
    mlp.h = None
