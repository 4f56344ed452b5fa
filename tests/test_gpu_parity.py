"""Parity of the HIP path (through the C ABI) with the oracle on the GPU.

Bar: bit-exact.  Filter words are byte-compared with the oracle's build of the
same keys; match lists must equal the oracle's batch_contains element for
element, including match_offset.  The benchmark workloads at full size are
checked in tests/test_gpu_fullsize.py.
"""
import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dtl_hip as D  # noqa: E402


def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D.load()
    return "cuda"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).to(dev())


def mixed_keys(build, n, hit_rate, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(2**31, 2**32, n, dtype=np.uint64).astype(np.uint32)
    nh = int(round(hit_rate * n))
    if nh and len(build):
        idx = rng.choice(n, nh, replace=False)
        k[idx] = build[rng.integers(0, len(build), nh)]
    return k


def gpu_build(desc, keys_np):
    words = D.new_filter(desc)
    if len(keys_np):
        D.bbf_build(desc, words, to_dev(keys_np))
    torch.cuda.synchronize()
    return words


def oracle_words(f, build):
    w = f.new_words()
    f.insert(w, build)
    return w


def check_cfg(cfg, n_build, n_probe, seed=1, hit=0.2, offset=0, params=None):
    wb, m, k, w, s, z = cfg
    f = O.Bloom(wb, m, k, w, s, z)
    desc = D.bbf_desc(wb, m, k, w, s, z)
    rng = np.random.default_rng(seed)
    build = rng.integers(0, 2**31, n_build, dtype=np.uint64).astype(np.uint32)
    words = gpu_build(desc, build)
    ow = oracle_words(f, build)
    got = words.cpu().numpy().view(ow.dtype)
    assert np.array_equal(got, ow), "filter words differ for %s" % (cfg,)
    probe = mixed_keys(build, n_probe, hit, seed + 1)
    out, cnt = D.bbf_probe(desc, words, to_dev(probe), offset, params=params)
    torch.cuda.synchronize()
    pos = D.positions(out, cnt)
    exp = f.batch_contains(ow, probe, offset)
    assert np.array_equal(pos, exp), "match list differs for %s" % (cfg,)
    return pos


def test_reference_kat_on_gpu():
    """src/dtl/filter/filter_test.cpp:16-40 through bbf_build / bbf_probe."""
    desc = D.bbf_desc(32, 1024, 8, 1, 1)
    words = gpu_build(desc, np.array([13, 37, 42, 88], dtype=np.uint32))
    out, cnt = D.bbf_probe(desc, words, to_dev(np.array([13, 37, 42, 88], np.uint32)), 0)
    assert D.positions(out, cnt).tolist() == [0, 1, 2, 3]
    out, cnt = D.bbf_probe(desc, words, to_dev(np.array([0, 1, 42, 2], np.uint32)), 1000)
    assert D.positions(out, cnt).tolist() == [1002]


BASE_CFGS = [
    (32, 1 << 20, 4, 1, 1, 1),        # plain register-blocked (128 KiB, LDS)
    (64, 1 << 17, 8, 8, 8, 1),        # cache-sectorized 16 KiB (LDS)
    (64, 1 << 25, 8, 8, 8, 1),        # cache-sectorized 4 MiB (global/L2)
    (32, 10 * 20000, 6, 16, 16, 2),   # zoned MAGIC (batched semi-join config)
    (32, 10 * 700000, 6, 16, 16, 2),  # zoned MAGIC, larger than LDS
]


@pytest.mark.parametrize("cfg", BASE_CFGS)
def test_benchmark_configs_small(cfg):
    m_bits = cfg[1]
    n_build = max(100, m_bits // 12)
    check_cfg(cfg, min(n_build, 400000), 300000, seed=3, hit=0.1, offset=17)


@pytest.mark.parametrize("cfg", BASE_CFGS[:2])
def test_lds_and_global_paths_agree(cfg):
    a = check_cfg(cfg, 5000, 100000, seed=5, params=D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL))
    b = check_cfg(cfg, 5000, 100000, seed=5)
    c = check_cfg(cfg, 5000, 100000, seed=5, params=D.ProbeParams(0, 1024, 7, 0, 0))
    assert np.array_equal(a, b) and np.array_equal(a, c)


def all_configs():
    cfgs = []
    for wb in (32, 64):
        for w in (1, 2, 4, 8, 16):
            for s in (1, 2, 4, 8, 16):
                if s > w * wb // 8:
                    continue
                for k in range(1, 17):
                    if k % s == 0:
                        cfgs.append((wb, w, s, 1, k))
        for w in (4, 8, 16):
            for z in (2, 4, 8):
                if z < w:
                    for k in range(2, 17, 2):
                        if k % z == 0:
                            cfgs.append((wb, w, w, z, k))
    return cfgs


def test_every_configuration_parity():
    """Every configuration the reference dispatches, POW2 and MAGIC sizes,
    including the k=8 word-block configurations (scalar semantics)."""
    cfgs = all_configs()
    assert len(cfgs) == D.load().dtl_bbf_config_count()
    for i, (wb, w, s, z, k) in enumerate(cfgs):
        m = (1 << 16) if i % 2 == 0 else 50021
        check_cfg((wb, m, k, w, s, z), 800, 6000, seed=100 + i, hit=0.3, offset=i)


def test_every_configuration_lds_tile_groups():
    """Every configuration through the LDS kernel asked for 1024-thread
    workgroups (taken for 32-bit words, whose workgroups then take the keys
    of 8 tiles at a time; 64-bit words run 256-thread workgroups, which take 2): grids of
    1, 3 and 7 workgroups and the default over 15 or 16 tiles, so that tile
    groups end past the last tile at every position; POW2 and MAGIC sizes;
    default flags and DTL_PROBE_ONE_LANE (both take the LDS kernel; 64- and
    128-byte blocks with statically indexed words use its swizzled layout)."""
    for i, (wb, w, s, z, k) in enumerate(all_configs()):
        m = (1 << 17) if i % 2 == 0 else 100003
        cfg = (wb, m, k, w, s, z)
        f = O.Bloom(*cfg)
        desc = D.bbf_desc(*cfg)
        rng = np.random.default_rng(900 + i)
        build = rng.integers(0, 2**31, 2000, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        ow = oracle_words(f, build)
        probe = mixed_keys(build, 30000 + 2048 * (i % 2) + i, 0.3, 950 + i)
        exp = f.batch_contains(ow, probe, 3)
        keys = to_dev(probe)
        for grid in (1, 3, 7, 0):
            for flags in (0, D.PROBE_ONE_LANE):
                out, cnt = D.bbf_probe(desc, words, keys, 3, params=D.ProbeParams(0, 1024, grid, 0, flags))
                torch.cuda.synchronize()
                assert np.array_equal(D.positions(out, cnt), exp), (cfg, grid, flags)


def test_every_configuration_global_paths():
    """Every configuration through the global-memory kernels (filter larger
    than the LDS limit given): the lane-split kernel where the configuration
    has one and the one-lane kernel, POW2 and MAGIC sizes, ragged last tile."""
    L = D.load()
    split_default = D.ProbeParams(1024, 0, 0, 0, D.PROBE_DIRECT)
    one_lane = D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_DIRECT | D.PROBE_ONE_LANE)
    n_split = 0
    for i, (wb, w, s, z, k) in enumerate(all_configs()):
        m = (1 << 18) if i % 2 == 0 else 200003
        cfg = (wb, m, k, w, s, z)
        f = O.Bloom(*cfg)
        desc = D.bbf_desc(*cfg)
        rng = np.random.default_rng(500 + i)
        build = rng.integers(0, 2**31, 3000, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        ow = oracle_words(f, build)
        assert np.array_equal(words.cpu().numpy().view(ow.dtype), ow), cfg
        probe = mixed_keys(build, 5000, 0.3, 600 + i)
        exp = f.batch_contains(ow, probe, i)
        keys = to_dev(probe)
        for params in (split_default, one_lane):
            out, cnt = D.bbf_probe(desc, words, keys, i, params=params)
            torch.cuda.synchronize()
            assert np.array_equal(D.positions(out, cnt), exp), (cfg, params.flags)
        n_split += L.dtl_bbf_split_lanes(wb, w, s, z, k) > 0
    assert n_split == 107  # the configurations with a lane-split kernel (host test pins the rule)


@pytest.mark.parametrize("cfg", [
    (64, 1 << 25, 8, 8, 8, 1),        # 4 lanes per key, POW2
    (64, 10 * 300001, 16, 16, 16, 1),  # 128 B blocks: 8 lanes per key, MAGIC
    (32, 1 << 24, 8, 8, 8, 1),        # 32 B blocks: 2 lanes per key
    (64, 1 << 24, 8, 8, 4, 1),        # sectors of 16 B, dynamic word inside the piece
    (32, 10 * 400000, 12, 16, 16, 4),  # zones of 16 B
    (64, 1 << 24, 6, 4, 4, 2),        # u64 zones of 16 B
    (32, 10 * 300007, 6, 16, 16, 2),  # zones of 32 B span two pieces, MAGIC
    (64, 1 << 24, 8, 8, 8, 2),        # u64 zones of 32 B: 2 sets of 2 pieces
    (64, 10 * 250007, 8, 16, 16, 2),  # 128 B blocks, zones of 64 B: 2 sets of 4 pieces, MAGIC
    (32, 1 << 23, 4, 8, 8, 2),        # 32 B blocks, one zone of 16 B per piece
])
def test_lane_split_partitioned_and_direct(cfg):
    """The lane-split kernel over grouped keys (partitioned probe) and over
    the keys in order, against the oracle and the one-lane kernel."""
    assert D.load().dtl_bbf_split_lanes(cfg[0], cfg[3], cfg[4], cfg[5], cfg[2]) > 0
    a = check_cfg(cfg, 60000, 300001, seed=41, hit=0.1, offset=9, params=part_params(14))
    b = check_cfg(cfg, 60000, 300001, seed=41, hit=0.1, offset=9,
                  params=D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_DIRECT))
    c = check_cfg(cfg, 60000, 300001, seed=41, hit=0.1, offset=9,
                  params=D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_DIRECT | D.PROBE_ONE_LANE))
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 2047, 2048, 2049, 4097, 100003])
def test_edge_sizes(n):
    cfg = (64, 1 << 17, 8, 8, 8, 1)
    desc = D.bbf_desc(*cfg[:1], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5])
    f = O.Bloom(*cfg)
    build = np.arange(1, 3001, dtype=np.uint32) * np.uint32(2654435761 & 0x7FFFFFFF)
    words = gpu_build(desc, build)
    probe = mixed_keys(build, n, 0.5, 9) if n else np.zeros(0, np.uint32)
    keys = to_dev(probe) if n else torch.zeros(0, dtype=torch.int32, device=dev())
    for params in (None, D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL)):
        out, cnt = D.bbf_probe(desc, words, keys, 5, params=params)
        torch.cuda.synchronize()
        exp = f.batch_contains(oracle_words(f, build), probe, 5) if n else np.zeros(0, np.uint32)
        assert np.array_equal(D.positions(out, cnt), exp)


@pytest.mark.parametrize("n", [(1 << 24) - 1, 1 << 24, (1 << 24) + 1, (1 << 24) + 2049])
def test_compaction_at_fused_scan_limit(n):
    """Lists of up to 8192 tiles (2^24 keys) are compacted by one launch that
    adds up the tile counts itself; longer ones by a scan and an expand.  The
    plain register-blocked filter (LDS path) on either side of the limit,
    against the oracle, with a match offset."""
    cfg = (32, 1 << 20, 4, 1, 1, 1)
    desc = D.bbf_desc(*cfg[:1], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5])
    f = O.Bloom(*cfg)
    build = np.arange(1, (1 << 16) + 1, dtype=np.uint32) * np.uint32(2654435761 & 0x7FFFFFFF)
    words = gpu_build(desc, build)
    probe = mixed_keys(build, n, 0.1, 17)
    out, cnt = D.bbf_probe(desc, words, to_dev(probe), 77)
    torch.cuda.synchronize()
    exp = f.batch_contains(oracle_words(f, build), probe, 77)
    assert np.array_equal(D.positions(out, cnt), exp)


def test_compaction_dense_and_sparse_groups():
    """A list of more than 8192 tiles (scan + expand) whose density changes
    along the batch: runs of 100%, ~25%, ~12% (around the density from which
    a workgroup stages each tile's output in LDS) and 1% hits, and a last
    partial tile, so that both store forms of the expand and the switch
    between neighbouring groups are exercised, against the oracle."""
    cfg = (32, 1 << 20, 4, 1, 1, 1)
    desc = D.bbf_desc(*cfg[:1], cfg[1], cfg[2], cfg[3], cfg[4], cfg[5])
    f = O.Bloom(*cfg)
    build = np.arange(1, (1 << 16) + 1, dtype=np.uint32) * np.uint32(2654435761 & 0x7FFFFFFF)
    words = gpu_build(desc, build)
    parts = []
    for i, (size, hit) in enumerate([(3 << 20, 1.0), (1 << 21, 0.25), (1 << 21, 0.11), (1 << 21, 0.01),
                                     (5 << 20, 0.5), ((1 << 22) + 777, 0.12), ((1 << 21) + 5, 0.9)]):
        parts.append(mixed_keys(build, size, hit, 40 + i))
    probe = np.concatenate(parts)
    assert len(probe) > 8192 * 2048
    out, cnt = D.bbf_probe(desc, words, to_dev(probe), 5)
    torch.cuda.synchronize()
    exp = f.batch_contains(oracle_words(f, build), probe, 5)
    assert np.array_equal(D.positions(out, cnt), exp)


def part_params(slice_log2):
    p = D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_PARTITIONED)
    p.slice_log2 = slice_log2
    return p


PART_CASES = [
    ((64, 1 << 22, 8, 8, 8, 1), 12, 300000),       # 512 KiB filter, 4 KiB slices -> 128 partitions
    ((64, 1 << 25, 8, 8, 8, 1), 8, 500003),        # 4 MiB, slices widened until 2048 partitions
    ((32, 10 * 200000, 6, 16, 16, 2), 13, 123457),  # zoned MAGIC, ragged partitions
    ((32, 1 << 21, 4, 1, 1, 1), 10, 70000),        # register-blocked u32 w1
    ((64, 3000011, 7, 4, 1, 1), 11, 99999),        # MAGIC, word-blocked u64 w4
    ((32, 5000000, 16, 8, 2, 1), 14, 4097),        # sltw-style, n just over 2 tiles
]


@pytest.mark.parametrize("cfg,slice_log2,n_probe", PART_CASES)
def test_partitioned_probe_matches_oracle(cfg, slice_log2, n_probe):
    """Partitioned probe (keys grouped by filter slice) equals the oracle."""
    a = check_cfg(cfg, 40000, n_probe, seed=31, hit=0.15, offset=77, params=part_params(slice_log2))
    b = check_cfg(cfg, 40000, n_probe, seed=31, hit=0.15, offset=77,
                  params=D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_DIRECT))
    assert np.array_equal(a, b)


BIN_CASES = [
    ((64, 1 << 25, 8, 8, 8, 1), 19, 1 << 20),        # 4 MiB, 512 KiB slices: 64 bins of 8 tiles
    ((32, 1 << 25, 4, 1, 1, 1), 18, (1 << 20) + 777),  # register-blocked u32 w1, ragged last tile
    ((32, 30000011, 6, 16, 16, 2), 18, 1200001),     # zoned MAGIC, lane-split spanning zones
    # 512 and 1024 partitions: the scatter's 32K-key tiles (32 keys per thread)
    ((64, 1 << 25, 8, 8, 8, 1), 13, (1 << 25) + 777),
    ((64, 1 << 25, 8, 8, 8, 1), 12, (1 << 26) + 4097),
]


@pytest.mark.parametrize("cfg,slice_log2,n_probe", BIN_CASES)
def test_partitioned_capacity_bins_match_histogram_and_oracle(cfg, slice_log2, n_probe):
    """Partitioned probe with fixed-capacity bins (no histogram pass; taken
    when the bins hold 4 tiles or more) equals the histogram-sized form, the
    direct kernel and the oracle."""
    a = check_cfg(cfg, 60000, n_probe, seed=61, hit=0.1, offset=5, params=part_params(slice_log2))
    p = part_params(slice_log2)
    p.flags |= D.PROBE_PART_HISTOGRAM
    b = check_cfg(cfg, 60000, n_probe, seed=61, hit=0.1, offset=5, params=p)
    c = check_cfg(cfg, 60000, n_probe, seed=61, hit=0.1, offset=5,
                  params=D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL | D.PROBE_DIRECT))
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("dup_share,slice_log2,n", [(0.3, 19, 1 << 21), (0.9, 19, 1 << 21),
                                                   (0.5, 13, (1 << 25) + 5)])
def test_partitioned_capacity_bins_spill_exact(dup_share, slice_log2, n):
    """A batch in which most keys are one value (all in one partition) fills
    that partition's bins far past their capacity: the keys that do not fit
    take slots of other bins, and the list still equals the oracle's (8
    partitions; and 512, where the scatter takes 32K-key tiles)."""
    cfg = (64, 1 << 25, 8, 8, 8, 1)
    f = O.Bloom(*cfg)
    desc = D.bbf_desc(*cfg)
    rng = np.random.default_rng(62)
    build = rng.integers(0, 2**31, 60000, dtype=np.uint64).astype(np.uint32)
    words = gpu_build(desc, build)
    ow = oracle_words(f, build)
    probe = mixed_keys(build, n, 0.05, 63)
    idx = rng.choice(n, int(dup_share * n), replace=False)
    probe[idx[: len(idx) // 2]] = build[7]            # a member: every copy matches
    probe[idx[len(idx) // 2:]] = np.uint32(2**31 + 5)  # a non-member
    exp = f.batch_contains(ow, probe, 11)
    out, cnt = D.bbf_probe(desc, words, to_dev(probe), 11, params=part_params(slice_log2))
    torch.cuda.synchronize()
    assert np.array_equal(D.positions(out, cnt), exp)


@pytest.mark.parametrize("n", [1, 2, 63, 2047, 2049, 4095, 4096, 4097, 10000])
def test_partitioned_probe_small_batches(n):
    cfg = (64, 1 << 20, 8, 8, 8, 1)
    check_cfg(cfg, 5000, n, seed=n, hit=0.4, offset=3, params=part_params(10))


def test_partitioned_probe_repeated_calls_reuse_scratch():
    """Back-to-back partitioned probes of different sizes on one stream."""
    cfg = (64, 1 << 23, 8, 8, 8, 1)
    for n, seed in ((200000, 1), (50000, 2), (300001, 3)):
        check_cfg(cfg, 60000, n, seed=seed, hit=0.05, params=part_params(14))


def test_match_offset_wraps_like_u32():
    cfg = (32, 1 << 14, 4, 1, 1, 1)
    off = 2**32 - 10
    pos = check_cfg(cfg, 300, 5000, seed=11, hit=0.5, offset=off)
    assert pos.dtype == np.uint32 and len(pos) > 20


def test_build_is_order_independent_and_idempotent():
    cfg = (64, 1 << 22, 8, 8, 8, 1)
    desc = D.bbf_desc(64, 1 << 22, 8, 8, 8, 1)
    rng = np.random.default_rng(21)
    keys = rng.integers(0, 2**32, 300000, dtype=np.uint64).astype(np.uint32)
    a = gpu_build(desc, keys)
    b = gpu_build(desc, keys[::-1].copy())
    D.bbf_build(desc, b, to_dev(keys))
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    f = O.Bloom(*cfg)
    assert np.array_equal(a.cpu().numpy().view(np.uint64), oracle_words(f, keys))


# ---------------------------------------------------------------------------
# cuckoo filter
# ---------------------------------------------------------------------------
def cf_part_params(slice_log2=10):
    p = D.ProbeParams(0, 0, 0, 0, D.PROBE_PARTITIONED)
    p.slice_log2 = slice_log2
    return p


def check_cf(m, t, b, load, n_probe, hit, seed, offset=0, overfill=False, params=(None,)):
    """The GPU list equals the oracle's for each probe parameter set given
    (None = defaults)."""
    d = D.cf_desc(m, t, b)
    o = O.Cuckoo(m, t, b)
    rng = np.random.default_rng(seed)
    n = int(load * d.bucket_cnt * b)
    keys = np.unique(rng.integers(0, 2**31, int(n * 1.2) + 10, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(keys)
    keys = keys[:n]
    buf, ins = D.cf_build_host(d, keys, seed=seed)
    if not overfill:
        assert ins == len(keys)
    assert ins > 0
    dbuf = torch.from_numpy(buf).to(dev())
    probe = mixed_keys(keys, n_probe, hit, seed + 7)
    exp = o.batch_contains(buf, probe, offset)
    kd = to_dev(probe)
    for prm in params:
        out, cnt = D.cf_probe(d, dbuf, kd, offset, params=prm)
        torch.cuda.synchronize()
        pos = D.positions(out, cnt)
        assert np.array_equal(pos, exp), (t, b, None if prm is None else (prm.flags, prm.slice_log2))
    return d, buf, pos


@pytest.mark.parametrize("t,b", [(16, 4), (8, 4), (12, 4), (4, 4), (32, 4), (16, 2), (8, 2), (12, 2),
                                 (32, 2), (4, 2), (16, 1), (8, 1), (12, 1), (4, 1), (32, 1)])
def test_cuckoo_all_configs(t, b):
    load = {1: 0.3, 2: 0.8, 4: 0.9}[b] if t > 4 else 0.3
    bpb = (t * b + 7) // 8
    # tables with one slot per bucket or 4-bit tags may overflow into the
    # victim cache; the probe must still match the oracle on the same bytes
    over = t == 4 or b == 1
    # direct kernel and the two-pass partitioned probe (slices of 1 KiB: tens of partitions)
    prm = (D.ProbeParams(0, 0, 0, 0, D.PROBE_DIRECT), cf_part_params(10))
    check_cf((1 << 13) * bpb * 8, t, b, load, 50000, 0.5, seed=t * 10 + b, overfill=over, params=prm)
    check_cf(5003 * bpb * 8, t, b, load, 30000, 0.3, seed=t * 10 + b + 1, offset=99, overfill=over, params=prm)


def test_cuckoo_victim_cache_hits():
    d, buf, pos = check_cf(64 * 64, 16, 4, 1.5, 20000, 0.5, seed=3, overfill=True,
                           params=(None, cf_part_params(7)))
    assert buf[d.victim_offset + 8] == 1


def test_cuckoo_benchmark_config_small():
    """Config 3 shape at 1/64 scale: 16-bit tags, 4 slots, pow2, 95% load."""
    check_cf(1 << 26, 16, 4, 0.95, 1 << 20, 0.01, seed=44, params=(None, cf_part_params(14)))
    check_cf(1 << 26, 16, 4, 0.95, 1 << 20, 0.5, seed=45, params=(None, cf_part_params(14)))


@pytest.mark.parametrize("n", [1, 2, 63, 2047, 2049, 4097, 100003])
def test_cuckoo_partitioned_small_batches(n):
    check_cf(3001 * 8 * 8, 16, 4, 0.9, n, 0.4, seed=n, offset=5, params=(cf_part_params(9),))


def test_cuckoo_partitioned_large_batch():
    """2^25 probe keys against a 32 MiB table (default-sized slices), both
    passes and the hit sort at scale, against the oracle's threaded
    restatement (pinned to the scalar oracle by tests/test_cpu_baseline.py);
    the passes' fixed-capacity bins (the default at this size) and
    histogram-sized partitions give the same list."""
    import cpu_baseline as CB
    m = (1 << 22) * 8 * 8
    d = D.cf_desc(m, 16, 4)
    o = O.Cuckoo(m, 16, 4)
    rng = np.random.default_rng(8)
    n = int(0.9 * d.bucket_cnt * 4)
    keys = np.unique(rng.integers(0, 2**31, int(n * 1.2), dtype=np.uint64).astype(np.uint32))
    rng.shuffle(keys)
    keys = keys[:n]
    buf, ins = D.cf_build_host(d, keys, seed=8)
    assert ins == len(keys)
    probe = big_probe_keys(keys, (1 << 25) - 3, 9)
    p = D.ProbeParams(0, 0, 0, 0, D.PROBE_PARTITIONED)
    out, cnt = D.cf_probe(d, torch.from_numpy(buf).to(dev()), to_dev(probe), 3, params=p)
    torch.cuda.synchronize()
    got = D.positions(out, cnt).astype(np.uint64)
    _, _, lists = CB.cf_probe(o, buf, probe, min(16, CB.physical_cores()))
    exp = np.concatenate(lists).astype(np.uint64) + 3
    assert len(got) == len(exp) and np.array_equal(got, exp)
    ph = D.ProbeParams(0, 0, 0, 0, D.PROBE_PARTITIONED | D.PROBE_PART_HISTOGRAM)
    out, cnt = D.cf_probe(d, torch.from_numpy(buf).to(dev()), to_dev(probe), 3, params=ph)
    torch.cuda.synchronize()
    assert np.array_equal(D.positions(out, cnt).astype(np.uint64), exp)


def test_cuckoo_partitioned_many_partitions():
    """More than 1024 partitions (4 KiB slices of an 8 MiB table): the two
    passes count their partitions in separate histogram passes."""
    d, buf, pos = check_cf((1 << 20) * 64, 16, 4, 0.9, (1 << 20) + 7, 0.3, seed=17, offset=3,
                           params=(cf_part_params(12),))
    assert len(pos) > 0


def test_cuckoo_default_path_at_break_even_size():
    """A 32 MiB table probed with 2^23 keys takes the partitioned two-pass
    probe by default: same list as the forced direct kernel and the oracle's."""
    import cpu_baseline as CB
    m = (1 << 22) * 64
    d = D.cf_desc(m, 16, 4)
    o = O.Cuckoo(m, 16, 4)
    keys = distinct_keys(int(0.9 * d.bucket_cnt * 4), 21)
    buf, ins = D.cf_build_host(d, keys, seed=21)
    assert ins == len(keys)
    probe = big_probe_keys(keys, (1 << 23) + 3, 22)
    dbuf = torch.from_numpy(buf).to(dev())
    kd = to_dev(probe)
    _, _, parts = CB.cf_probe(o, buf, probe, min(16, CB.physical_cores()))
    exp = np.concatenate(parts).astype(np.uint64) + 1
    for prm in (None, D.ProbeParams(0, 0, 0, 0, D.PROBE_DIRECT)):
        out, cnt = D.cf_probe(d, dbuf, kd, 1, params=prm)
        torch.cuda.synchronize()
        got = D.positions(out, cnt).astype(np.uint64)
        assert len(got) == len(exp) and np.array_equal(got, exp)


@pytest.mark.parametrize("hit", [0.0, 0.01, 0.03, 0.1, 0.5, 1.0, "dup"])
def test_cuckoo_pass2_leaves_out_found_keys(hit):
    """Partitioned cuckoo probe with fixed-capacity bins (16 partitions of an
    8 MiB table at 95% load, 2^21 + 5 keys): where pass 1 finds more than
    n/32 keys (about 59% of a table's keys sit in their first bucket, so hit
    rates of 10% and more), pass 2 regroups only the keys pass 1 did not
    find and its matches are ORed into pass 1's sorted bitmap; below that
    both passes take every key.  Either way, and with every key regrouped
    twice (PROBE_CF_PASS2_ALL), the list is the oracle's.  "dup": 90% of the
    keys are one table key (pass 2 leaves nearly all out; pass 1's bins
    spill)."""
    m = (1 << 20) * 64
    d = D.cf_desc(m, 16, 4)
    o = O.Cuckoo(m, 16, 4)
    keys = distinct_keys(int(0.95 * d.bucket_cnt * 4), 31)
    buf, ins = D.cf_build_host(d, keys, seed=31)
    assert ins == len(keys)
    n = (1 << 21) + 5
    if hit == "dup":
        probe = mixed_keys(keys, n, 0.05, 32)
        rng = np.random.default_rng(33)
        probe[rng.choice(n, int(0.9 * n), replace=False)] = keys[12345]
    else:
        probe = mixed_keys(keys, n, hit, 32)
    exp = o.batch_contains(buf, probe, 7)
    dbuf = torch.from_numpy(buf).to(dev())
    kd = to_dev(probe)
    for flags in (D.PROBE_PARTITIONED, D.PROBE_PARTITIONED | D.PROBE_CF_PASS2_ALL):
        prm = D.ProbeParams(0, 0, 0, 0, flags)
        prm.slice_log2 = 19
        for _ in range(2):  # the second call finds the first one's scratch contents
            out, cnt = D.cf_probe(d, dbuf, kd, 7, params=prm)
            torch.cuda.synchronize()
            assert np.array_equal(D.positions(out, cnt), exp), (hit, flags)


# ---------------------------------------------------------------------------
# cuckoo GPU build: concurrent inserts give a different table than the
# sequential insert, so parity is "valid table + same probe semantics": the
# oracle's placement check (each key's tag in one of its two buckets or the
# victim cache, exactly one tag per key), every key found, and the GPU probe
# list equal to the oracle's on the GPU-built bytes.
# ---------------------------------------------------------------------------
CF_GPU_BUILD = [(16, 4), (8, 4), (4, 4), (32, 4), (16, 2), (8, 2), (32, 2), (4, 2), (16, 1), (8, 1), (12, 1),
                (4, 1), (32, 1)]


def distinct_keys(n, seed):
    rng = np.random.default_rng(seed)
    keys = np.unique(rng.integers(0, 2**31, int(n * 1.2) + 10, dtype=np.uint64).astype(np.uint32))
    rng.shuffle(keys)
    return keys[:n]


def gpu_cf_build(d, keys, seed, max_kicks=0, buf=None):
    if buf is None:
        buf = D.new_cf_table(d)
    failed = D.cf_build(d, buf, to_dev(keys), seed=seed, max_kicks=max_kicks)
    torch.cuda.synchronize()
    return buf, int(failed.item())


def check_gpu_built(d, o, buf, keys, seed, offset=0):
    hb = buf.cpu().numpy()
    assert o.placement_valid(hb, keys)
    assert o.batch_contains(hb, keys).size == len(keys)       # no false negatives
    probe = mixed_keys(keys, 30000, 0.4, seed + 1)
    exp = o.batch_contains(hb, probe, offset)
    for prm in (None, cf_part_params(9)):
        out, cnt = D.cf_probe(d, buf, to_dev(probe), offset, params=prm)
        torch.cuda.synchronize()
        assert np.array_equal(D.positions(out, cnt), exp)
    return hb


@pytest.mark.parametrize("t,b", CF_GPU_BUILD)
def test_cuckoo_gpu_build_valid_table(t, b):
    """Power-of-two and MAGIC tables at loads the sequential insert fills
    without the victim cache."""
    load = {1: 0.3, 2: 0.4 if t == 4 else 0.8, 4: 0.8 if t == 4 else 0.9}[b]
    bpb = (t * b + 7) // 8
    for m, seed in (((1 << 12) * bpb * 8, t * 10 + b), (3001 * bpb * 8, t * 10 + b + 500)):
        d = D.cf_desc(m, t, b)
        o = O.Cuckoo(m, t, b)
        keys = distinct_keys(int(load * d.bucket_cnt * b), seed)
        buf, failed = gpu_cf_build(d, keys, seed)
        assert failed == 0, (t, b, m)
        check_gpu_built(d, o, buf, keys, seed, offset=7)


def test_cuckoo_gpu_build_rejects_straddling_buckets():
    d = D.cf_desc(1 << 16, 12, 4)
    dev()
    with pytest.raises(RuntimeError):
        D.cf_build(d, D.new_cf_table(d), to_dev(np.arange(10, dtype=np.uint32)))


def test_cuckoo_gpu_build_in_two_calls():
    """A second call adds keys to the table the first call built."""
    m = 5003 * 64
    d = D.cf_desc(m, 16, 4)
    o = O.Cuckoo(m, 16, 4)
    keys = distinct_keys(int(0.93 * d.bucket_cnt * 4), 31)
    half = len(keys) // 2
    buf, f1 = gpu_cf_build(d, keys[:half], 31)
    buf, f2 = gpu_cf_build(d, keys[half:], 32, buf=buf)
    assert f1 == 0 and f2 == 0
    check_gpu_built(d, o, buf, keys, 31)


def test_cuckoo_gpu_build_overflow_accounts_for_every_key():
    """Past capacity, every walk ends in a slot, the victim cache or the
    failed count; a table whose failed count is 0 stays valid."""
    d = D.cf_desc(64 * 64, 16, 4)           # 64 buckets, 256 slots
    o = O.Cuckoo(64 * 64, 16, 4)
    keys = distinct_keys(400, 5)
    buf, failed = gpu_cf_build(d, keys, 5, max_kicks=50)
    hb = buf.cpu().numpy()
    assert o.victim(hb)[2]
    assert failed > 0
    assert o.count_occupied(hb) + failed == len(keys)
    keys = distinct_keys(200, 6)
    buf, failed = gpu_cf_build(d, keys, 6)
    assert failed == 0
    check_gpu_built(d, o, buf, keys, 6)


@pytest.mark.parametrize("buckets,t,b", [(300007, 16, 4), (250007, 32, 4), (200003, 8, 2)])
def test_cuckoo_gpu_build_magic_high_load(buckets, t, b):
    """MAGIC tables at high load (95% for 4 slots, 85% for 2): the 16-byte
    buckets of 32-bit tags (two compare-and-swap words per bucket) and
    sub-word buckets sharing a 4-byte word with their neighbours.  Odd
    bucket counts: with an even one the reference's MAGIC alternate index
    caps the fill at 74-85% (its sequential insert overflows too)."""
    bpb = (t * b + 7) // 8
    d = D.cf_desc(buckets * bpb * 8, t, b)
    assert d.addr_mode == D.ADDR_MAGIC and d.bucket_cnt % 2 == 1
    o = O.Cuckoo(buckets * bpb * 8, t, b)
    keys = distinct_keys(int((0.95 if b == 4 else 0.85) * d.bucket_cnt * b), buckets)
    buf, failed = gpu_cf_build(d, keys, 3)
    assert failed == 0
    check_gpu_built(d, o, buf, keys, 4)


def test_cuckoo_gpu_build_benchmark_shape():
    """The benchmark table's shape at 1/256 scale: 16-bit tags, 4 slots,
    2^18 buckets, 95% load (996,147 keys)."""
    m = (1 << 18) * 64
    d = D.cf_desc(m, 16, 4)
    o = O.Cuckoo(m, 16, 4)
    keys = distinct_keys(int(0.95 * d.bucket_cnt * 4), 12)
    buf, failed = gpu_cf_build(d, keys, 12)
    assert failed == 0
    check_gpu_built(d, o, buf, keys, 12)


# ---------------------------------------------------------------------------
# batched
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("params", [None, D.ProbeParams(0, 0, 0, 4, 0), D.ProbeParams(0, 1024, 0, 2, 0),
                                    D.ProbeParams(0, 0, 0, 0, D.PROBE_FORCE_GLOBAL),
                                    D.ProbeParams(100000, 1024, 64, 8, 0)])
def test_batched_matches_per_filter(params):
    rng = np.random.default_rng(77)
    tasks, expect = [], []
    for i in range(40):
        n = int(2 ** rng.uniform(6, 16))
        m = 10 * n
        f = O.Bloom(32, m, 6, 16, 16, 2)
        desc = D.bbf_desc(32, m, 6, 16, 16, 2)
        build = rng.integers(0, 2**31, n, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        nprobe = int(2 ** rng.uniform(0, 15)) if i % 9 else 0
        probe = mixed_keys(build, nprobe, rng.uniform(0.001, 0.3), i)
        keys = to_dev(probe) if nprobe else torch.zeros(0, dtype=torch.int32, device=dev())
        off = int(rng.integers(0, 1 << 20))
        tasks.append((desc, words, keys, off))
        expect.append(f.batch_contains(oracle_words(f, build), probe, off) if nprobe else np.zeros(0, np.uint32))
    plan = D.BatchPlan(tasks, params)
    for rep in range(2):
        plan.run()
        torch.cuda.synchronize()
        got = plan.results()
        for g, e in zip(got, expect):
            assert np.array_equal(g, e)
    plan.close()


def test_batched_every_configuration():
    """One plan with a task of every configuration the reference dispatches
    (one group each), POW2 and MAGIC sizes, ragged batches: staged in LDS and
    from global memory, each with the full and the halved number of keys
    (pieces) per lane in flight; every list equals the oracle's."""
    tasks, expect = [], []
    for i, (wb, w, s, z, k) in enumerate(all_configs()):
        m = (1 << 16) if i % 2 == 0 else 70001
        cfg = (wb, m, k, w, s, z)
        f = O.Bloom(*cfg)
        desc = D.bbf_desc(*cfg)
        rng = np.random.default_rng(3000 + i)
        build = rng.integers(0, 2**31, 1500, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        probe = mixed_keys(build, 2048 * (1 + i % 3) + 7 * i, 0.3, 3500 + i)
        tasks.append((desc, words, to_dev(probe), i))
        expect.append(f.batch_contains(oracle_words(f, build), probe, i))
    for flags in (0, D.PROBE_HALF_IN_FLIGHT, D.PROBE_FORCE_GLOBAL, D.PROBE_FORCE_GLOBAL | D.PROBE_HALF_IN_FLIGHT):
        plan = D.BatchPlan(tasks, D.ProbeParams(0, 0, 0, 1, flags))
        assert plan.group_count() == len(tasks)
        plan.run()
        torch.cuda.synchronize()
        for j, (g, e) in enumerate(zip(plan.results(), expect)):
            assert np.array_equal(g, e), (tasks[j][0].word_bits, tasks[j][0].w, tasks[j][0].s, tasks[j][0].z,
                                          tasks[j][0].k, flags)
        plan.close()


def test_batched_graph_replay():
    """The batch plan runs under hipGraph capture; replays match the eager run."""
    rng = np.random.default_rng(5)
    tasks, expect = [], []
    for i in range(12):
        n = int(2 ** rng.uniform(8, 14))
        f = O.Bloom(32, 10 * n, 6, 16, 16, 2)
        desc = D.bbf_desc(32, 10 * n, 6, 16, 16, 2)
        build = rng.integers(0, 2**31, n, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        probe = mixed_keys(build, 5000 + 100 * i, 0.1, i)
        tasks.append((desc, words, to_dev(probe), i))
        expect.append(f.batch_contains(oracle_words(f, build), probe, i))
    plan = D.BatchPlan(tasks)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.run(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.run(s)
    for rep in range(3):
        for c in plan.counts:
            c.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for got, e in zip(plan.results(), expect):
            assert np.array_equal(got, e)
    plan.close()


# ---------------------------------------------------------------------------
# full size (1 GiB cache-sectorized filter): properties + exact sample
# ---------------------------------------------------------------------------
def big_probe_keys(build, n, seed):
    """n probe keys (1% of them build keys) without materialising a
    per-key random stream larger than needed."""
    rng = np.random.default_rng(seed)
    k = (np.arange(n, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(seed)).astype(np.uint32) | np.uint32(1 << 31)
    nh = n // 100
    k[rng.integers(0, n, nh)] = build[rng.integers(0, len(build), nh)]
    return k


@pytest.mark.parametrize("ntiles,path", [(4096, "direct"), (16384, "direct"), (65536, "direct"),
                                         (16384, "partitioned"), (65536, "partitioned")])
def test_large_batches_exact(ntiles, path):
    """Batches long enough for the multi-tile compaction groups (2, 4 and 8
    tiles per workgroup) and for the hit sort at scale, against the oracle."""
    cfg = (64, 1 << 25, 8, 8, 8, 1) if path == "direct" else (64, 1 << 29, 8, 8, 8, 1)
    f = O.Bloom(*cfg)
    desc = D.bbf_desc(*cfg)
    rng = np.random.default_rng(ntiles)
    build = rng.integers(0, 2**31, 300000, dtype=np.uint64).astype(np.uint32)
    words = gpu_build(desc, build)
    ow = oracle_words(f, build)
    n = ntiles * 2048 - 37
    probe = big_probe_keys(build, n, ntiles + 1)
    flags = D.PROBE_DIRECT if path == "direct" else D.PROBE_PARTITIONED
    out, cnt = D.bbf_probe(desc, words, to_dev(probe), 11, params=D.ProbeParams(0, 0, 0, 0, flags))
    torch.cuda.synchronize()
    got = D.positions(out, cnt)
    # checker: the oracle's threaded AVX-512 restatement of batch_contains
    # (tests/test_cpu_baseline.py pins it to the scalar oracle); offset added here
    import cpu_baseline as CB
    threads = min(16, CB.physical_cores())
    _, _, lists = CB.bbf_probe(f, ow, probe, threads)
    exp = np.concatenate(lists).astype(np.uint64) + 11
    assert len(got) == len(exp)
    assert np.array_equal(got.astype(np.uint64), exp)


@pytest.mark.parametrize("log2_m,log2_n", [(27, 24), (28, 23), (26, 25)])
def test_default_path_at_break_even_sizes(log2_m, log2_n):
    """16 MiB with 2^24 + 3 keys, 32 MiB with 2^23 + 3 and 8 MiB with 2^25 + 3
    take the partitioned probe by default (fixed-capacity bins): the default
    call's list equals the forced direct and forced partitioned ones and the
    oracle's on every key."""
    cfg = (64, 1 << log2_m, 8, 8, 8, 1)
    f = O.Bloom(*cfg)
    desc = D.bbf_desc(*cfg)
    build = np.random.default_rng(log2_m).integers(0, 2**31, (1 << log2_m) // 12, dtype=np.uint64).astype(np.uint32)
    words = gpu_build(desc, build)
    ow = oracle_words(f, build)
    probe = big_probe_keys(build, (1 << log2_n) + 3, log2_n)
    kd = to_dev(probe)
    lists = []
    for flags in (0, D.PROBE_DIRECT, D.PROBE_PARTITIONED):
        out, cnt = D.bbf_probe(desc, words, kd, 2, params=D.ProbeParams(0, 0, 0, 0, flags))
        torch.cuda.synchronize()
        lists.append(D.positions(out, cnt).astype(np.uint64))
    import cpu_baseline as CB
    _, _, parts = CB.bbf_probe(f, ow, probe, min(16, CB.physical_cores()))
    exp = np.concatenate(parts).astype(np.uint64) + 2
    for got in lists:
        assert len(got) == len(exp) and np.array_equal(got, exp)


def test_batched_autotune_keeps_results():
    rng = np.random.default_rng(123)
    tasks, expect = [], []
    for i in range(20):
        n = int(2 ** rng.uniform(8, 15))
        f = O.Bloom(32, 10 * n, 6, 16, 16, 2)
        desc = D.bbf_desc(32, 10 * n, 6, 16, 16, 2)
        build = rng.integers(0, 2**31, n, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        probe = mixed_keys(build, int(2 ** rng.uniform(10, 16)), 0.05, i)
        tasks.append((desc, words, to_dev(probe), 3 * i))
        expect.append(f.batch_contains(oracle_words(f, build), probe, 3 * i))
    plan = D.BatchPlan(tasks)
    best = plan.tune(reps=1)
    assert best.block_threads in (256, 1024) and best.chunk_tiles >= 1
    # every group's tuned shape is one of the tuner's candidates; the counts add up
    n_tasks = n_keys = 0
    for gi in range(plan.group_count()):
        sh, nt, nk, mfb = plan.group_shape(gi)
        assert sh.block_threads in (256, 1024) and sh.chunk_tiles in (1, 2, 4, 8, 16) and sh.grid_blocks >= 1
        assert sh.lds_max_bytes in (0, (mfb + 15) // 16 * 16)
        assert (sh.flags & ~(D.PROBE_FORCE_GLOBAL | D.PROBE_HALF_IN_FLIGHT)) == 0
        n_tasks += nt
        n_keys += nk
    assert n_tasks == len(tasks) and n_keys == sum(int(t[2].numel()) for t in tasks)
    with pytest.raises(D.DtlError):
        plan.group_shape(plan.group_count())
    plan.run()
    torch.cuda.synchronize()
    for g, e in zip(plan.results(), expect):
        assert np.array_equal(g, e)
    plan.close()


MIXED_CFGS = [(32, 6, 16, 16, 2), (64, 8, 8, 8, 1), (32, 4, 1, 1, 1), (64, 12, 8, 8, 4), (32, 3, 2, 1, 1),
              (64, 10, 4, 2, 1)]


def test_batched_mixed_configurations():
    """One plan over tasks of several configurations (interleaved), tuned and
    graph-replayed: every list equals the oracle's."""
    rng = np.random.default_rng(99)
    tasks, expect = [], []
    for i in range(30):
        wb, k, w, s, z = MIXED_CFGS[i % len(MIXED_CFGS)]
        n = int(2 ** rng.uniform(7, 15))
        m = int(rng.integers(6, 16)) * n
        f = O.Bloom(wb, m, k, w, s, z)
        desc = D.bbf_desc(wb, m, k, w, s, z)
        build = rng.integers(0, 2**31, n, dtype=np.uint64).astype(np.uint32)
        words = gpu_build(desc, build)
        probe = mixed_keys(build, int(2 ** rng.uniform(8, 15)), rng.uniform(0.01, 0.3), i)
        tasks.append((desc, words, to_dev(probe), 7 * i))
        expect.append(f.batch_contains(oracle_words(f, build), probe, 7 * i))
    plan = D.BatchPlan(tasks)
    # one launch per (configuration, size class): filters up to 32 KiB, up to
    # the LDS capacity (160 KiB less 2 KiB), larger
    cls = lambda nb: 0 if nb <= 32 * 1024 else 1 if nb <= 160 * 1024 - 2048 else 2
    groups = {((t[0].word_bits, t[0].w, t[0].s, t[0].z, t[0].k), cls(t[0].size_bytes)) for t in tasks}
    assert len({g[0] for g in groups}) == len(MIXED_CFGS) and len({g[1] for g in groups}) > 1
    assert plan.group_count() == len(groups)
    plan.tune(reps=1)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.run(s)
    torch.cuda.synchronize()
    for g, e in zip(plan.results(), expect):
        assert np.array_equal(g, e)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.run(s)
    for rep in range(2):
        for c in plan.counts:
            c.fill_(-1)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for g, e in zip(plan.results(), expect):
            assert np.array_equal(g, e)
    plan.close()
    # every filter from global memory: lane-split groups (u64 w8 s8, zoned u64
    # w8 z4) and one-lane groups, 256- and 1024-thread workgroups, with the
    # full and the halved number of keys (pieces) per lane in flight; and the
    # halved count with filters staged in LDS
    for threads in (256, 1024):
        for flags in (D.PROBE_FORCE_GLOBAL, D.PROBE_FORCE_GLOBAL | D.PROBE_HALF_IN_FLIGHT, D.PROBE_HALF_IN_FLIGHT):
            plan = D.BatchPlan(tasks, D.ProbeParams(0, threads, 0, 2, flags))
            for gi in range(plan.group_count()):
                sh = plan.group_shape(gi)[0]
                assert bool(sh.flags & D.PROBE_HALF_IN_FLIGHT) == bool(flags & D.PROBE_HALF_IN_FLIGHT)
            plan.run()
            torch.cuda.synchronize()
            for g, e in zip(plan.results(), expect):
                assert np.array_equal(g, e), (threads, flags)
            plan.close()


# ---------------------------------------------------------------------------
# roofline microbenchmark: every access pattern runs (it reports only time)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode,log2_bytes", [(m, 22) for m in range(8)] + [(8, 14), (8, 17), (9, 14), (9, 17)])
def test_gather_bench_modes_run(mode, log2_bytes):
    dev()
    buf = torch.randint(0, 2**31, (1 << (log2_bytes - 2),), dtype=torch.int32, device="cuda")
    sink = torch.zeros(1, dtype=torch.int64, device="cuda")
    D.gather_bench(buf, mode, (1 << 20) + 5, sink)
    torch.cuda.synchronize()


def test_scratch_growth_after_capture_is_refused():
    """A probe captured into a hipGraph pins its stream's scratch: a later,
    larger call on that stream returns DTL_ERUNTIME without enqueuing work or
    freeing the buffer, replays keep giving the eager result, and after
    dtl_scratch_release (once the graph is gone) the larger call runs."""
    cfg = (64, 1 << 25, 8, 8, 8, 1)
    f = O.Bloom(*cfg)
    desc = D.bbf_desc(*cfg)
    build = np.random.default_rng(31).integers(0, 2**31, 100000, dtype=np.uint64).astype(np.uint32)
    words = gpu_build(desc, build)
    ow = oracle_words(f, build)
    small = mixed_keys(build, 50000, 0.2, 32)
    large = mixed_keys(build, 1 << 23, 0.2, 33)   # needs more than the 1 MiB minimum scratch
    ks, kl = to_dev(small), to_dev(large)
    s = torch.cuda.Stream()
    out = torch.empty(1 << 23, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    with torch.cuda.stream(s):
        D.scratch_release(s)
        D.bbf_probe(desc, words, ks, 0, stream=s, out=out, count=cnt)   # warm-up sizes the scratch
    torch.cuda.synchronize()
    before, pinned = D.scratch_bytes(s)
    assert before > 0 and not pinned
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        D.bbf_probe(desc, words, ks, 0, stream=s, out=out, count=cnt)
    assert D.scratch_bytes(s) == (before, True)
    exp_small = f.batch_contains(ow, small, 0)
    with torch.cuda.stream(s):
        with pytest.raises(D.DtlError) as ei:
            D.bbf_probe(desc, words, kl, 0, stream=s, out=out, count=cnt)
    assert ei.value.status == D.ERUNTIME
    assert D.scratch_bytes(s) == (before, True)
    for _ in range(3):
        cnt.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(D.positions(out, cnt), exp_small)
    del g
    torch.cuda.synchronize()
    D.scratch_release(s)
    assert D.scratch_bytes(s) == (0, False)
    with torch.cuda.stream(s):
        D.bbf_probe(desc, words, kl, 0, stream=s, out=out, count=cnt)
    torch.cuda.synchronize()
    assert np.array_equal(D.positions(out, cnt), f.batch_contains(ow, large, 0))
    D.scratch_release(s)


def test_partitioned_probes_replay_from_a_graph():
    """The partitioned probes decide on the device alone (the cuckoo probe's
    gate on the first pass's match count, the Bloom probe's hit sort writing
    the positions and the count): captured into one hipGraph -- the cuckoo
    table at 1% and 50% hits (gate closed and open) and a Bloom filter --
    every replay, with eager work between replays, gives the oracle's lists."""
    m = (1 << 20) * 64
    cd = D.cf_desc(m, 16, 4)
    co = O.Cuckoo(m, 16, 4)
    ckeys = distinct_keys(int(0.95 * cd.bucket_cnt * 4), 51)
    cbuf, ins = D.cf_build_host(cd, ckeys, seed=51)
    assert ins == len(ckeys)
    n = (1 << 21) + 5
    cprobes = [mixed_keys(ckeys, n, hit, 52 + i) for i, hit in enumerate((0.01, 0.5))]
    cexp = [co.batch_contains(cbuf, p, 9) for p in cprobes]
    cfg = (64, 1 << 25, 8, 8, 8, 1)
    bd = D.bbf_desc(*cfg)
    bo = O.Bloom(*cfg)
    bkeys = distinct_keys(1 << 18, 53)
    words = gpu_build(bd, bkeys)
    bprobe = mixed_keys(bkeys, n, 0.05, 54)
    bexp = bo.batch_contains(oracle_words(bo, bkeys), bprobe, 11)
    dbuf = torch.from_numpy(cbuf).to(dev())
    ckd = [to_dev(p) for p in cprobes]
    bkd = to_dev(bprobe)
    cp = D.ProbeParams(0, 0, 0, 0, D.PROBE_PARTITIONED)
    cp.slice_log2 = 19      # 16 partitions with fixed-capacity bins
    bp = D.ProbeParams(0, 0, 0, 0, D.PROBE_PARTITIONED)
    bp.slice_log2 = 17      # 32 partitions with fixed-capacity bins
    outs = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(3)]
    cnts = [torch.zeros(1, dtype=torch.int64, device="cuda") for _ in range(3)]
    s = torch.cuda.Stream()

    def calls():
        for i in range(2):
            D.cf_probe(cd, dbuf, ckd[i], 9, stream=s, out=outs[i], count=cnts[i], params=cp)
        D.bbf_probe(bd, words, bkd, 11, stream=s, out=outs[2], count=cnts[2], params=bp)

    with torch.cuda.stream(s):
        calls()       # sizes the stream's scratch before the capture
    torch.cuda.synchronize()
    exp = cexp + [bexp]
    for i in range(3):
        assert np.array_equal(D.positions(outs[i], cnts[i]), exp[i]), ("eager", i)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        calls()
    for rep in range(3):
        for o, c in zip(outs, cnts):
            o.fill_(-1)
            c.fill_(-1)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for i in range(3):
            assert np.array_equal(D.positions(outs[i], cnts[i]), exp[i]), ("replay", rep, i)
    del g
    torch.cuda.synchronize()
    D.scratch_release(s)
