"""Full-size parity: every benchmark workload at the size it is benchmarked at,
the whole output compared with the oracle.

* cache-sectorized filter (u64, w=8, s=8, k=8) at 16 KiB, 4 MiB, 128 MiB and
  1 GiB, built from m/12 keys: filter words byte-compared with the oracle's
  build (threaded, block-sharded: cpu_baseline.bbf_build) and the list of all
  2^28 probe keys (1% hits) compared element by element with the oracle's
  threaded batch_contains (cpu_baseline.bbf_probe, pinned to the scalar
  oracle by tests/test_cpu_baseline.py), default and direct paths;
* the longest batches: 2^31 + 2^20 + 5 keys against a 64 MiB Bloom filter
  and a 16 MiB cuckoo table, with a match offset that wraps positions past
  2^32, default and direct paths;
* cuckoo filter (16-bit tags, 4 slots, 2^26 buckets) loaded to 95% by the
  host build, 2^28 probes at 1% and 50% hits, default (two-pass
  partitioned) and direct paths against the oracle's threaded probe;
* the 4096-filter batched semi-join (u32, w=16, z=2, k=6, MAGIC): every
  filter's words and every filter's list, from one graph-captured batch plan;
* the calibration grid of the learned configuration model: every candidate
  configuration at every calibrated size 2^13..2^33 bits, build and probe
  checked through the oracle's sparse build and its batch_contains on the
  gathered blocks (no dense host copy of the large filters).

Semantics: blocked_bloomfilter_batch_dispatch.hpp:28-69 (ascending i +
match_offset), cuckoofilter_logic.hpp:346-362, 390-479.  The inputs are the
benchmark's own seeded generators (bench.py).
"""
import concurrent.futures as cf_futures
import sys

import numpy as np
import pytest

import oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import bench  # noqa: E402
import cpu_baseline as CB  # noqa: E402
import dtl_hip as D  # noqa: E402


def say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def threads():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    D.load()
    return max(1, min(16, CB.physical_cores()))


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def cpu_list(fn, obj, data, keys, threads, offset=0):
    _, _, parts = fn(obj, data, keys, threads)
    out = np.concatenate(parts).astype(np.uint64) if parts else np.zeros(0, np.uint64)
    return out + offset


@pytest.mark.parametrize("log2_m", [17, 25, 30, 33])
def test_sectorized_filter_full_batch(log2_m, threads):
    cfg = (64, 1 << log2_m, 8, 8, 8, 1)
    desc = D.bbf_desc(*cfg)
    f = O.Bloom(*cfg)
    n_probe = 1 << 28
    say("fullsize: sectorized 2^%d bits, building" % log2_m)
    build, probe, n_build = bench.make_inputs(torch, log2_m, n_probe, 0.01, 1)
    words = D.new_filter(desc)
    D.bbf_build(desc, words, build)
    torch.cuda.synchronize()
    hb = host_u32(build)
    ow = CB.bbf_build(f, hb, threads)
    assert np.array_equal(words.cpu().numpy().view(np.uint64), ow), "filter words differ"
    del hb
    say("fullsize: sectorized 2^%d bits, probing" % log2_m)
    hp = host_u32(probe)
    exp = cpu_list(CB.bbf_probe, f, ow, hp, threads, 5)
    out = torch.empty(n_probe, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for flags in (0, D.PROBE_DIRECT):
        D.bbf_probe(desc, words, probe, 5, params=D.ProbeParams(0, 0, 0, 0, flags), out=out, count=cnt)
        torch.cuda.synchronize()
        got = D.positions(out, cnt).astype(np.uint64)
        assert len(got) == len(exp), (flags, len(got), len(exp))
        assert np.array_equal(got, exp), flags
    # size-independent properties of the same list: every planted build key
    # is found (no false negatives), positions strictly ascending
    assert len(exp) >= int(round(0.01 * n_probe))
    assert bool(np.all(exp[1:] > exp[:-1]))
    planted = hp[hp < (1 << 31)]          # probe keys drawn from the build set (absent ones have bit 31 set)
    assert len(planted) == int(round(0.01 * n_probe))
    assert len(f.batch_contains(ow, planted[:100000])) == min(len(planted), 100000)


def test_batch_beyond_2_31_keys(threads):
    """The longest batches: 2^31 + 2^20 + 5 keys against a 64 MiB filter, so
    that positions pass 2^31 and the partitioned probe sizes its partitions
    with the histogram pass (fixed-capacity bins stop at 2^30 keys), with a
    match offset of 2^31 + 12345 so that i + match_offset wraps past 2^32 as
    in the reference's 32-bit arithmetic (blocked_bloomfilter_batch_dispatch.hpp:
    47-67): the default (partitioned) and the direct path against the
    oracle's threaded list, element by element."""
    log2_m = 29
    cfg = (64, 1 << log2_m, 8, 8, 8, 1)
    desc = D.bbf_desc(*cfg)
    f = O.Bloom(*cfg)
    n_probe = (1 << 31) + (1 << 20) + 5
    off = (1 << 31) + 12345
    say("fullsize: 2^31 + 2^20 + 5 keys, building")
    # bench.make_inputs draws the hit positions with randperm, which stops at
    # 2^31 - 1 elements: here they are drawn with replacement (about 1% hits)
    n_build = (1 << log2_m) // 12
    build = bench.distinct_keys(torch, n_build, 6, 0)
    probe = bench.distinct_keys(torch, n_probe, 7, 1)
    g = torch.Generator(device="cuda").manual_seed(3)
    hits = n_probe // 100
    idx = torch.randint(0, n_probe, (hits,), device="cuda", generator=g)
    probe[idx] = build[torch.randint(0, n_build, (hits,), device="cuda", generator=g)]
    del idx
    words = D.new_filter(desc)
    D.bbf_build(desc, words, build)
    torch.cuda.synchronize()
    ow = CB.bbf_build(f, host_u32(build), threads)
    assert np.array_equal(words.cpu().numpy().view(np.uint64), ow), "filter words differ"
    del build
    say("fullsize: 2^31 + 2^20 + 5 keys, oracle")
    exp = cpu_list(CB.bbf_probe, f, ow, host_u32(probe), threads, off) & np.uint64(0xFFFFFFFF)
    assert len(exp) >= int(0.009 * n_probe)
    out = torch.empty(n_probe, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for flags in (0, D.PROBE_DIRECT):
        say("fullsize: 2^31 + 2^20 + 5 keys, probe flags %d" % flags)
        D.bbf_probe(desc, words, probe, off, params=D.ProbeParams(0, 0, 0, 0, flags), out=out, count=cnt)
        torch.cuda.synchronize()
        got = D.positions(out, cnt).astype(np.uint64)
        assert len(got) == len(exp), (flags, len(got), len(exp))
        assert np.array_equal(got, exp), flags
    del out, probe, words
    torch.cuda.empty_cache()


def test_cuckoo_batch_beyond_2_31_keys(threads):
    """The same batch length for the cuckoo probe: 2^31 + 2^20 + 5 keys
    (about 30% hits) against a 16 MiB table (2^18 buckets of 16-bit tags x 4
    slots) at 95% load from the host build, match offset 2^31 + 12345: the
    default two-pass partitioned probe (partitions sized by one histogram
    pass at this length; a match list of up to 2n positions before the hit
    sort merges the passes) and the direct probe against the oracle's
    threaded list (cuckoofilter_logic.hpp:390-479)."""
    log2_buckets = 18
    desc = D.cf_desc((1 << log2_buckets) * 64, 16, 4)
    cf = O.Cuckoo(desc.m, 16, 4)
    n_build = int(0.95 * (1 << log2_buckets) * 4)
    keys = bench.cuckoo_build_keys(n_build)
    buf, inserted = D.cf_build_host(desc, keys, seed=2)
    assert inserted == n_build
    dbuf = torch.from_numpy(buf).cuda()
    build_dev = torch.from_numpy(keys.view(np.int32)).cuda()
    n = (1 << 31) + (1 << 20) + 5
    off = (1 << 31) + 12345
    say("fullsize: cuckoo, 2^31 + 2^20 + 5 keys")
    probe = bench.distinct_keys(torch, n, 78, 1)
    g = torch.Generator(device="cuda").manual_seed(4)
    hits = int(0.3 * n)
    idx = torch.randint(0, n, (hits,), device="cuda", generator=g)
    probe[idx] = build_dev[torch.randint(0, n_build, (hits,), device="cuda", generator=g)]
    del idx
    exp = cpu_list(CB.cf_probe, cf, buf, host_u32(probe), threads, off) & np.uint64(0xFFFFFFFF)
    assert len(exp) >= int(0.25 * n)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for flags in (0, D.PROBE_DIRECT):
        say("fullsize: cuckoo, 2^31 + 2^20 + 5 keys, probe flags %d" % flags)
        D.cf_probe(desc, dbuf, probe, off, out=out, count=cnt, params=D.ProbeParams(0, 0, 0, 0, flags))
        torch.cuda.synchronize()
        got = D.positions(out, cnt).astype(np.uint64)
        assert len(got) == len(exp), (flags, len(got), len(exp))
        assert np.array_equal(got, exp), flags
    del out, probe, dbuf
    torch.cuda.empty_cache()


def test_cuckoo_benchmark_table_full_batch(threads):
    log2_buckets = 26
    desc = D.cf_desc((1 << log2_buckets) * 64, 16, 4)
    cf = O.Cuckoo(desc.m, 16, 4)
    n_build = int(0.95 * (1 << log2_buckets) * 4)
    say("fullsize: cuckoo host build of %d keys" % n_build)
    keys = bench.cuckoo_build_keys(n_build)
    buf, inserted = D.cf_build_host(desc, keys, seed=1)
    assert inserted == n_build
    vi, vt, vused = cf.victim(buf)
    assert cf.count_occupied(buf) + int(vused) == n_build
    dbuf = torch.from_numpy(buf).cuda()
    build_dev = torch.from_numpy(keys.view(np.int32)).cuda()
    n = 1 << 28
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    for hit in (0.01, 0.5):
        say("fullsize: cuckoo probe at %.0f%% hits" % (hit * 100))
        probe, hits = bench.cuckoo_probe_keys(torch, build_dev, n, hit)
        exp = cpu_list(CB.cf_probe, cf, buf, host_u32(probe), threads, 3)
        assert len(exp) >= hits
        for flags in (0, D.PROBE_DIRECT):
            D.cf_probe(desc, dbuf, probe, 3, out=out, count=cnt, params=D.ProbeParams(0, 0, 0, 0, flags))
            torch.cuda.synchronize()
            got = D.positions(out, cnt).astype(np.uint64)
            assert len(got) == len(exp), (hit, flags, len(got), len(exp))
            assert np.array_equal(got, exp), (hit, flags)
        del probe


def test_semi_join_4096_filters_every_list(threads):
    n_filters = 4096
    say("fullsize: semi-join inputs (%d filters)" % n_filters)
    meta, probes = bench.semi_join_inputs(torch, n_filters, 11)
    wb, k, w, s_, z = bench.FIXED_SEMI_JOIN_CFG
    tasks = bench.build_filters(torch, D, meta, probes, lambda f, nb: (wb, k, w, s_, z, 10 * nb))
    say("fullsize: semi-join plan (%d probe keys)" % sum(mt["n_probe"] for mt in meta))
    plan = D.BatchPlan(tasks)
    plan.tune(reps=1)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.run(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.run(s)
    for c in plan.counts:
        c.fill_(-1)
    torch.cuda.synchronize()
    graph.replay()
    torch.cuda.synchronize()
    say("fullsize: semi-join checking every filter")

    def check(i):
        desc, words, keys, off = tasks[i]
        nb = meta[i]["n_build"]
        f = O.Bloom(wb, 10 * nb, k, w, s_, z)
        assert f.block_cnt == desc.block_cnt
        ow = f.new_words()
        f.insert(ow, host_u32(bench.build_keys(torch, i, nb)))
        if not np.array_equal(words.cpu().numpy().view(np.uint32), ow):
            return "filter %d: words differ" % i
        exp = f.batch_contains(ow, host_u32(keys), off)
        got = D.positions(plan.outs[i], plan.counts[i])
        if not np.array_equal(got, exp):
            return "filter %d: list differs (%d vs %d matches)" % (i, len(got), len(exp))
        return None

    with cf_futures.ThreadPoolExecutor(threads) as ex:
        errors = [e for e in ex.map(check, range(n_filters)) if e]
    plan.close()
    del graph
    assert not errors, errors[:5]


def candidate_configs():
    """The learned model's candidate space (model/util.hpp:16-63 with
    z in {1, 2, 4}): word size x w <= 64 B, s <= w, s <= k, k % s == 0;
    zoned w >= 4, z < w, k even."""
    out = []
    for k in range(1, 17):
        for ws in (4, 8):
            for w in (1, 2, 4, 8, 16):
                if ws * w > 64:
                    continue
                for s in (1, 2, 4, 8, 16):
                    if s > w or s > k or k % s:
                        continue
                    out.append((ws * 8, w, s, 1, k))
                if w >= 4:
                    for z in (2, 4):
                        if z >= w or z > k or k % z or k & 1:
                            continue
                        out.append((ws * 8, w, w, z, k))
    return out


def test_learned_model_grid_every_configuration_and_size(threads):
    """Every (configuration, size) cell the calibration sweep measures: the
    candidate configurations at m = 2^b - 2 x block bits (MAGIC), b = 13..33,
    plus the power-of-two size 2^b (POW2).  Build: the GPU words at the keys'
    touched indexes equal the oracle's sparse build and no other word is
    non-zero.  Probe: the list equals the oracle's batch_contains on the
    gathered blocks; from 64 MiB on also through the partitioned probe."""
    cfgs = candidate_configs()
    assert len(cfgs) == 279
    big = torch.zeros((1 << 30) + 4096, dtype=torch.uint8, device="cuda")
    out = torch.empty(1 << 14, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    rng = np.random.default_rng(2026)
    cells = 0
    for b in range(13, 34):
        say("fullsize: learned-model grid, 2^%d bits" % b)
        for ci, (wb, w, s, z, k) in enumerate(cfgs):
            block_bits = wb * w
            for m in ((1 << b) - 2 * block_bits, 1 << b) if (ci + b) % 4 == 0 else ((1 << b) - 2 * block_bits,):
                m = max(m, 3 * block_bits)
                desc = D.bbf_desc(wb, m, k, w, s, z)
                f = O.Bloom(wb, m, k, w, s, z)
                nbytes = int(desc.size_bytes)
                words = big[:nbytes]
                words.zero_()
                n_build = int(min(max(m // 10, 16), 8192))
                build = rng.integers(0, 2**32, n_build, dtype=np.uint64).astype(np.uint32)
                D.bbf_build(desc, words, torch.from_numpy(build.view(np.int32)).cuda())
                idx, vals = f.sparse_words(build)
                wv = words.view(torch.int64 if wb == 64 else torch.int32)
                got = wv[torch.from_numpy(idx.astype(np.int64)).cuda()].cpu().numpy().view(vals.dtype)
                assert np.array_equal(got, vals), ("build", wb, w, s, z, k, m)
                assert int(torch.count_nonzero(wv).item()) == len(idx), ("stray bits", wb, w, s, z, k, m)
                n_probe = 4096
                probe = rng.integers(0, 2**32, n_probe, dtype=np.uint64).astype(np.uint32)
                probe[rng.integers(0, n_probe, n_probe // 4)] = build[rng.integers(0, n_build, n_probe // 4)]
                gathered = wv.view(-1, w)[torch.from_numpy(f.blocks_of(probe).astype(np.int64)).cuda()]
                exp = f.batch_contains_blocks(gathered.cpu().numpy().view(vals.dtype), probe, 17)
                kd = torch.from_numpy(probe.view(np.int32)).cuda()
                flag_sets = (0, D.PROBE_PARTITIONED) if nbytes >= (64 << 20) else (0,)
                for flags in flag_sets:
                    D.bbf_probe(desc, words, kd, 17, params=D.ProbeParams(0, 0, 0, 0, flags), out=out, count=cnt)
                    torch.cuda.synchronize()
                    assert np.array_equal(D.positions(out, cnt), exp), ("probe", flags, wb, w, s, z, k, m)
                cells += 1
    assert cells >= 279 * 21
    del big
