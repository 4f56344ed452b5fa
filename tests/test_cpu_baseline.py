"""The CPU baseline timed by bench.py (multithreaded AVX-512 / AVX2
restatement of the reference's SIMD batch_contains, workers pinned one per
physical core) returns exactly the scalar oracle's match list when its
per-thread lists are concatenated: for each hand-vectorized configuration
(the four benchmark filters), each instruction set, thread counts that leave
ragged slices, and batch lengths that leave scalar tails."""
import numpy as np
import pytest

import cpu_baseline as CB
import oracle as O

ISAS = [isa for isa, bit in (("avx512", 2), ("avx2", 1))]

BBF_CASES = [
    # (config, build keys): plain register-blocked, cache-sectorized, zoned MAGIC
    ((32, 1 << 20, 4, 1, 1, 1), 1 << 16),
    ((64, 1 << 22, 8, 8, 8, 1), (1 << 22) // 12),
    ((32, 10 * 30011, 6, 16, 16, 2), 30011),
    ((32, 10 * 1000003, 6, 16, 16, 2), 1000003),
]


def _needs(isa, avail):
    bit = 2 if isa == "avx512" else 1
    if not avail & bit:
        pytest.skip("%s not available on this CPU" % isa)


@pytest.mark.parametrize("cfg,n_build", BBF_CASES)
@pytest.mark.parametrize("isa", ISAS)
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_bbf_baseline_matches_oracle(cfg, n_build, isa, threads):
    f = O.Bloom(*cfg)
    _needs(isa, CB.bbf_simd_available(f))
    rng = np.random.default_rng(n_build)
    build = rng.integers(0, 2**31, n_build, dtype=np.uint64).astype(np.uint32)
    words = f.new_words()
    f.insert(words, build)
    probe = rng.integers(0, 2**32, 200001, dtype=np.uint64).astype(np.uint32)
    probe[::7] = build[rng.integers(0, n_build, len(probe[::7]))]
    r = CB.run("bbf", f, words, probe, threads, isa)
    assert r.isa == {"avx512": "AVX-512", "avx2": "AVX2"}[isa]
    got = np.concatenate(r.lists)
    assert r.count == len(got)
    exp = f.batch_contains(words, probe)
    assert len(exp) > 20000
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("isa", ISAS)
@pytest.mark.parametrize("threads", [1, 5])
def test_cf_baseline_matches_oracle(isa, threads):
    cf = O.Cuckoo(1 << 22, 16, 4)
    _needs(isa, CB.cf_simd_available(cf))
    buf = cf.new_buffer()
    rng = np.random.default_rng(2)
    keys = np.unique(rng.integers(0, 2**31, 80000, dtype=np.uint64).astype(np.uint32))[:62000]
    cf.insert(buf, keys, O.MT19937(3))
    probe = rng.integers(0, 2**32, 100003, dtype=np.uint64).astype(np.uint32)
    probe[::2] = keys[: len(probe[::2])]
    r = CB.run("cf", cf, buf, probe, threads, isa)
    assert np.array_equal(np.concatenate(r.lists), cf.batch_contains(buf, probe))


@pytest.mark.parametrize("isa", ISAS)
def test_cf_baseline_victim_cache(isa):
    """A table filled past its capacity uses the victim cache; the SIMD
    bodies report victim hits like the oracle."""
    cf = O.Cuckoo(1 << 13, 16, 4)
    _needs(isa, CB.cf_simd_available(cf))
    buf = cf.new_buffer()
    keys = np.unique(np.random.default_rng(4).integers(0, 2**31, 1400, dtype=np.uint64).astype(np.uint32))
    cf.insert(buf, keys, O.MT19937(5))
    assert cf.victim(buf)[2]
    probe = np.concatenate([keys, np.random.default_rng(6).integers(0, 2**32, 5003, dtype=np.uint64).astype(np.uint32)])
    r = CB.run("cf", cf, buf, probe, 3, isa)
    assert np.array_equal(np.concatenate(r.lists), cf.batch_contains(buf, probe))


def test_generic_configs_fall_back_to_scalar():
    f = O.Bloom(32, 10 * 5000, 8, 16, 8, 1)
    assert not CB.bbf_simd_available(f)
    rng = np.random.default_rng(3)
    build = rng.integers(0, 2**31, 5000, dtype=np.uint64).astype(np.uint32)
    words = f.new_words()
    f.insert(words, build)
    probe = np.concatenate([build, rng.integers(0, 2**32, 9000, dtype=np.uint64).astype(np.uint32)])
    r = CB.run("bbf", f, words, probe, 4)
    assert r.isa == "scalar"
    assert np.array_equal(np.concatenate(r.lists), f.batch_contains(words, probe))


def test_workers_are_pinned_one_per_core():
    import os
    assert 1 <= CB.physical_cores() <= len(os.sched_getaffinity(0))


@pytest.mark.parametrize("cfg", [(64, 1 << 22, 8, 8, 8, 1), (32, 10 * 30011, 6, 16, 16, 2),
                                 (32, 1 << 20, 4, 1, 1, 1), (64, 3 << 20, 16, 16, 2, 1)])
def test_threaded_build_matches_sequential(cfg):
    """The block-sharded threaded build (a checker for full-size GPU builds)
    equals the oracle's sequential insert."""
    f = O.Bloom(*cfg)
    keys = np.random.default_rng(4).integers(0, 2**32, 70000, dtype=np.uint64).astype(np.uint32)
    seq = f.new_words()
    f.insert(seq, keys)
    for threads in (1, 3, 8):
        assert np.array_equal(CB.bbf_build(f, keys, threads), seq)


@pytest.mark.parametrize("cfg", [(64, 1 << 26, 8, 8, 8, 1), (32, (1 << 27) - 5, 6, 16, 16, 2),
                                 (64, 1 << 20, 5, 2, 1, 1), (32, 1 << 30, 16, 8, 2, 1)])
def test_sparse_and_block_checkers_match_dense_oracle(cfg):
    """The oracle's sparse build (touched words only) and its batch_contains
    on gathered blocks agree with the dense build and batch_contains."""
    f = O.Bloom(*cfg)
    rng = np.random.default_rng(8)
    keys = rng.integers(0, 2**31, 5000, dtype=np.uint64).astype(np.uint32)
    dense = f.new_words()
    f.insert(dense, keys)
    idx, vals = f.sparse_words(keys)
    assert np.count_nonzero(dense) == len(idx)
    assert np.array_equal(dense[idx.astype(np.int64)], vals)
    probe = np.concatenate([keys[:2000], rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)])
    rng.shuffle(probe)
    blocks = f.blocks_of(probe).astype(np.int64)
    w = f.d.w
    gathered = dense.reshape(-1, w)[blocks]
    assert np.array_equal(f.batch_contains_blocks(gathered, probe, 9), f.batch_contains(dense, probe, 9))


@pytest.mark.parametrize("isa", ISAS)
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_probe_tasks_matches_oracle_per_filter(isa, threads):
    """The semi-join baseline (pinned workers taking whole filters) returns
    each filter's own list, positions from 0, equal to the scalar oracle's."""
    rng = np.random.default_rng(threads)
    blooms, words, keys, exp = [], [], [], []
    for i, n_build in enumerate([1 << 10, 3001, 1 << 14, 70001, 60, 1 << 17]):
        f = O.Bloom(32, 10 * n_build, 6, 16, 16, 2)
        if i == 0:
            _needs(isa, CB.bbf_simd_available(f))
        build = rng.integers(0, 2**31, n_build, dtype=np.uint64).astype(np.uint32)
        w = f.new_words()
        f.insert(w, build)
        n_probe = int(rng.integers(1, 40000))
        probe = rng.integers(0, 2**32, n_probe, dtype=np.uint64).astype(np.uint32)
        probe[::5] = build[rng.integers(0, n_build, len(probe[::5]))]
        blooms.append(f)
        words.append(w.view(np.uint32))
        keys.append(probe)
        exp.append(f.batch_contains(w, probe))
    lists, secs = CB.probe_tasks(blooms, words, keys, threads, isa)
    assert secs > 0
    assert len(lists) == len(exp)
    for got, e in zip(lists, exp):
        assert np.array_equal(got, e)
