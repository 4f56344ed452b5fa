"""The CPU baseline timed by bench.py (multithreaded AVX-512 restatement of
the reference's SIMD batch_contains) returns exactly the scalar oracle's
match list when its per-thread lists are concatenated."""
import numpy as np
import pytest

import cpu_baseline as CB
import oracle as O


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_bbf_baseline_matches_oracle(threads):
    f = O.Bloom(64, 1 << 22, 8, 8, 8)
    rng = np.random.default_rng(1)
    build = rng.integers(0, 2**31, (1 << 22) // 12, dtype=np.uint64).astype(np.uint32)
    words = f.new_words()
    f.insert(words, build)
    probe = rng.integers(2**31, 2**32, 200001, dtype=np.uint64).astype(np.uint32)
    probe[::40] = build[: len(probe[::40])]
    total, secs, lists = CB.bbf_probe(f, words, probe, threads)
    got = np.concatenate(lists)
    assert total == len(got)
    assert np.array_equal(got, f.batch_contains(words, probe))


@pytest.mark.parametrize("threads", [1, 5])
def test_cf_baseline_matches_oracle(threads):
    cf = O.Cuckoo(1 << 22, 16, 4)
    buf = cf.new_buffer()
    rng = np.random.default_rng(2)
    keys = np.unique(rng.integers(0, 2**31, 80000, dtype=np.uint64).astype(np.uint32))[:62000]
    cf.insert(buf, keys, O.MT19937(3))
    probe = rng.integers(0, 2**32, 100003, dtype=np.uint64).astype(np.uint32)
    probe[::2] = keys[: len(probe[::2])]
    total, secs, lists = CB.cf_probe(cf, buf, probe, threads)
    assert np.array_equal(np.concatenate(lists), cf.batch_contains(buf, probe))


def test_generic_configs_fall_back_to_scalar():
    f = O.Bloom(32, 10 * 5000, 6, 16, 16, 2)
    assert not CB.bbf_simd_available(f)
    rng = np.random.default_rng(3)
    build = rng.integers(0, 2**31, 5000, dtype=np.uint64).astype(np.uint32)
    words = f.new_words()
    f.insert(words, build)
    probe = np.concatenate([build, rng.integers(0, 2**32, 9000, dtype=np.uint64).astype(np.uint32)])
    total, secs, lists = CB.bbf_probe(f, words, probe, 4)
    assert np.array_equal(np.concatenate(lists), f.batch_contains(words, probe))


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
