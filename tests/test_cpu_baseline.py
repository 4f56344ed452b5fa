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
