"""bbf_probe_noalloc (the probe with a fixed per-GPU scratch array and no
device allocation) against bbf_probe, through the C ABI: two calls in a row
per filter, the second with a smaller batch, for LDS-staged, lane-split,
one-lane and MAGIC-addressed filters; the size limit is refused on the host;
the per-stream scratch of bbf_probe is never touched; the call replays under
hipGraph capture."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import dtl_hip as D  # noqa: E402


def dev_keys(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def keys_for(build, n, seed):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    nh = n // 5
    k[rng.integers(0, n, nh)] = build[rng.integers(0, len(build), nh)]
    return k


CASES = [
    ((64, 1 << 17, 8, 8, 8, 1), 10000),          # 16 KiB: staged in LDS
    ((32, 1 << 20, 4, 1, 1, 1), 1 << 16),        # 128 KiB plain filter: LDS, 1024-thread workgroups
    ((64, 1 << 25, 8, 8, 8, 1), 200000),         # 4 MiB: lane-split from L2
    ((32, 10 * 300007, 6, 16, 16, 2), 300007),   # zoned, MAGIC addressing
    ((64, 1 << 32, 8, 8, 8, 1), 1 << 20),        # 512 MiB: one lane per key from HBM
]


@pytest.mark.parametrize("cfg,n_build", CASES)
def test_noalloc_matches_bbf_probe(cfg, n_build):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    desc = D.bbf_desc(*cfg)
    rng = np.random.default_rng(n_build)
    build = rng.integers(0, 2**31, n_build, dtype=np.uint64).astype(np.uint32)
    words = D.new_filter(desc)
    D.bbf_build(desc, words, dev_keys(build))
    s = torch.cuda.Stream()
    for i, n in enumerate(((1 << 22) + 17, 100003)):   # the second call is smaller
        k = dev_keys(keys_for(build, n, 10 * n_build + i))
        with torch.cuda.stream(s):
            ref_out, ref_cnt = D.bbf_probe(desc, words, k, 7 + i)
            out, cnt = D.bbf_probe_noalloc(desc, words, k, 7 + i)
        torch.cuda.synchronize()
        ref = D.positions(ref_out, ref_cnt)
        got = D.positions(out, cnt)
        assert len(ref) > 0
        assert np.array_equal(got, ref), (cfg, n)


def test_noalloc_limit_zero_batch_and_no_library_scratch():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = D.load()
    assert L.dtl_noalloc_max_keys() == 1 << 28
    desc = D.bbf_desc(64, 1 << 25, 8, 8, 8, 1)
    words = D.new_filter(desc)
    s = torch.cuda.Stream()
    cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    import ctypes as C
    fake = C.c_void_p(16)   # never dereferenced: the size check comes first
    with torch.cuda.stream(s):
        rc = L.bbf_probe_noalloc(C.byref(desc), C.c_void_p(words.data_ptr()), fake, (1 << 28) + 1, 0, fake,
                                 C.c_void_p(cnt.data_ptr()), C.c_void_p(s.cuda_stream))
        assert rc == D.EINVAL
        assert "2^28" in L.dtl_last_error().decode()
        empty = torch.empty(0, dtype=torch.int32, device="cuda")
        out, cnt = D.bbf_probe_noalloc(desc, words, empty, 0, count=cnt)
        k = dev_keys(np.arange(5000, dtype=np.uint32))
        D.bbf_probe_noalloc(desc, words, k, 0)
    torch.cuda.synchronize()
    assert int(cnt.item()) == 0
    assert D.scratch_bytes(s) == (0, False)     # bbf_probe's per-stream scratch never allocated


def test_noalloc_replays_under_graph_capture():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    desc = D.bbf_desc(32, 10 * 50021, 6, 16, 16, 2)
    build = np.random.default_rng(3).integers(0, 2**31, 50021, dtype=np.uint64).astype(np.uint32)
    words = D.new_filter(desc)
    D.bbf_build(desc, words, dev_keys(build))
    k = dev_keys(keys_for(build, 300001, 4))
    s = torch.cuda.Stream()
    out = torch.empty(300001, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    with torch.cuda.stream(s):
        D.bbf_probe_noalloc(desc, words, k, 3, out=out, count=cnt)   # first call on the device: address lookup
    torch.cuda.synchronize()
    ref = D.positions(out, cnt)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        D.bbf_probe_noalloc(desc, words, k, 3, out=out, count=cnt)
    other = dev_keys(keys_for(build, 70001, 5))
    for _ in range(3):
        out.fill_(-1)
        cnt.fill_(-1)
        with torch.cuda.stream(s):
            D.bbf_probe_noalloc(desc, words, other, 0)   # eager work on the same array between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(D.positions(out, cnt), ref)
