#!/usr/bin/env python3
"""Benchmark of the filter probe path on one MI355X.

Prints one JSON line.  The headline ("value", one "step" = one bbf_probe call
over one batch): the cache-sectorized blocked Bloom filter -- 64-bit words,
w=8 words (512-bit blocks), s=8 sectors, k=8, m = 2^33 bits (1 GiB), built
from m/12 distinct uint32 keys -- probed with 2^28 uint32 keys at a 1% hit
rate; output is the ascending match-position list.  Inputs are resident in
HBM before timing.  The same line carries every other workload of the metric,
each with its roofline, a CPU baseline timed on this host's cores and an
element-wise comparison of its list with the CPU baseline's (outside the
timed region):
  regimes            the same filter at 16 KiB, 4 MiB and 128 MiB (2^28 keys), and
                     the plain register-blocked filter (u32, w=1, k=4, 128 KiB,
                     2^16 keys; 2^24 probes at 10% hits), the latter also as 16
                     calls captured in one hipGraph;
  cuckoo             16-bit tags, 4 slots, 2^26 buckets at 95% load (host build),
                     2^28 probes at 1% and 50% hits;
  learned_config     the calibration sweep, the MLP and its regret against the
                     exhaustive choice;
  batched_semi_join  4096 zoned partition filters: plain per-filter calls of the
                     fixed configuration vs one tuned, graph-captured batch plan,
                     and the learned vs exhaustive configurations in batch plans,
                     total filter cost at tw = 1, 10, 100 and 1000 GPU cycles
                     (at 2.4 GHz) per qualifying tuple.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bloomfilter-bsd_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
GPU_CLOCK_GHZ = 2.4     # MI355X peak engine clock: converts tw in GPU cycles to ns
TW_CYCLES = (1.0, 10.0, 100.0, 1000.0)
LDS_STAGE_MAX = 160 * 1024 - 2048   # largest filter a batch plan stages in LDS by default
METRIC = "G probes/s on 1 MI355X + % random-access roofline; tuned vs plain config"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def distinct_keys(torch, n, seed, high_bit):
    """n distinct uint32 keys (as int32) from a seeded bijection of [0, 2^31):
    x = (i + 2^20 * seed) mod 2^31, then two odd multiplications and xor-shifts
    modulo 2^31 (each step is invertible), bit 31 set to `high_bit`.  Build
    keys use high_bit = 0, keys known to be absent use high_bit = 1."""
    M = (1 << 31) - 1
    x = (torch.arange(n, dtype=torch.int64, device="cuda") + (seed << 20)) & M
    x = (x * 0x5851F42D) & M
    x = x ^ (x >> 16)
    x = (x * 0x2C1B3C6D) & M
    x = x ^ (x >> 13)
    x = (x * 0x297A2D39) & M
    if high_bit:
        x = x | (1 << 31)
    return x.to(torch.int32)


def make_inputs(torch, log2_m, n_probe, hit_rate, seed, n_build=None):
    """Seeded synthetic inputs on the GPU: n_build (default m/12) distinct
    build keys in [0, 2^31), probe keys distinct in [2^31, 2^32) with
    round(hit_rate * n) of them replaced (at random positions) by random build
    keys."""
    if n_build is None:
        n_build = (1 << log2_m) // 12
    g = torch.Generator(device="cuda").manual_seed(seed)
    build = distinct_keys(torch, n_build, 2 * seed, 0)
    probe = distinct_keys(torch, n_probe, 2 * seed + 1, 1)
    hits = int(round(hit_rate * n_probe))
    idx = torch.randperm(n_probe, device="cuda", generator=g)[:hits]
    probe[idx] = build[torch.randint(0, n_build, (hits,), device="cuda", generator=g)]
    return build, probe, n_build


def time_probe(torch, D, desc, words, probe, steps, warmup, params=None, per_step=None, warm_ms=0.0, out=None):
    """Returns (ms per step, avg ms of the probe phase, matches, positions
    tensor).  If per_step is a dict it receives the median, min and max of
    the per-call HIP-event times (ms).  Warm-up runs `warmup` calls and, with
    warm_ms, keeps calling until that much wall time has passed (GPU clocks
    ramp up over tens of milliseconds after idle)."""
    n = probe.numel()
    if out is None:
        out = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in evs:  # materialise the events
        a.record(stream)
        b.record(stream)
    for _ in range(warmup):
        D.bbf_probe(desc, words, probe, 0, params=params, out=out, count=cnt)
    torch.cuda.synchronize()
    w0 = time.time()
    while (time.time() - w0) * 1e3 < warm_ms:
        D.bbf_probe(desc, words, probe, 0, params=params, out=out, count=cnt)
        torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for i in range(steps):
        p = D.ProbeParams() if params is None else D.ProbeParams(
            params.lds_max_bytes, params.block_threads, params.grid_blocks, params.chunk_tiles, params.flags)
        if params is not None:
            p.slice_log2 = params.slice_log2
        p.ev_start = evs[i][0].cuda_event
        p.ev_stop = evs[i][1].cuda_event
        D.bbf_probe(desc, words, probe, 0, params=p, out=out, count=cnt)
    t1.record(stream)
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    each = sorted(a.elapsed_time(b) for a, b in evs)
    kern_ms = sum(each) / steps
    if per_step is not None:
        mid = len(each) // 2
        per_step.update({"median": round(each[mid] if len(each) % 2 else 0.5 * (each[mid - 1] + each[mid]), 4),
                         "min": round(each[0], 4), "max": round(each[-1], 4), "calls": len(each)})
    c = int(cnt.item())
    return total_ms / steps, kern_ms, c, out[:c]


def host_cpu_topology():
    """CPUs of this process's share, physical cores among them and SMT
    threads per core (from /sys, as lscpu reports them)."""
    cpus = sorted(os.sched_getaffinity(0))
    sib = {}
    for c in cpus:
        try:
            with open("/sys/devices/system/cpu/cpu%d/topology/thread_siblings_list" % c) as fh:
                sib.setdefault(fh.read().strip(), []).append(c)
        except OSError:
            sib.setdefault(str(c), []).append(c)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    per_core = max(len(v) for v in sib.values()) if sib else 1
    return {"cpus": len(cpus), "physical_cores": len(sib), "smt_threads_per_core": per_core,
            "smt": per_core > 1, "model": model}


def cpu_best(run_once, min_seconds=0.25, min_reps=3, max_reps=50):
    """Best wall time over repetitions that last at least min_seconds in all
    (benchmark.cpp:166-193); returns (best seconds, last result, reps)."""
    best, total, reps, last = None, 0.0, 0, None
    while reps < min_reps or (total < min_seconds and reps < max_reps):
        last = run_once()
        best = last.seconds if best is None else min(best, last.seconds)
        total += last.seconds
        reps += 1
    return best, last, reps


def cpu_baseline(CB, kind, obj, data, keys_np, gpu_positions, what, one_thread=False):
    """The oracle's threaded AVX-512 / AVX2 restatement of the reference's
    batch_contains (oracle/cpu_simd.cpp) over the whole batch `keys_np`
    against the same filter in host memory: one pinned worker per physical
    core of this process's CPU share (at most 16), each probing a contiguous
    slice in 1024-key batches into its own output (benchmark.cpp:123-164);
    best of repetitions lasting >= 250 ms.  The concatenated list is compared
    element by element with the GPU's (`gpu_positions`, uint64)."""
    import numpy as np
    threads = min(CB.physical_cores(), 16)
    n = len(keys_np)
    rates = {}
    last = None
    for isa in ("avx512", "avx2"):
        best, r, reps = cpu_best(lambda: CB.run(kind, obj, data, keys_np, threads, isa))
        rates[r.isa] = (n / best / 1e9, best, reps)
        if last is None:
            last = r
    identical = bool(np.array_equal(last.positions(), gpu_positions))
    top = max(rates, key=lambda k: rates[k][0])
    res = {"value": round(rates[top][0], 4), "unit": "G probes/s", "cores": threads, "kind": "port",
           "isa": top, "by_isa_G_probes_per_s": {k: round(v[0], 4) for k, v in rates.items()},
           "host_cpu": host_cpu_topology(), "lists_identical_to_gpu": identical,
           "sample": "the whole batch: %d probe keys against %s in host memory; %s batch_contains, %d pinned "
                     "threads, best of %d runs = %.4f s" % (n, what, top, threads, rates[top][2], rates[top][1])}
    if one_thread:
        r1 = CB.run(kind, obj, data, keys_np, 1, "avx512")
        res["one_thread_G_probes_per_s"] = round(n / r1.seconds / 1e9, 4)
    return res


GATHER_MODES = {0: "1 lane / 64 B sector, 4x16 B loads", 1: "as 0, non-temporal",
                2: "4 lanes / 64 B sector, 1x16 B each", 3: "as 2, non-temporal",
                4: "1 lane / 32 B sector, 2x16 B loads", 5: "2 lanes / 32 B sector",
                6: "as 2, uniform random sectors", 7: "as 0, uniform random sectors",
                8: "LDS-staged buffer, 1 lane / pseudo-random 4 B word (per-lane LCG)",
                9: "LDS-staged buffer, 1 lane / pseudo-random 64 B block, 4x16 B reads in the probe's rotated layout"}


def gather_table(torch, D, modes, n=1 << 28, reps=5, log2_bytes=30):
    """Random-sector read microbenchmark over a 2^log2_bytes buffer (dtl_gather_bench)."""
    buf = torch.randint(0, 2**31, (1 << (log2_bytes - 2),), dtype=torch.int32, device="cuda")
    sink = torch.zeros(1, dtype=torch.int64, device="cuda")
    out = {}
    for mode in modes:
        sector = 4 if mode == 8 else (32 if mode in (4, 5) else 64)
        for _ in range(2):
            D.gather_bench(buf, mode, n, sink)
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            D.gather_bench(buf, mode, n, sink)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / reps
        out[mode] = {"pattern": GATHER_MODES[mode], "ms": round(ms, 4),
                     "G_sectors_per_s": round(n / ms / 1e6, 2),
                     "useful_GB_per_s": round(n * sector / ms / 1e6, 1)}
    del buf
    return out


def graph_replay_ms(torch, D, desc, words, probe, calls, reps, expect):
    """Per-call time of `calls` back-to-back bbf_probe calls captured in one
    hipGraph and replayed `reps` times (the caller's launch gaps between calls
    go away; the kernels are the same).  The list of the last call is checked
    against `expect` after the replays."""
    n = probe.numel()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        D.bbf_probe(desc, words, probe, 0, out=out, count=cnt)   # sizes the stream's scratch before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for _ in range(calls):
            D.bbf_probe(desc, words, probe, 0, out=out, count=cnt)
    g.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        g.replay()
    b.record()
    torch.cuda.synchronize()
    c = int(cnt.item())
    same = c == expect.numel() and bool(torch.equal(out[:c], expect))
    del g
    return a.elapsed_time(b) / (reps * calls), same


def regime(torch, D, CB, O, word_bits, log2_m, k, w, s, z, n_probe, hit_rate, steps, warmup, label,
           n_build=None, no_cpu=False, graph_calls=0):
    """Whole-job probe rate of one filter configuration and size, with the
    CPU baseline on the same batch and their lists compared.  With
    graph_calls, also the rate of that many calls captured in one hipGraph."""
    import numpy as np
    desc = D.bbf_desc(word_bits, 1 << log2_m, k, w, s, z)
    build, probe, n_build = make_inputs(torch, log2_m, n_probe, hit_rate, 7, n_build=n_build)
    words = D.new_filter(desc)
    D.bbf_build(desc, words, build)
    torch.cuda.synchronize()
    st = {}
    ms, kern_ms, matches, pos = time_probe(torch, D, desc, words, probe, steps, warmup, per_step=st, warm_ms=200.0)
    graph = None
    if graph_calls:
        g_ms, g_same = graph_replay_ms(torch, D, desc, words, probe, graph_calls, 20, pos)
        graph = {"calls_per_graph": graph_calls, "replays": 20, "ms_per_call": round(g_ms, 4),
                 "G_probes_per_s": round(n_probe / (g_ms * 1e-3) / 1e9, 2), "list_identical": g_same}
    bpk = 4.0 + 4.0 * matches / n_probe   # key + match from HBM; the filter is read once per workgroup or from cache
    res = {"config": label, "filter_bytes": int(desc.size_bytes),
           "path": "lds" if desc.size_bytes <= 158 * 1024 else "global",
           "G_probes_per_s": round(n_probe / (ms * 1e-3) / 1e9, 2),
           "kernel_G_probes_per_s": round(n_probe / (kern_ms * 1e-3) / 1e9, 2),
           "call_ms": st, "matches": matches,
           "hbm_bytes_per_probe": round(bpk + (0.0 if desc.size_bytes <= 158 * 1024 else 64.0), 3)}
    if graph is not None:
        res["graph_replay"] = graph
    if not no_cpu:
        f = O.Bloom(word_bits, 1 << log2_m, k, w, s, z)
        gpu = pos.cpu().numpy().view(np.uint32).astype(np.uint64)
        res["cpu_baseline"] = cpu_baseline(CB, "bbf", f, words.cpu().numpy().view(f.dtype),
                                           probe.cpu().numpy().view(np.uint32), gpu, "the same filter")
    del words, build, probe, pos
    torch.cuda.empty_cache()
    return res


def run_regimes(torch, D, CB, O, steps_arg, no_cpu=False):
    """The filter-size sweep of the sectorized filter: 16 KiB (LDS), 4 MiB
    (L2), 128 MiB (Infinity Cache), and the plain register-blocked 128 KiB
    filter (2^16 build keys, 2^24 probes at 10% hits), with the LDS random
    4-byte read rate beside the two LDS-staged ones."""
    log("bench: filter-size regimes")
    steps = max(3, steps_arg // 2)
    warm = 3
    regs = [regime(torch, D, CB, O, 64, 17, 8, 8, 8, 1, 1 << 28, 0.01, steps, warm, "u64 w8 s8 k8, 16 KiB",
                   no_cpu=no_cpu),
            regime(torch, D, CB, O, 64, 25, 8, 8, 8, 1, 1 << 28, 0.01, steps, warm, "u64 w8 s8 k8, 4 MiB",
                   no_cpu=no_cpu),
            regime(torch, D, CB, O, 64, 30, 8, 8, 8, 1, 1 << 28, 0.01, steps, warm, "u64 w8 s8 k8, 128 MiB",
                   no_cpu=no_cpu),
            regime(torch, D, CB, O, 32, 20, 4, 1, 1, 1, 1 << 24, 0.10, steps, warm,
                   "plain register-blocked u32 w1 s1 k4, 128 KiB (2^16 keys), 2^24 probes at 10% hits",
                   n_build=1 << 16, no_cpu=no_cpu, graph_calls=16)]
    # the LDS-staged regimes against the measured LDS read rate of their own
    # access: one random 64 B block per probe for the 16 KiB u64 w8 filter
    # (mode 9), one random 4-byte word for the 128 KiB u32 w1 filter (mode 8)
    for r, lb, mode, what in ((regs[0], 14, 9, "64 B block"), (regs[3], 17, 8, "4-byte word")):
        lds = gather_table(torch, D, [mode], reps=3, log2_bytes=lb)[mode]
        r["roofline"] = {"bound": "lds", "achieved": round(r["kernel_G_probes_per_s"], 2),
                         "peak": lds["G_sectors_per_s"], "unit": "G random %s LDS reads/s" % what,
                         "frac": round(r["kernel_G_probes_per_s"] / lds["G_sectors_per_s"], 3),
                         "lds_GB_per_s": lds["useful_GB_per_s"],
                         "note": "one random LDS %s read per probe; the LDS rate is measured "
                                 "(dtl_gather_bench mode %d, buffer of %d bytes); the key stream "
                                 "from HBM (4 B per probe) comes on top" % (what, mode, 1 << lb)}
        hbm = r["kernel_G_probes_per_s"] * r["hbm_bytes_per_probe"]
        r["roofline"]["key_stream_hbm"] = {"achieved_GB_per_s": round(hbm, 1), "peak": HBM_PEAK_GBS,
                                           "frac": round(hbm / HBM_PEAK_GBS, 4)}
    for r in regs[1:3]:
        bpk = 4.0 + 64.0 + 4.0 * r["matches"] / (1 << 28)
        ach = bpk * r["kernel_G_probes_per_s"]
        r["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_probe": round(bpk, 3),
                         "note": "4 B key + 64 B block + 4 B per match; the block is read from L2 / the "
                                 "Infinity Cache, so this is an equivalent figure"}
    return regs


FIXED_SEMI_JOIN_CFG = (32, 6, 16, 16, 2)   # (word_bits, k, w, s, z): u32 words, w=16, z=2, k=6


def semi_join_sizes(n_filters, seed, log2_build=(10, 22), log2_probe=(12, 24), hit=(0.001, 0.30)):
    """Per-partition sizes of the batched semi-join workload, seeded: build
    sizes log-uniform in 2^10..2^22 keys, probe batches log-uniform in
    2^12..2^24 keys, hit rates uniform in [0.1%, 30%].  Returns a list of
    (n_build, n_probe, hit_rate)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_filters):
        nb = int(2 ** rng.uniform(*log2_build))
        npb = int(2 ** rng.uniform(*log2_probe))
        h = float(rng.uniform(*hit))
        out.append((nb, npb, h))
    return out


def semi_join_inputs(torch, n_filters, seed):
    """Keys of the batched semi-join workload (sizes from semi_join_sizes):
    probe keys distinct and known absent, round(rate x n_probe) of them
    replaced at random positions by random build keys.  Build keys are
    regenerated on demand (build_keys) so that only probe keys stay
    resident."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    meta, probes = [], []
    for f, (nb, npb, h) in enumerate(semi_join_sizes(n_filters, seed)):
        build = build_keys(torch, f, nb)
        probe = distinct_keys(torch, npb, 3 * f + 100001, 1)
        hits = int(round(h * npb))
        if hits:
            idx = torch.randperm(npb, device="cuda", generator=g)[:hits]
            probe[idx] = build[torch.randint(0, nb, (hits,), device="cuda", generator=g)]
        probes.append(probe)
        meta.append({"n_build": nb, "n_probe": npb, "hits": hits})
    torch.cuda.synchronize()
    return meta, probes


def build_keys(torch, f, nb):
    return distinct_keys(torch, nb, 3 * f + 100000, 0)


def build_filters(torch, D, meta, probes, choose):
    """One filter per partition; choose(f, n_build) -> (word_bits, k, w, s, z, m)."""
    tasks = []
    for f, (mt, probe) in enumerate(zip(meta, probes)):
        wb, k, w, s_, z, m = choose(f, mt["n_build"])
        desc = D.bbf_desc(wb, m, k, w, s_, z)
        words = D.new_filter(desc)
        D.bbf_build(desc, words, build_keys(torch, f, mt["n_build"]))
        tasks.append((desc, words, probe, 0))
    torch.cuda.synchronize()
    return tasks


def time_plan_graph(torch, plan, steps):
    """Captures one run of a (tuned) batch plan into a hipGraph and times
    replays; returns (ms per replay, per-task counts of the last replay)."""
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.run(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        plan.run(s)
    graph.replay()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    stream = torch.cuda.current_stream()
    a.record(stream)
    for _ in range(steps):
        graph.replay()
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    counts = [int(c.item()) for c in plan.counts]
    del graph
    return ms, counts


def semi_join_cpu_baseline(torch, CB, O, tasks, meta, gpu_lists, sample, seed=5):
    """The CPU baseline of the semi-join on a seeded sample of `sample`
    filters (their whole probe batches): pinned workers, one per physical
    core, each probing whole filters (largest first) with the AVX-512 body of
    the fixed configuration; best of repetitions lasting >= 250 ms.  Each
    filter's list is compared with the GPU's."""
    import numpy as np
    rng = np.random.default_rng(seed)
    idx = sorted(rng.choice(len(tasks), size=min(sample, len(tasks)), replace=False).tolist(),
                 key=lambda i: -meta[i]["n_probe"])
    wb, k, w, s_, z = FIXED_SEMI_JOIN_CFG
    blooms = [O.Bloom(wb, 10 * meta[i]["n_build"], k, w, s_, z) for i in idx]
    words = [tasks[i][1].cpu().numpy().view(np.uint32) for i in idx]
    keys = [tasks[i][2].cpu().numpy().view(np.uint32) for i in idx]
    n = sum(len(x) for x in keys)
    threads = min(CB.physical_cores(), 16)
    best, reps, lists, total = None, 0, None, 0.0
    while reps < 3 or (total < 0.25 and reps < 20):
        lists, secs = CB.probe_tasks(blooms, words, keys, threads, "avx512")
        best = secs if best is None else min(best, secs)
        total += secs
        reps += 1
    identical = all(np.array_equal(l, gpu_lists[i]) for l, i in zip(lists, idx))
    return {"value": round(n / best / 1e9, 4), "unit": "G probes/s", "cores": threads, "kind": "port",
            "isa": "AVX-512" if CB.bbf_simd_available(blooms[0]) & 2 else "scalar",
            "lists_identical_to_gpu": bool(identical),
            "sample": "%d of the %d filters (seeded), %d probe keys: pinned workers probe whole filters, "
                      "best of %d runs = %.4f s" % (len(idx), len(tasks), n, reps, best)}


def run_batched(torch, D, n_filters, steps, seed=11, tw_cycles=TW_CYCLES, model=None, tune_reps=3, CB=None, O=None,
                cpu_sample=512):
    """The batched semi-join: plain per-filter launches of the fixed-config
    kernel vs one autotuned, graph-captured batch plan over all filters (same
    filters, same lists); with a trained cost model, also the learned form:
    each filter's (config, m) chosen by the model for its n and tw, all
    filters in one mixed-configuration plan, beside the exhaustive table's
    choice in the same kind of plan.  Total filter cost = probe time + tw x
    false positives (measured: matches - true hits); tw in GPU cycles (at
    GPU_CLOCK_GHZ) and ns."""
    import numpy as np
    log("batched: generating %d partitions" % n_filters)
    meta, probes = semi_join_inputs(torch, n_filters, seed)
    n_keys = sum(m["n_probe"] for m in meta)
    true_hits = sum(m["hits"] for m in meta)
    wb, k, w, s_, z = FIXED_SEMI_JOIN_CFG
    tasks = build_filters(torch, D, meta, probes, lambda f, nb: (wb, k, w, s_, z, 10 * nb))
    outs = [torch.empty(max(t[2].numel(), 1), dtype=torch.int32, device="cuda") for t in tasks]
    cnts = [torch.zeros(1, dtype=torch.int64, device="cuda") for _ in tasks]
    stream = torch.cuda.current_stream()

    def plain_once():
        for (desc, words, keys, off), o, c in zip(tasks, outs, cnts):
            D.bbf_probe(desc, words, keys, off, out=o, count=c)

    plain_once()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        plain_once()
    b.record(stream)
    torch.cuda.synchronize()
    plain_ms = a.elapsed_time(b) / steps
    plain_counts = [int(c.item()) for c in cnts]
    plain_matches = sum(plain_counts)

    log("batched: autotuning the batch plan")
    plan = D.BatchPlan(tasks)
    best = plan.tune(reps=tune_reps)
    group_shapes = []
    for gi in range(plan.group_count()):
        sh, nt, nk, mfb = plan.group_shape(gi)
        group_shapes.append({"tasks": nt, "probe_keys": nk, "largest_filter_bytes": mfb,
                             "lds_max_bytes": sh.lds_max_bytes, "block_threads": sh.block_threads,
                             "keys_in_flight": "half" if sh.flags & D.PROBE_HALF_IN_FLIGHT else "full",
                             "chunk_tiles": sh.chunk_tiles, "grid_blocks": sh.grid_blocks})
    batch_ms, counts = time_plan_graph(torch, plan, steps)
    plan_lists = plan.results()
    plain_lists = [D.positions(o, c) for o, c in zip(outs, cnts)]
    lists_identical = all(np.array_equal(x, y) for x, y in zip(plan_lists, plain_lists))
    plan.close()
    del outs, cnts, plan_lists
    filter_bytes = int(sum(t[0].size_bytes for t in tasks))
    # Two byte counts per filter f with n keys and c matches.  Block model:
    # 4 B per key, one 64 B block per key where the filter is read from
    # memory or the whole filter once per 2048-key tile where it is staged in
    # LDS (the plan stages filters up to the LDS capacity by default: the
    # launches of its 32 KiB and 158 KiB size classes), 4 B per match.  Compulsory HBM bytes: 4 B per
    # key, min(filter bytes, 64 B per key) since a filter's words come from
    # HBM at most once, 4 B per match.  Filters that fit the L2 or the
    # Infinity Cache serve most block reads from there, so the block model
    # can exceed the HBM peak; the compulsory count is the HBM floor.
    block_bytes = 0
    hbm_bytes = 0
    for (desc, _, keys, _), c in zip(tasks, plain_counts):
        nk = keys.numel()
        staged = desc.size_bytes <= LDS_STAGE_MAX
        tiles = (nk + 2047) // 2048
        block_bytes += 4 * nk + (desc.size_bytes * tiles if staged else 64 * nk) + 4 * c
        hbm_bytes += 4 * nk + min(desc.size_bytes, 64 * nk) + 4 * c
    cpu = None
    if CB is not None:
        log("batched: CPU baseline on %d filters" % cpu_sample)
        cpu = semi_join_cpu_baseline(torch, CB, O, tasks, meta, plain_lists, cpu_sample)
        cpu["full_workload_ms_at_this_rate"] = round(n_keys / cpu["value"] / 1e6, 1)
    del tasks, plain_lists
    torch.cuda.empty_cache()
    fp = plain_matches - true_hits
    res = {"workload": "%d zoned partition filters (u32 words, w=16, z=2, k=6, MAGIC, m = 10 bits per build "
                       "key), build sizes log-uniform 2^10..2^22 keys, probe batches log-uniform 2^12..2^24 "
                       "keys at hit rates uniform in [0.1%%, 30%%]" % n_filters,
           "filters": n_filters, "probe_keys": n_keys, "true_hits": true_hits,
           "fixed_config": "u32 words, w=16, z=2, k=6, m=10 bits/key (MAGIC)",
           "filter_bytes": filter_bytes,
           "plain_ms": round(plain_ms, 3), "batched_graph_ms": round(batch_ms, 3),
           "plain_G_probes_per_s": round(n_keys / plain_ms / 1e6, 2),
           "batched_G_probes_per_s": round(n_keys / batch_ms / 1e6, 2),
           "speedup": round(plain_ms / batch_ms, 3),
           "matches": plain_matches, "false_positives": fp,
           "counts_identical": bool(counts == plain_counts),
           "lists_identical": bool(lists_identical),
           "tuned": {"lds_max_bytes": best.lds_max_bytes, "block_threads": best.block_threads,
                     "chunk_tiles": best.chunk_tiles, "groups": group_shapes},
           "roofline": {"bound": "hbm", "bytes_per_probe": round(hbm_bytes / n_keys, 2),
                        "achieved": round(hbm_bytes / batch_ms / 1e6, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(hbm_bytes / batch_ms / 1e6 / HBM_PEAK_GBS, 4),
                        "note": "compulsory HBM bytes: per filter 4 B per key + min(filter bytes, "
                                "64 B per key) + 4 B per match",
                        "block_model": {"bytes_per_probe": round(block_bytes / n_keys, 2),
                                        "achieved": round(block_bytes / batch_ms / 1e6, 1),
                                        "frac": round(block_bytes / batch_ms / 1e6 / HBM_PEAK_GBS, 4),
                                        "note": "4 B per key + 64 B per key read from memory or the "
                                                "filter once per 2048-key tile when staged in LDS + "
                                                "4 B per match; above 1 where block reads hit L2 or "
                                                "the Infinity Cache"}},
           "total_filter_cost_ms": {}}
    if cpu is not None:
        res["cpu_baseline"] = cpu
    for tc in tw_cycles:
        t = tc / GPU_CLOCK_GHZ
        row = {"tw_cycles": tc, "tw_ns": round(t, 4),
               "plain": round(plain_ms + t * fp * 1e-6, 3), "batched": round(batch_ms + t * fp * 1e-6, 3)}
        if model is not None:
            for leg, use_mlp in (("learned_batched", True), ("exhaustive_batched", False)):
                log("batched: %s configurations for tw=%g cycles" % (leg, tc))
                choices = {}

                def chosen(f, nb):
                    ch = choices.get(nb)
                    if ch is None:
                        ch = choices[nb] = model.choose(nb, t, use_mlp=use_mlp)
                    return ch.word_bits, ch.k, ch.w, ch.s, ch.z, ch.m

                lt = build_filters(torch, D, meta, probes, chosen)
                lplan = D.BatchPlan(lt)
                lplan.tune(reps=tune_reps)
                l_ms, l_counts = time_plan_graph(torch, lplan, steps)
                groups = lplan.group_count()
                lplan.close()
                l_fp = sum(l_counts) - true_hits
                used = {}
                for desc, _, _, _ in lt:
                    key = "u%d w%d s%d z%d k%d" % (desc.word_bits, desc.w, desc.s, desc.z, desc.k)
                    used[key] = used.get(key, 0) + 1
                l_bytes = int(sum(x[0].size_bytes for x in lt))
                del lt
                torch.cuda.empty_cache()
                row[leg] = {"total_ms": round(l_ms + t * l_fp * 1e-6, 3), "probe_ms": round(l_ms, 3),
                            "false_positives": l_fp, "filter_bytes": l_bytes, "config_groups": groups,
                            "top_configs": dict(sorted(used.items(), key=lambda kv: -kv[1])[:4])}
            row["plain_over_learned_batched"] = round(row["plain"] / row["learned_batched"]["total_ms"], 3)
            row["learned_regret_vs_exhaustive"] = round(
                row["learned_batched"]["total_ms"] / row["exhaustive_batched"]["total_ms"] - 1.0, 4)
        res["total_filter_cost_ms"]["tw=%gcycles" % tc] = row
    del probes
    torch.cuda.empty_cache()
    return res


def run_learned(torch, D, train_fraction=1.0, samples=2000000, epochs=64, model_path=None):
    """Calibration sweep (every candidate configuration x filter sizes 2^13..2^33
    bits, timed on the GPU; FPR measured per configuration and bits per key),
    MLP trained on a seeded `train_fraction` of the (configuration, size)
    cells, then the learned choice against the exhaustive optimum over
    n = 2^10..2^28 and tw = 1..1000 GPU cycles per qualifying tuple."""
    import numpy as np
    log("learned: calibration sweep")
    M = D.CostModel.calibrate()
    calib_s = M.calibration_seconds()
    log("learned: calibrated %d configurations in %.1f s; training" % (M.config_count(), calib_s))
    t0 = time.time()
    M.train(train_fraction=train_fraction, samples=samples, epochs=epochs, seed=1)
    train_s = time.time() - t0
    regrets, rows = [], []
    t_ml = t_ex = 0.0
    tws = (1.0, 2.0, 5.0, 10.0, 20.0, 50.0, 100.0, 200.0, 500.0, 1000.0)   # GPU cycles per qualifying tuple
    for ln in range(10, 29, 2):
        for tc in tws:
            tw = tc / GPU_CLOCK_GHZ
            t0 = time.time()
            ml = M.choose(1 << ln, tw, use_mlp=True)
            t1 = time.time()
            ex = M.choose(1 << ln, tw, use_mlp=False)
            t2 = time.time()
            t_ml += t1 - t0
            t_ex += t2 - t1
            r = ml.cost_ns / ex.cost_ns - 1.0
            regrets.append(r)
            if tc in (1.0, 100.0) and ln in (10, 20, 28):
                rows.append({"n": 1 << ln, "tw_cycles": tc, "tw_ns": round(tw, 4), "learned": ml.as_dict(),
                             "exhaustive": ex.as_dict(), "regret": round(r, 4)})
    regrets = np.array(regrets)
    if model_path:
        M.save(model_path)
    res = {"configurations": M.config_count(), "filter_sizes_log2": [13, 33],
           "calibration_s": round(calib_s, 2), "train_s": round(train_s, 2),
           "train_fraction_of_cells": train_fraction, "train_samples": samples,
           "mlp": "8-64-64-1 FP32 (word size, w, s, z, k, log2 m, log2 tw, log2 calibrated FPR at m/n bits per key)",
           "queries": int(regrets.size), "query_grid": "n = 2^10..2^28 (x4) x tw = 1..1000 cycles (10 values)",
           "regret_mean": round(float(regrets.mean()), 4), "regret_median": round(float(np.median(regrets)), 4),
           "regret_p95": round(float(np.percentile(regrets, 95)), 4),
           "regret_max": round(float(regrets.max()), 4),
           "share_within_5pct": round(float((regrets <= 0.05).mean()), 3),
           "choose_ms_learned": round(1e3 * t_ml / regrets.size, 3),
           "choose_ms_exhaustive_table": round(1e3 * t_ex / regrets.size, 3),
           "examples": rows}
    return res, M


def cuckoo_build_keys(n_build):
    """n_build distinct uint32 keys below 2^31 for the cuckoo workload (numpy;
    the same seeded bijection as distinct_keys, seed 9)."""
    import numpy as np
    M = (1 << 31) - 1
    x = (np.arange(n_build, dtype=np.uint64) + np.uint64(9 << 20)) & np.uint64(M)
    x = (x * np.uint64(0x5851F42D)) & np.uint64(M)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x2C1B3C6D)) & np.uint64(M)
    x ^= x >> np.uint64(13)
    x = (x * np.uint64(0x297A2D39)) & np.uint64(M)
    return x.astype(np.uint32)


def cuckoo_probe_keys(torch, build_dev, n, hit):
    """n probe keys on the GPU, distinct and absent (bit 31 set), with
    round(hit x n) of them replaced at random positions by build keys;
    returns (keys, planted hits)."""
    g = torch.Generator(device="cuda").manual_seed(5)
    probe = distinct_keys(torch, n, 77, 1)
    hits = int(round(hit * n))
    idx = torch.randperm(n, device="cuda", generator=g)[:hits]
    probe[idx] = build_dev[torch.randint(0, build_dev.numel(), (hits,), device="cuda", generator=g)]
    torch.cuda.synchronize()
    return probe, hits


def run_cuckoo(torch, D, CB, O, steps, warmup, no_cpu, log2_buckets=26, load=0.95):
    """Cuckoo filter, 16-bit tags, 4 slots per bucket, 2^26 buckets (512 MiB +
    victim cache), loaded to 95% with distinct uniform keys by the host build
    (cf_build_host, the reference's insert with a seeded mt19937 for kicks);
    GPU probe of 2^28 keys at 1% and 50% hits; the GPU build (cf_build) of
    the same keys timed beside it."""
    import numpy as np
    desc = D.cf_desc((1 << log2_buckets) * 64, 16, 4)
    n_build = int(load * (1 << log2_buckets) * 4)
    log("cuckoo: generating %d distinct keys" % n_build)
    keys = cuckoo_build_keys(n_build)
    log("cuckoo: host build")
    t0 = time.time()
    buf, inserted = D.cf_build_host(desc, keys, seed=1)
    build_s = time.time() - t0
    dbuf = torch.from_numpy(buf).cuda()
    build_dev = torch.from_numpy(keys.view(np.int32)).cuda()
    del keys
    n = 1 << 28
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    res = {"workload": "cuckoo filter batch_contains: 16-bit tags, 4 slots/bucket, 2^%d buckets (%d MiB), "
                       "%d keys (%.0f%% load) built on the host; 2^28 probe keys" %
                       (log2_buckets, desc.buffer_bytes >> 20, n_build, load * 100),
           "host_build_s": round(build_s, 1), "inserted": int(inserted), "legs": []}
    log("cuckoo: GPU build of the same keys (cf_build)")
    gbuf = D.new_cf_table(desc)
    failed = D.cf_build(desc, gbuf, build_dev, seed=1)    # untimed first run (loads the kernel)
    torch.cuda.synchronize()
    gbuf.zero_()
    failed.zero_()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    D.cf_build(desc, gbuf, build_dev, seed=1, failed=failed)
    b.record()
    torch.cuda.synchronize()
    g_ms = a.elapsed_time(b)
    found_out = torch.empty(n_build, dtype=torch.int32, device="cuda")
    found_cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    D.cf_probe(desc, gbuf, build_dev, 0, out=found_out, count=found_cnt)
    torch.cuda.synchronize()
    res["gpu_build"] = {"ms": round(g_ms, 2), "M_keys_per_s": round(n_build / g_ms / 1e3, 1),
                        "failed": int(failed.item()), "build_keys_found": int(found_cnt.item()),
                        "vs_host_build": round(build_s * 1e3 / g_ms, 1)}
    del gbuf, found_out
    cf = O.Cuckoo(desc.m, 16, 4) if not no_cpu else None
    for hit in (0.01, 0.5):
        log("cuckoo: probing at %.0f%% hits" % (hit * 100))
        probe, hits = cuckoo_probe_keys(torch, build_dev, n, hit)

        def timed(params):
            for _ in range(warmup):
                D.cf_probe(desc, dbuf, probe, 0, out=out, count=cnt, params=params)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(steps):
                D.cf_probe(desc, dbuf, probe, 0, out=out, count=cnt, params=params)
            b.record()
            torch.cuda.synchronize()
            c = int(cnt.item())
            return a.elapsed_time(b) / steps, c, out[:c].cpu().numpy().view(np.uint32).astype(np.uint64)

        ms, matches, pos = timed(None)   # default: partitioned two-pass probe for this size
        d_ms, d_matches, d_pos = timed(D.ProbeParams(0, 0, 0, 0, D.PROBE_DIRECT))
        bpk = 4.0 + 2 * 64.0 + 4.0 * matches / n
        leg = {"hit_rate": hit, "ms": round(ms, 4), "G_probes_per_s": round(n / ms / 1e6, 2),
               "matches": matches, "false_positives": matches - hits,
               "path": "partitioned (keys grouped by first-bucket slice, then by alternate-bucket slice)",
               "roofline": {"bound": "hbm", "bytes_per_probe": round(bpk, 2),
                            "achieved": round(bpk * n / ms / 1e6, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(bpk * n / ms / 1e6 / HBM_PEAK_GBS, 4),
                            "note": "4 B key + two 64 B bucket sectors + 4 B per match"},
               "direct_kernel_call": {"ms": round(d_ms, 4), "G_probes_per_s": round(n / d_ms / 1e6, 2),
                                      "frac": round(bpk * n / d_ms / 1e6 / HBM_PEAK_GBS, 4),
                                      "lists_identical": bool(np.array_equal(d_pos, pos))}}
        del d_pos
        if not no_cpu:
            leg["cpu_baseline"] = cpu_baseline(CB, "cf", cf, buf, probe.cpu().numpy().view(np.uint32), pos,
                                               "the same 512 MiB table")
        res["legs"].append(leg)
        del probe, pos
    del dbuf, build_dev, out
    torch.cuda.empty_cache()
    return res


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def summarize_traffic(fetch_csv, write_csv, commit):
    """HBM bytes of the headline call from two `rocprofv3 --pmc` passes (FETCH_SIZE,
    then WRITE_SIZE, KiB per dispatch) of `bench.py --headline-only --calib-copy`:
    each counter is scaled by true bytes / counter bytes of the 1 GiB device copy
    that --calib-copy runs first; the phase is the last partitioned call
    (k_part_init_bins or, with a histogram pass, k_fill_zero_u32 .. k_hit_bits),
    the direct kernel its last launch."""
    import csv

    def rows(path, counter):
        out = []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                out.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"]) * 1024.0))
        return sorted(out)

    f, w = rows(fetch_csv, "FETCH_SIZE"), rows(write_csv, "WRITE_SIZE")
    if [x[1] for x in f] != [x[1] for x in w]:
        raise SystemExit("the two passes dispatched different kernels")
    copy_bytes = 1 << 30
    ci = next(i for i, x in enumerate(f) if x[1].startswith("__amd_rocclr_copyBuffer"))
    fs, ws = copy_bytes / f[ci][2], copy_bytes / w[ci][2]

    def short(name):
        name = name.split("(")[0].replace("void ", "").replace("dtlh::", "")
        return name.replace("unsigned long", "u64")

    scat = [i for i, x in enumerate(f) if "k_part_scatter<" in x[1]][-1]
    first = max(i for i in range(scat) if "k_fill_zero_u32" in f[i][1] or "k_part_init_bins" in f[i][1])
    last = next(i for i in range(scat, len(f)) if "k_hit_bits" in f[i][1])
    kern = [{"kernel": short(f[i][1]), "fetch_bytes": round(f[i][2] * fs), "write_bytes": round(w[i][2] * ws)}
            for i in range(first, last + 1)]
    fb, wb = sum(k["fetch_bytes"] for k in kern), sum(k["write_bytes"] for k in kern)
    d = [i for i, x in enumerate(f) if "k_bbf_probe_global<" in x[1]][-1]
    fdir, wdir = os.path.dirname(fetch_csv) or ".", os.path.dirname(write_csv) or "."
    return {
        "workload": "cache-sectorized BBF u64 w8 s8 k8, 1 GiB filter, 2^28 probe keys at 1% hits (bench.py defaults)",
        "commit": commit,
        "commands": ["rocprofv3 --pmc FETCH_SIZE -d %s -o %s --output-format csv -- python3 bench.py --steps 2 "
                     "--warmup 1 --no-cpu --headline-only --calib-copy"
                     % (fdir, os.path.basename(fetch_csv).split("_counter_collection")[0]),
                     "rocprofv3 --pmc WRITE_SIZE -d %s -o %s --output-format csv -- python3 bench.py --steps 2 "
                     "--warmup 1 --no-cpu --headline-only --calib-copy"
                     % (wdir, os.path.basename(write_csv).split("_counter_collection")[0]),
                     "python3 bench.py --summarize-traffic %s %s --commit %s" % (fetch_csv, write_csv, commit)],
        "copy_calibration": {"copy_bytes": copy_bytes, "fetch_kib": f[ci][2] / 1024, "write_kib": w[ci][2] / 1024,
                             "fetch_scale": round(fs, 4), "write_scale": round(ws, 4)},
        "partitioned_phase": {"kernels": kern, "fetch_bytes": fb, "write_bytes": wb},
        "bytes_per_launch": fb + wb,
        "direct_kernel": {"kernel": short(f[d][1]), "fetch_bytes": round(f[d][2] * fs),
                          "write_bytes": round(w[d][2] * ws),
                          "bytes_per_launch": round(f[d][2] * fs) + round(w[d][2] * ws)},
        "algorithmic_bytes_per_launch": (4.0 + 64.0) * (1 << 28) + 4.0 * 2763103,
    }


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize-traffic":
        sp = argparse.ArgumentParser()
        sp.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
        sp.add_argument("--commit", default=None)
        a = sp.parse_args()
        print(json.dumps(summarize_traffic(a.summarize_traffic[0], a.summarize_traffic[1], a.commit), indent=1))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-m", type=int, default=33)
    ap.add_argument("--log2-probe", type=int, default=28)
    ap.add_argument("--hit-rate", type=float, default=0.01)
    ap.add_argument("--no-cpu", action="store_true", help="skip every CPU baseline")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline (no regimes, cuckoo, learned model or semi-join)")
    ap.add_argument("--no-regimes", action="store_true", help="skip the filter-size sweep")
    ap.add_argument("--no-cuckoo", action="store_true", help="skip the cuckoo workload")
    ap.add_argument("--no-semi-join", action="store_true", help="skip the learned model and the semi-join")
    ap.add_argument("--semi-join-filters", type=int, default=4096)
    ap.add_argument("--gather-only", action="store_true",
                    help="only run the random-sector gather microbenchmark (all access patterns)")
    ap.add_argument("--regimes-only", action="store_true", help="only run the filter-size sweep")
    ap.add_argument("--batched", type=int, default=0, metavar="FILTERS",
                    help="only run the batched semi-join workload with this many filters (fixed configuration)")
    ap.add_argument("--cuckoo", action="store_true", help="only run the cuckoo filter workload")
    ap.add_argument("--learned", action="store_true",
                    help="only run the learned configuration model; with --batched also the learned + batched "
                         "semi-join")
    ap.add_argument("--train-fraction", type=float, default=1.0)
    ap.add_argument("--model-path", default=None, help="save the calibrated + trained model here")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (a single MI355X); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import dtl_hip as D
    D.load()
    torch.cuda.set_device(0)
    CB = O = None
    if not args.no_cpu:
        import cpu_baseline as CB
        import oracle as O

    if args.cuckoo:
        print(json.dumps({"cuckoo": run_cuckoo(torch, D, CB, O, args.steps, args.warmup, args.no_cpu)}), flush=True)
        return
    if args.learned:
        res, model = run_learned(torch, D, train_fraction=args.train_fraction, model_path=args.model_path)
        print(json.dumps({"learned_config": res}), flush=True)
        if args.batched:
            b = run_batched(torch, D, args.batched, max(2, args.steps // 4), model=model, tune_reps=1, CB=CB, O=O)
            print(json.dumps({"batched_semi_join": b}), flush=True)
        model.close()
        return
    if args.batched:
        print(json.dumps({"batched_semi_join": run_batched(torch, D, args.batched, max(2, args.steps // 4),
                                                           CB=CB, O=O)}), flush=True)
        return
    if args.regimes_only:
        print(json.dumps({"regimes": run_regimes(torch, D, CB, O, args.steps, args.no_cpu)}), flush=True)
        return
    if args.gather_only:
        res = {"random_sector_gather_1GiB": gather_table(torch, D, range(8))}
        for lb in (21, 22, 24, 27, 30):
            res["by_size_%dMiB" % (1 << (lb - 20))] = {
                m: r["G_sectors_per_s"] for m, r in gather_table(torch, D, (0, 2, 6, 7), reps=3, log2_bytes=lb).items()}
        for lb in (14, 17):
            res["lds_%dKiB" % (1 << (lb - 10))] = gather_table(torch, D, (8, 9), reps=3, log2_bytes=lb)
        print(json.dumps(res), flush=True)
        return

    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log("bench: building inputs (m=2^%d bits, 2^%d probe keys)" % (args.log2_m, args.log2_probe))
    desc = D.bbf_desc(64, 1 << args.log2_m, 8, 8, 8, 1)
    build, probe, n_build = make_inputs(torch, args.log2_m, 1 << args.log2_probe, args.hit_rate, 1)
    words = D.new_filter(desc)
    D.bbf_build(desc, words, build)
    torch.cuda.synchronize()
    del build

    log("bench: probing (%d warmup + %d timed steps)" % (args.warmup, args.steps))
    per_call = {}
    ms, kern_ms, matches, pos = time_probe(torch, D, desc, words, probe, args.steps, args.warmup, per_step=per_call)
    n = probe.numel()
    value = n / (ms * 1e-3) / 1e9
    partitioned = desc.size_bytes >= (64 << 20) and n >= (1 << 22)
    # algorithmic bytes per probe: 4 B key + one 64 B block + 4 B per match
    bytes_per_key = 4.0 + 64.0 + 4.0 * matches / n
    achieved = bytes_per_key * n / (kern_ms * 1e-3) / 1e9
    traffic = load_traffic("traffic_bbf_probe_u64_w8_s8_k8_1gib.json")
    gpu_pos = pos.cpu().numpy().view(np.uint32).astype(np.uint64)
    del pos
    log("bench: direct kernel leg")
    dparams = D.ProbeParams(0, 0, 0, 0, D.PROBE_DIRECT)
    d_ms, d_kern_ms, d_matches, d_pos = time_probe(torch, D, desc, words, probe, max(3, args.steps // 2), 1,
                                                   params=dparams)
    d_identical = bool(np.array_equal(d_pos.cpu().numpy().view(np.uint32).astype(np.uint64), gpu_pos))
    del d_pos
    d_achieved = bytes_per_key * n / (d_kern_ms * 1e-3) / 1e9

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "G probes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 keys, u64 filter words",
        "data": "synthetic (seeded uniform uint32 keys)",
        "config": {
            "workload": "cache-sectorized blocked Bloom filter batch_contains: 64-bit words, w=8, s=8, "
                        "k=8, m=2^%d bits, built from m/12 keys; 2^%d probe keys at %.1f%% hits"
                        % (args.log2_m, args.log2_probe, args.hit_rate * 100),
            "filter_bytes": int(desc.size_bytes),
            "probe_keys": n,
            "matches": matches,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("bytes_per_launch") if traffic else None,
            "traffic_measured_at": traffic.get("commit") if traffic else None,
            "kernel": ("partitioned probe phase (k_part_init_bins, k_part_scatter into fixed-capacity "
                       "bins, k_bbf_probe_split over the grouped keys: 4 lanes per key, one 16 B piece "
                       "of the 64 B block each; match reorder, whose hit sort writes the ascending positions), HIP "
                       "events around the phase" if partitioned
                       else "k_bbf_probe_global<u64,w8,s8,k8>"),
            "kernel_ms": round(kern_ms, 4),
            "kernel_ms_per_call": per_call,
            "bytes_per_probe": round(bytes_per_key, 4),
            "direct_kernel": {
                "kernel": "k_bbf_probe_global<u64,w8,s8,k8> (one random 64 B block per key from HBM)",
                "kernel_ms": round(d_kern_ms, 4), "achieved": round(d_achieved, 1),
                "frac": round(d_achieved / HBM_PEAK_GBS, 4),
                "G_probes_per_s": round(n / (d_ms * 1e-3) / 1e9, 3),
                "lists_identical": d_identical},
        },
    }
    if not args.no_cpu:
        log("bench: CPU baseline on the whole batch (%d keys)" % n)
        f = O.Bloom(64, 1 << args.log2_m, 8, 8, 8, 1)
        result["cpu_baseline"] = cpu_baseline(CB, "bbf", f, words.cpu().numpy().view(np.uint64),
                                              probe.cpu().numpy().view(np.uint32), gpu_pos,
                                              "the same 1 GiB filter", one_thread=True)
    del probe, words, gpu_pos
    torch.cuda.empty_cache()
    g = gather_table(torch, D, [0, 2])
    best = max(g.values(), key=lambda r: r["G_sectors_per_s"])
    result["roofline"]["measured_random_64B_gather"] = {
        "G_sectors_per_s": best["G_sectors_per_s"], "pattern": best["pattern"],
        "direct_kernel_vs_gather": round(n / (d_kern_ms * 1e-3) / 1e9 / best["G_sectors_per_s"], 3),
        "probe_phase_vs_gather": round(n / (kern_ms * 1e-3) / 1e9 / best["G_sectors_per_s"], 3),
        "note": "gfx950 fills a whole 128 B L2 line per random 64 B block (TCC_EA0_RDREQ_128B = all reads)"}
    if args.headline_only:
        print(json.dumps(result), flush=True)
        return
    if not args.no_regimes:
        result["regimes"] = run_regimes(torch, D, CB, O, args.steps, args.no_cpu)
    if not args.no_cuckoo:
        result["cuckoo"] = run_cuckoo(torch, D, CB, O, max(3, args.steps // 2), 2, args.no_cpu)
    if not args.no_semi_join:
        lres, model = run_learned(torch, D, train_fraction=args.train_fraction, model_path=args.model_path)
        result["learned_config"] = lres
        result["batched_semi_join"] = run_batched(torch, D, args.semi_join_filters, max(2, args.steps // 4),
                                                  model=model, tune_reps=1, CB=CB, O=O)
        model.close()
        result["tuned_vs_plain"] = {
            k: {"plain_ms": v["plain"], "learned_batched_ms": v["learned_batched"]["total_ms"],
                "speedup": v["plain_over_learned_batched"],
                "learned_regret_vs_exhaustive": v["learned_regret_vs_exhaustive"]}
            for k, v in result["batched_semi_join"]["total_filter_cost_ms"].items()}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
