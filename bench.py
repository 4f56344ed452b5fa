#!/usr/bin/env python3
"""Benchmark of the filter probe path on one MI355X.

Headline workload (one "step" = one bbf_probe call over one batch):
  cache-sectorized blocked Bloom filter, 64-bit words, w=8 words (512-bit
  blocks), s=8 sectors, k=8, m = 2^33 bits (1 GiB), built from m/12 uniform
  uint32 keys; probe batch of 2^28 uint32 keys at a 1% hit rate; output is the
  ascending match-position list.  Inputs are resident in HBM before timing.

Prints one JSON line (see the `result` dict) with the whole-job probe rate,
the roofline of the dominant kernel (the probe kernel, timed with HIP events
on the stream it runs on) and a CPU baseline measured on this host.
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
METRIC = "G probes/s on 1 MI355X + % random-access roofline; tuned vs plain config"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(torch, log2_m, n_probe, hit_rate, seed):
    """Seeded synthetic inputs on the GPU: build keys uniform in [0, 2^31),
    probe keys uniform in [2^31, 2^32) with round(hit_rate * n) of them
    replaced by build keys at random positions."""
    n_build = (1 << log2_m) // 12
    g = torch.Generator(device="cuda").manual_seed(seed)
    build = torch.randint(0, 2**31, (n_build,), dtype=torch.int64, device="cuda", generator=g).to(torch.int32)
    probe = torch.randint(-(2**31), 0, (n_probe,), dtype=torch.int64, device="cuda", generator=g).to(torch.int32)
    hits = int(round(hit_rate * n_probe))
    idx = torch.randperm(n_probe, device="cuda", generator=g)[:hits]
    probe[idx] = build[torch.randint(0, n_build, (hits,), device="cuda", generator=g)]
    return build, probe, n_build


def time_probe(torch, D, desc, words, probe, steps, warmup, params=None):
    """Returns (ms per step, avg ms of the probe kernel, matches)."""
    n = probe.numel()
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
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for i in range(steps):
        p = D.ProbeParams() if params is None else D.ProbeParams(
            params.lds_max_bytes, params.block_threads, params.grid_blocks, params.chunk_tiles, params.flags)
        p.ev_start = evs[i][0].cuda_event
        p.ev_stop = evs[i][1].cuda_event
        D.bbf_probe(desc, words, probe, 0, params=p, out=out, count=cnt)
    t1.record(stream)
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / steps
    return total_ms / steps, kern_ms, int(cnt.item())


def cpu_baseline(torch, O, CB, words, probe, log2_m, sample_keys):
    """The multithreaded AVX-512 restatement of the reference's batch_contains
    (oracle/cpu_simd.cpp) on the first `sample_keys` probe keys against the
    same 1 GiB filter copied to host memory: one thread per physical core of
    this process's CPU share (at most 16), each probing a contiguous slice in
    1024-key batches into its own output (benchmark.cpp:123-164); best of 3."""
    import numpy as np
    f = O.Bloom(64, 1 << log2_m, 8, 8, 8, 1)
    host_words = words.cpu().numpy().view(np.uint64)
    keys = probe[:sample_keys].cpu().numpy().view(np.uint32).copy()
    threads = min(CB.physical_cores(), 16)
    simd = CB.bbf_simd_available(f)
    best = None
    for _ in range(3):
        _, secs, _ = CB.bbf_probe(f, host_words, keys, threads, simd)
        best = secs if best is None else min(best, secs)
    one = keys[: max(sample_keys // 16, 1 << 20)]
    _, secs1, _ = CB.bbf_probe(f, host_words, one, 1, simd)
    return {"value": round(sample_keys / best / 1e9, 4), "unit": "G probes/s", "cores": threads,
            "kind": "port",
            "sample": "%d probe keys (1%% hits) against the same 1 GiB filter in host memory; %s "
                      "batch_contains, %d threads, best of 3 = %.3f s; 1 thread: %.4f G probes/s"
                      % (sample_keys, "AVX-512" if simd else "scalar", threads, best,
                         len(one) / secs1 / 1e9)}


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2-m", type=int, default=33)
    ap.add_argument("--log2-probe", type=int, default=28)
    ap.add_argument("--hit-rate", type=float, default=0.01)
    ap.add_argument("--cpu-sample", type=int, default=1 << 26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every GPU box in this setup has a single MI355X, "
              "so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    import dtl_hip as D
    D.load()
    torch.cuda.set_device(0)

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
    ms, kern_ms, matches = time_probe(torch, D, desc, words, probe, args.steps, args.warmup)
    n = probe.numel()
    value = n / (ms * 1e-3) / 1e9
    bytes_per_key = 4.0 + 64.0 + 1.0 / 8.0   # key + one 64 B block + match bitmap bit
    achieved = bytes_per_key * n / (kern_ms * 1e-3) / 1e9
    traffic = load_traffic("traffic_bbf_probe_u64_w8_s8_k8_1gib.json")

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
            "kernel": "k_bbf_probe_global<u64,w8,s8,k8>",
            "kernel_ms": round(kern_ms, 4),
            "bytes_per_probe": bytes_per_key,
        },
    }
    if not args.no_cpu:
        import cpu_baseline as CB
        import oracle as O
        log("bench: CPU baseline on %d keys" % args.cpu_sample)
        result["cpu_baseline"] = cpu_baseline(torch, O, CB, words, probe, args.log2_m, args.cpu_sample)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
