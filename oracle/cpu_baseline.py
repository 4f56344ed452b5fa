"""ctypes binding of the multithreaded AVX-512 / AVX2 CPU baseline (cpu_simd.cpp).

BENCH/TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it; tests/
check it against the scalar oracle.
"""
import ctypes as C
import os

import numpy as np

import oracle as O

_lib = None


ISAS = {"best": -1, "scalar": 0, "avx2": 1, "avx512": 2}
ISA_NAMES = {0: "scalar", 1: "AVX2", 2: "AVX-512"}


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libcpubase.so")
        if not os.path.exists(path):
            raise RuntimeError("CPU baseline not built (%s)" % path)
        L = C.CDLL(path)
        L.cpub_bbf_simd_available.restype = C.c_int
        L.cpub_bbf_simd_available.argtypes = [C.POINTER(O.BBF)]
        L.cpub_cf_simd_available.restype = C.c_int
        L.cpub_cf_simd_available.argtypes = [C.POINTER(O.CF)]
        L.cpub_physical_cores.restype = C.c_int
        L.cpub_physical_cores.argtypes = []
        L.cpub_set_pinning.restype = None
        L.cpub_set_pinning.argtypes = [C.c_int]
        for fn, desc in (("cpub_bbf_probe", O.BBF), ("cpub_cf_probe", O.CF)):
            getattr(L, fn).restype = C.c_uint64
            getattr(L, fn).argtypes = [C.POINTER(desc), C.c_void_p, C.c_void_p, C.c_uint64, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.cpub_bbf_probe_tasks.restype = C.c_uint64
        L.cpub_bbf_probe_tasks.argtypes = [C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_uint64), C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.cpub_bbf_build.restype = None
        L.cpub_bbf_build.argtypes = [C.POINTER(O.BBF), C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
        _lib = L
    return _lib


def physical_cores():
    """Physical cores of this process's CPU set (SMT siblings counted once);
    the probe workers are pinned one per core."""
    return max(1, lib().cpub_physical_cores())


class Probe:
    """Result of one threaded probe: total count, wall seconds of the parallel
    region, per-thread position lists (concatenated they are the whole
    list), and the body that ran ("AVX-512", "AVX2" or "scalar")."""

    def __init__(self, count, seconds, lists, isa):
        self.count, self.seconds, self.lists, self.isa = count, seconds, lists, isa

    def positions(self, offset=0):
        out = np.concatenate(self.lists).astype(np.uint64) if self.lists else np.zeros(0, np.uint64)
        return out + offset


def _isa_code(simd):
    if simd is True:
        return -1
    if simd is False:
        return 0
    return ISAS[simd]


def run(kind, obj, data, keys, threads, simd=True):
    """Threaded probe of `keys` against a Bloom filter (kind "bbf", data = its
    words) or a cuckoo table (kind "cf", data = its buffer).  simd: True (the
    best body for this configuration and CPU), False (scalar), or one of
    "avx512", "avx2", "scalar"."""
    fn = lib().cpub_bbf_probe if kind == "bbf" else lib().cpub_cf_probe
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    per = (len(keys) + threads - 1) // threads
    outs = [np.zeros(max(per, 1), dtype=np.uint32) for _ in range(threads)]
    ptrs = (C.c_void_p * threads)(*[o.ctypes.data for o in outs])
    counts = (C.c_uint64 * threads)()
    secs = C.c_double(0)
    used = C.c_int(0)
    total = fn(C.byref(obj.d), data.ctypes.data, keys.ctypes.data, len(keys), threads, _isa_code(simd), ptrs,
               counts, C.byref(secs), C.byref(used))
    lists = [o[: counts[t]] for t, o in enumerate(outs)]
    return Probe(total, secs.value, lists, ISA_NAMES[used.value])


def bbf_probe(bloom, words, keys, threads, simd=True):
    """Returns (count, seconds, per-thread position lists)."""
    r = run("bbf", bloom, words, keys, threads, simd)
    return r.count, r.seconds, r.lists


def cf_probe(cuckoo, buf, keys, threads, simd=True):
    r = run("cf", cuckoo, buf, keys, threads, simd)
    return r.count, r.seconds, r.lists


def bbf_simd_available(bloom):
    """Bit 1: an AVX2 body covers the configuration on this CPU; bit 2: AVX-512."""
    return lib().cpub_bbf_simd_available(C.byref(bloom.d))


def cf_simd_available(cuckoo):
    return lib().cpub_cf_simd_available(C.byref(cuckoo.d))


def bbf_build(bloom, keys, threads):
    """The oracle's filter for `keys`, built on `threads` host threads (block
    ranges per thread; equals bloom.insert on a zeroed filter)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    words = bloom.new_words()
    lib().cpub_bbf_build(C.byref(bloom.d), words.ctypes.data, keys.ctypes.data, len(keys), threads)
    return words


def probe_tasks(blooms, words, keys, threads, simd=True):
    """Probes many filters, each with its own keys, on `threads` pinned
    workers that take whole tasks in the given order (largest first balances
    best).  Returns (per-task position lists, wall seconds)."""
    n = len(blooms)
    keys = [np.ascontiguousarray(k, dtype=np.uint32) for k in keys]
    outs = [np.zeros(max(len(k), 1), dtype=np.uint32) for k in keys]
    arr = lambda xs: (C.c_void_p * n)(*xs)
    counts = (C.c_uint64 * n)()
    secs = C.c_double(0)
    lib().cpub_bbf_probe_tasks(n, arr([C.addressof(b.d) for b in blooms]), arr([w.ctypes.data for w in words]),
                               arr([k.ctypes.data for k in keys]), (C.c_uint64 * n)(*[len(k) for k in keys]),
                               arr([o.ctypes.data for o in outs]), counts, threads, _isa_code(simd),
                               C.byref(secs))
    return [o[: counts[i]] for i, o in enumerate(outs)], secs.value
