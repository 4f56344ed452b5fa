"""ctypes binding of the multithreaded AVX-512 CPU baseline (cpu_simd.cpp).

BENCH/TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it; tests/
check it against the scalar oracle.
"""
import ctypes as C
import os

import numpy as np

import oracle as O

_lib = None


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
        for fn, desc in (("cpub_bbf_probe", O.BBF), ("cpub_cf_probe", O.CF)):
            getattr(L, fn).restype = C.c_uint64
            getattr(L, fn).argtypes = [C.POINTER(desc), C.c_void_p, C.c_void_p, C.c_uint64, C.c_int,
                                       C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_double)]
        L.cpub_bbf_build.restype = None
        L.cpub_bbf_build.argtypes = [C.POINTER(O.BBF), C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]
        _lib = L
    return _lib


def physical_cores():
    """Physical cores usable by this process (SMT siblings counted once)."""
    cpus = sorted(os.sched_getaffinity(0))
    seen = set()
    for c in cpus:
        path = "/sys/devices/system/cpu/cpu%d/topology/thread_siblings_list" % c
        try:
            with open(path) as fh:
                seen.add(fh.read().strip())
        except OSError:
            seen.add(str(c))
    return max(1, len(seen))


def _probe(fn, d, buf, keys, threads, simd):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    per = (len(keys) + threads - 1) // threads
    outs = [np.zeros(max(per, 1), dtype=np.uint32) for _ in range(threads)]
    ptrs = (C.c_void_p * threads)(*[o.ctypes.data for o in outs])
    counts = (C.c_uint64 * threads)()
    secs = C.c_double(0)
    total = fn(C.byref(d), buf.ctypes.data, keys.ctypes.data, len(keys), threads, int(simd), ptrs,
               counts, C.byref(secs))
    lists = [o[: counts[t]] for t, o in enumerate(outs)]
    return total, secs.value, lists


def bbf_probe(bloom, words, keys, threads, simd=True):
    """Returns (count, seconds, per-thread position lists)."""
    return _probe(lib().cpub_bbf_probe, bloom.d, words, keys, threads, simd)


def cf_probe(cuckoo, buf, keys, threads, simd=True):
    return _probe(lib().cpub_cf_probe, cuckoo.d, buf, keys, threads, simd)


def bbf_simd_available(bloom):
    return bool(lib().cpub_bbf_simd_available(C.byref(bloom.d)))


def cf_simd_available(cuckoo):
    return bool(lib().cpub_cf_simd_available(C.byref(cuckoo.d)))


def bbf_build(bloom, keys, threads):
    """The oracle's filter for `keys`, built on `threads` host threads (block
    ranges per thread; equals bloom.insert on a zeroed filter)."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    words = bloom.new_words()
    lib().cpub_bbf_build(C.byref(bloom.d), words.ctypes.data, keys.ctypes.data, len(keys), threads)
    return words
