"""ctypes binding of libdtl_hip (include/dtl_hip.h) for tests and benchmarks.

Device memory and streams come from PyTorch (plumbing only); every filter
operation runs in the library's gfx950 kernels.  Loading fails loudly when the
library has not been built, and every call raises on a non-OK status: there
is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdtl_hip.so")

OK, EINVAL, ERUNTIME, ELOGIC, ENOMEM, EHIP, ENODEVICE = 0, -1, -2, -3, -4, -5, -6
ADDR_POW2, ADDR_MAGIC = 0, 1
PROBE_FORCE_GLOBAL = 1
PROBE_PARTITIONED = 2
PROBE_DIRECT = 4
PROBE_LANE_SPLIT = 8
PROBE_ONE_LANE = 16
PROBE_PART_HISTOGRAM = 32
PROBE_HALF_IN_FLIGHT = 64
PROBE_CF_PASS2_ALL = 128


class BbfDesc(C.Structure):
    _fields_ = [("word_bits", C.c_uint32), ("w", C.c_uint32), ("s", C.c_uint32), ("z", C.c_uint32),
                ("k", C.c_uint32), ("addr_mode", C.c_uint32), ("block_cnt", C.c_uint32),
                ("block_shift", C.c_uint32), ("magic", C.c_uint32), ("magic_shift", C.c_uint32),
                ("m", C.c_uint64), ("m_actual", C.c_uint64), ("word_cnt", C.c_uint64),
                ("size_bytes", C.c_uint64)]


class CfDesc(C.Structure):
    _fields_ = [("tag_bits", C.c_uint32), ("tags_per_bucket", C.c_uint32),
                ("bytes_per_bucket", C.c_uint32), ("addr_mode", C.c_uint32),
                ("bucket_cnt", C.c_uint32), ("bucket_shift", C.c_uint32),
                ("bucket_mask", C.c_uint32), ("magic", C.c_uint32), ("magic_shift", C.c_uint32),
                ("reserved", C.c_uint32), ("m", C.c_uint64), ("m_actual", C.c_uint64),
                ("table_words", C.c_uint64), ("victim_offset", C.c_uint64),
                ("buffer_bytes", C.c_uint64)]


class ProbeParams(C.Structure):
    _fields_ = [("lds_max_bytes", C.c_uint32), ("block_threads", C.c_uint32),
                ("grid_blocks", C.c_uint32), ("chunk_tiles", C.c_uint32), ("flags", C.c_uint32),
                ("ev_start", C.c_void_p), ("ev_stop", C.c_void_p), ("slice_log2", C.c_uint32),
                ("reserved", C.c_uint32)]


class BbfTask(C.Structure):
    _fields_ = [("desc", BbfDesc), ("d_words", C.c_void_p), ("d_keys", C.c_void_p),
                ("n", C.c_uint64), ("match_offset", C.c_uint32), ("reserved", C.c_uint32),
                ("d_pos", C.c_void_p), ("d_count", C.c_void_p)]


class CalibParams(C.Structure):
    _fields_ = [("log2_m_min", C.c_uint32), ("log2_m_max", C.c_uint32), ("probe_keys_log2", C.c_uint32),
                ("reps", C.c_uint32), ("fpr_keys_log2", C.c_uint32), ("flags", C.c_uint32)]


class BbfChoice(C.Structure):
    _fields_ = [("word_bits", C.c_uint32), ("w", C.c_uint32), ("s", C.c_uint32), ("z", C.c_uint32),
                ("k", C.c_uint32), ("reserved", C.c_uint32), ("m", C.c_uint64), ("cost_ns", C.c_double),
                ("predicted", C.c_double)]

    def key(self):
        return (self.word_bits, self.w, self.s, self.z, self.k)

    def as_dict(self):
        return {"word_bits": self.word_bits, "w": self.w, "s": self.s, "z": self.z, "k": self.k,
                "m": self.m, "cost_ns": round(self.cost_ns, 5)}


class DtlError(RuntimeError):
    def __init__(self, fn, status):
        detail = load().dtl_last_error().decode() if status == EHIP else ""
        super().__init__("%s failed: %s (%d) %s" % (fn, status_string(status), status, detail))
        self.status = status


EXPORTS = [
    "dtl_bbf_desc_init", "dtl_cf_desc_init", "bbf_build", "bbf_probe", "bbf_probe_ex", "cf_probe", "cf_probe_ex",
    "cf_build_host", "cf_build", "bbf_batch_plan_create", "bbf_batch_plan_run", "bbf_batch_plan_destroy",
    "bbf_probe_batched", "bbf_batch_plan_tune", "dtl_status_string", "dtl_last_error", "dtl_bbf_supported",
    "dtl_bbf_config_count", "dtl_bbf_split_lanes", "dtl_version", "dtl_gather_bench",
    "bbf_batch_plan_group_count", "bbf_batch_plan_group_shape", "dtl_model_calibrate", "dtl_model_train", "dtl_model_choose", "dtl_model_cost", "dtl_model_lookup_ns",
    "dtl_model_fpr", "dtl_model_config_count", "dtl_model_get_config", "dtl_model_calibration_seconds",
    "dtl_model_save", "dtl_model_load", "dtl_model_destroy", "dtl_mlp_param_count", "dtl_mlp_create",
    "dtl_mlp_destroy", "dtl_mlp_train", "dtl_mlp_train_ex", "dtl_mlp_predict", "dtl_mlp_gradient", "dtl_mlp_get_params",
    "dtl_mlp_set_params", "dtl_mlp_input_count", "dtl_scratch_release", "dtl_scratch_bytes",
    "bbf_probe_noalloc", "dtl_noalloc_max_keys",
]

_lib = None


def load(path=LIB_PATH):
    """Loads the library.  torch is imported first so that the library binds to
    the HIP runtime torch already uses (one runtime per process)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libdtl_hip.so is not built (%s); run __graft_entry__.build()" % path)
    import torch  # noqa: F401  (binds the process to torch's HIP runtime)
    L = C.CDLL(path)
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.dtl_bbf_desc_init.argtypes = [C.POINTER(BbfDesc), u32, u64, u32, u32, u32, u32]
    L.dtl_cf_desc_init.argtypes = [C.POINTER(CfDesc), u64, u32, u32]
    L.bbf_build.argtypes = [C.POINTER(BbfDesc), vp, vp, u64, vp]
    L.bbf_probe.argtypes = [C.POINTER(BbfDesc), vp, vp, u64, u32, vp, vp, vp]
    L.bbf_probe_ex.argtypes = [C.POINTER(BbfDesc), vp, vp, u64, u32, vp, vp, C.POINTER(ProbeParams), vp]
    L.bbf_probe_noalloc.argtypes = [C.POINTER(BbfDesc), vp, vp, u64, u32, vp, vp, vp]
    L.dtl_noalloc_max_keys.argtypes = []
    L.dtl_noalloc_max_keys.restype = u64
    L.cf_probe.argtypes = [C.POINTER(CfDesc), vp, vp, u64, u32, vp, vp, vp]
    L.cf_probe_ex.argtypes = [C.POINTER(CfDesc), vp, vp, u64, u32, vp, vp, C.POINTER(ProbeParams), vp]
    L.cf_build_host.argtypes = [C.POINTER(CfDesc), vp, vp, u64, u32, C.POINTER(C.c_uint64)]
    L.cf_build.argtypes = [C.POINTER(CfDesc), vp, vp, u64, u32, u32, vp, vp]
    L.bbf_batch_plan_create.argtypes = [C.POINTER(BbfTask), u32, C.POINTER(ProbeParams), C.POINTER(vp)]
    L.bbf_batch_plan_run.argtypes = [vp, vp]
    L.bbf_batch_plan_destroy.argtypes = [vp]
    L.bbf_batch_plan_destroy.restype = None
    L.bbf_probe_batched.argtypes = [C.POINTER(BbfTask), u32, C.POINTER(ProbeParams), vp]
    L.bbf_batch_plan_tune.argtypes = [vp, vp, u32, C.POINTER(ProbeParams)]
    L.bbf_batch_plan_tune.restype = i32
    L.dtl_status_string.argtypes = [i32]
    L.dtl_status_string.restype = C.c_char_p
    L.dtl_last_error.argtypes = []
    L.dtl_last_error.restype = C.c_char_p
    L.dtl_bbf_supported.argtypes = [u32, u32, u32, u32, u32]
    L.dtl_bbf_config_count.argtypes = []
    L.dtl_bbf_split_lanes.argtypes = [u32, u32, u32, u32, u32]
    L.dtl_version.argtypes = []
    L.dtl_gather_bench.argtypes = [vp, u64, u32, u64, vp, vp]
    L.bbf_batch_plan_group_count.argtypes = [vp]
    L.bbf_batch_plan_group_count.restype = u32
    L.bbf_batch_plan_group_shape.argtypes = [vp, u32, C.POINTER(ProbeParams), C.POINTER(u32), C.POINTER(u64),
                                             C.POINTER(u32)]
    f64, f32 = C.c_double, C.c_float
    L.dtl_model_calibrate.argtypes = [C.POINTER(CalibParams), C.POINTER(vp), vp]
    L.dtl_model_train.argtypes = [vp, f64, u32, u32, u32, vp]
    L.dtl_model_choose.argtypes = [vp, u64, f64, i32, C.POINTER(BbfChoice), vp]
    L.dtl_model_cost.argtypes = [vp, C.POINTER(BbfChoice), u64, f64]
    L.dtl_model_cost.restype = f64
    L.dtl_model_lookup_ns.argtypes = [vp, i32, f64]
    L.dtl_model_lookup_ns.restype = f64
    L.dtl_model_fpr.argtypes = [vp, i32, f64]
    L.dtl_model_fpr.restype = f64
    L.dtl_model_config_count.argtypes = [vp]
    L.dtl_model_get_config.argtypes = [vp, i32, C.POINTER(BbfChoice)]
    L.dtl_model_calibration_seconds.argtypes = [vp]
    L.dtl_model_calibration_seconds.restype = f64
    L.dtl_model_save.argtypes = [vp, C.c_char_p]
    L.dtl_model_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.dtl_model_destroy.argtypes = [vp]
    L.dtl_model_destroy.restype = None
    L.dtl_mlp_param_count.argtypes = []
    L.dtl_mlp_input_count.argtypes = []
    L.dtl_scratch_release.argtypes = [vp]
    L.dtl_scratch_bytes.argtypes = [vp, C.POINTER(C.c_int)]
    L.dtl_scratch_bytes.restype = u64
    L.dtl_mlp_create.argtypes = [u32, C.POINTER(vp)]
    L.dtl_mlp_destroy.argtypes = [vp]
    L.dtl_mlp_destroy.restype = None
    L.dtl_mlp_train.argtypes = [vp, vp, vp, vp, u32, u32, u32, f32, u32, vp]
    L.dtl_mlp_train_ex.argtypes = [vp, vp, vp, vp, vp, f32, f32, u32, u32, u32, f32, u32, vp]
    L.dtl_mlp_predict.argtypes = [vp, vp, u32, vp, vp]
    L.dtl_mlp_gradient.argtypes = [vp, vp, vp, vp, u32, vp, vp]
    L.dtl_mlp_get_params.argtypes = [vp, C.POINTER(C.c_float)]
    L.dtl_mlp_set_params.argtypes = [vp, C.POINTER(C.c_float)]
    for name in ("dtl_model_calibrate", "dtl_model_train", "dtl_model_choose", "dtl_model_config_count",
                 "dtl_model_get_config", "dtl_model_save", "dtl_model_load", "dtl_mlp_param_count",
                 "dtl_mlp_input_count", "dtl_scratch_release",
                 "dtl_mlp_create", "dtl_mlp_train", "dtl_mlp_train_ex", "dtl_mlp_predict", "dtl_mlp_gradient",
                 "dtl_mlp_get_params", "dtl_mlp_set_params"):
        getattr(L, name).restype = i32
    for name in ("dtl_bbf_desc_init", "dtl_cf_desc_init", "bbf_build", "bbf_probe", "bbf_probe_ex", "bbf_probe_noalloc",
                 "cf_probe", "cf_probe_ex", "cf_build_host", "cf_build", "bbf_batch_plan_create", "bbf_batch_plan_run",
                 "bbf_probe_batched", "dtl_bbf_supported", "dtl_bbf_config_count", "dtl_bbf_split_lanes",
                 "dtl_version",
                 "dtl_gather_bench"):
        getattr(L, name).restype = i32
    _lib = L
    return L


def status_string(status):
    return load().dtl_status_string(status).decode()


def _check(fn, rc):
    if rc != OK:
        raise DtlError(fn, rc)


def bbf_desc(word_bits, m, k, w=1, s=1, z=1):
    d = BbfDesc()
    _check("dtl_bbf_desc_init", load().dtl_bbf_desc_init(C.byref(d), word_bits, m, k, w, s, z))
    return d


def bbf_desc_rc(word_bits, m, k, w=1, s=1, z=1):
    d = BbfDesc()
    return load().dtl_bbf_desc_init(C.byref(d), word_bits, m, k, w, s, z)


def cf_desc(m, tag_bits=16, tags_per_bucket=4):
    d = CfDesc()
    _check("dtl_cf_desc_init", load().dtl_cf_desc_init(C.byref(d), m, tag_bits, tags_per_bucket))
    return d


def _stream(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _dev(t, what, itemsize=None, min_bytes=None, device=None, dtype=None):
    """Checks that `t` is a contiguous GPU tensor (with elements of `itemsize`
    bytes, of at least `min_bytes` bytes, on `device`, of `dtype`, when
    given): the library reads and writes device memory only, so a host or
    undersized tensor would reach a kernel as an invalid address.  Raises
    ValueError before any call into the library."""
    if not getattr(t, "is_cuda", False):
        raise ValueError("%s must be a tensor in GPU memory" % what)
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % what)
    if itemsize is not None and t.element_size() != itemsize:
        raise ValueError("%s must have %d-byte elements (got %s)" % (what, itemsize, t.dtype))
    if dtype is not None and t.dtype != dtype:
        raise ValueError("%s must be %s (got %s)" % (what, dtype, t.dtype))
    if min_bytes is not None and t.numel() * t.element_size() < min_bytes:
        raise ValueError("%s holds %d bytes, needs %d" % (what, t.numel() * t.element_size(), min_bytes))
    if device is not None and t.device != device:
        raise ValueError("%s is on %s, the other operands on %s" % (what, t.device, device))
    return t


def _outputs(out, count, n, device):
    """The position buffer (>= max(n, 1) 4-byte entries) and the device
    uint64 count of a probe, allocated when not given."""
    import torch
    if out is None:
        out = torch.empty(max(n, 1), dtype=torch.int32, device=device)
    if count is None:
        count = torch.zeros(1, dtype=torch.int64, device=device)
    _dev(out, "out", 4, min_bytes=4 * max(n, 1), device=device)
    _dev(count, "count", 8, min_bytes=8, device=device)
    return out, count


def new_filter(desc, device="cuda"):
    """Zeroed device buffer for a blocked Bloom filter (uint8 tensor)."""
    import torch
    return torch.zeros(int(desc.size_bytes), dtype=torch.uint8, device=device)


def bbf_build(desc, words, keys, stream=None):
    _dev(words, "words", min_bytes=int(desc.size_bytes))
    _dev(keys, "keys", 4, device=words.device)
    _check("bbf_build", load().bbf_build(C.byref(desc), _ptr(words), _ptr(keys), keys.numel(),
                                         _stream(stream)))


def bbf_probe(desc, words, keys, match_offset=0, params=None, stream=None, out=None, count=None):
    """Returns (positions tensor of capacity n, device count tensor)."""
    _dev(words, "words", min_bytes=int(desc.size_bytes))
    _dev(keys, "keys", 4, device=words.device)
    n = keys.numel()
    out, count = _outputs(out, count, n, keys.device)
    L = load()
    if params is None:
        rc = L.bbf_probe(C.byref(desc), _ptr(words), _ptr(keys), n, match_offset, _ptr(out),
                         _ptr(count), _stream(stream))
    else:
        rc = L.bbf_probe_ex(C.byref(desc), _ptr(words), _ptr(keys), n, match_offset, _ptr(out),
                            _ptr(count), C.byref(params), _stream(stream))
    _check("bbf_probe", rc)
    return out, count


def bbf_probe_noalloc(desc, words, keys, match_offset=0, stream=None, out=None, count=None):
    """bbf_probe_noalloc: the probe with the library's fixed per-GPU scratch
    (no device allocation in the library); returns (positions, count)."""
    _dev(words, "words", min_bytes=int(desc.size_bytes))
    _dev(keys, "keys", 4, device=words.device)
    n = keys.numel()
    out, count = _outputs(out, count, n, keys.device)
    _check("bbf_probe_noalloc", load().bbf_probe_noalloc(C.byref(desc), _ptr(words), _ptr(keys), n, match_offset,
                                                         _ptr(out), _ptr(count), _stream(stream)))
    return out, count


def positions(out, count):
    """Host numpy copy (uint32) of the first `count` positions."""
    import numpy as np
    c = int(count.item())
    return out[:c].cpu().numpy().view(np.uint32).copy()


def cf_build_host(desc, keys_np, seed=0):
    """Host build into a zeroed numpy buffer of desc.buffer_bytes (padded to 8)."""
    import numpy as np
    keys_np = np.ascontiguousarray(keys_np, dtype=np.uint32)
    buf = np.zeros((desc.buffer_bytes + 7) // 8 * 8, dtype=np.uint8)
    ins = C.c_uint64(0)
    _check("cf_build_host", load().cf_build_host(C.byref(desc), C.c_void_p(buf.ctypes.data),
                                                 C.c_void_p(keys_np.ctypes.data), len(keys_np),
                                                 seed, C.byref(ins)))
    return buf, ins.value


def new_cf_table(desc):
    """Zeroed device table + victim cache (desc.buffer_bytes, padded to 8) as uint8."""
    import torch
    return torch.zeros((desc.buffer_bytes + 7) // 8 * 8, dtype=torch.uint8, device="cuda")


def cf_build(desc, buf, keys, seed=0, max_kicks=0, failed=None, stream=None):
    """GPU build (cf_build): inserts the device keys into the device table
    `buf`; returns the device uint64 count of walks whose tag found no place
    (added to `failed` when given).  0 means a valid table."""
    import torch
    _dev(buf, "table", min_bytes=int(desc.buffer_bytes))
    _dev(keys, "keys", 4, device=buf.device)
    if failed is None:
        failed = torch.zeros(1, dtype=torch.int64, device=buf.device)
    _dev(failed, "failed", 8, min_bytes=8, device=buf.device)
    _check("cf_build", load().cf_build(C.byref(desc), _ptr(buf), _ptr(keys), keys.numel(), seed, max_kicks,
                                       _ptr(failed), _stream(stream)))
    return failed


def cf_probe(desc, buf, keys, match_offset=0, stream=None, out=None, count=None, params=None):
    _dev(buf, "table", min_bytes=int(desc.buffer_bytes))
    _dev(keys, "keys", 4, device=buf.device)
    n = keys.numel()
    out, count = _outputs(out, count, n, keys.device)
    L = load()
    if params is None:
        rc = L.cf_probe(C.byref(desc), _ptr(buf), _ptr(keys), n, match_offset, _ptr(out), _ptr(count),
                        _stream(stream))
    else:
        rc = L.cf_probe_ex(C.byref(desc), _ptr(buf), _ptr(keys), n, match_offset, _ptr(out), _ptr(count),
                           C.byref(params), _stream(stream))
    _check("cf_probe", rc)
    return out, count


class BatchPlan:
    """A bbf batch plan over a list of (desc, words, keys, match_offset) tasks."""

    def __init__(self, tasks, params=None):
        import torch
        self.L = load()
        self.outs, self.counts = [], []
        arr = (BbfTask * len(tasks))()
        for i, (desc, words, keys, off) in enumerate(tasks):
            out = torch.empty(max(keys.numel(), 1), dtype=torch.int32, device=keys.device)
            cnt = torch.zeros(1, dtype=torch.int64, device=keys.device)
            self.outs.append(out)
            self.counts.append(cnt)
            arr[i].desc = desc
            arr[i].d_words = words.data_ptr()
            arr[i].d_keys = keys.data_ptr()
            arr[i].n = keys.numel()
            arr[i].match_offset = off
            arr[i].d_pos = out.data_ptr()
            arr[i].d_count = cnt.data_ptr()
        self._keep = tasks
        self.plan = C.c_void_p()
        pp = C.byref(params) if params is not None else None
        _check("bbf_batch_plan_create",
               self.L.bbf_batch_plan_create(arr, len(tasks), pp, C.byref(self.plan)))

    def run(self, stream=None):
        _check("bbf_batch_plan_run", self.L.bbf_batch_plan_run(self.plan, _stream(stream)))

    def tune(self, reps=3, stream=None):
        best = ProbeParams()
        _check("bbf_batch_plan_tune", self.L.bbf_batch_plan_tune(self.plan, _stream(stream), reps, C.byref(best)))
        return best

    def results(self):
        return [positions(o, c) for o, c in zip(self.outs, self.counts)]

    def group_count(self):
        return self.L.bbf_batch_plan_group_count(self.plan)

    def group_shape(self, group):
        """(ProbeParams of the group's launch shape, tasks, probe keys, largest filter bytes)."""
        shape, nt, keys, mfb = ProbeParams(), C.c_uint32(), C.c_uint64(), C.c_uint32()
        _check("bbf_batch_plan_group_shape", self.L.bbf_batch_plan_group_shape(
            self.plan, group, C.byref(shape), C.byref(nt), C.byref(keys), C.byref(mfb)))
        return shape, nt.value, keys.value, mfb.value

    def close(self):
        if self.plan:
            self.L.bbf_batch_plan_destroy(self.plan)
            self.plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scratch_release(stream=None):
    """Frees the probe scratch of the stream (dtl_scratch_release)."""
    _check("dtl_scratch_release", load().dtl_scratch_release(_stream(stream)))


def scratch_bytes(stream=None):
    """(bytes, pinned) of the probe scratch held for the stream."""
    cap = C.c_int(0)
    b = load().dtl_scratch_bytes(_stream(stream), C.byref(cap))
    return int(b), bool(cap.value)


def gather_bench(buf, mode, n_reads, sink, stream=None):
    _check("dtl_gather_bench", load().dtl_gather_bench(_ptr(buf), buf.numel() * buf.element_size(), mode,
                                                       n_reads, _ptr(sink), _stream(stream)))


class MLP:
    """The library's FP32 MLP (dtl_mlp.hip): input_count (8) inputs, two
    hidden layers of 64, one output."""

    def __init__(self, seed=0):
        self.L = load()
        self.h = C.c_void_p()
        _check("dtl_mlp_create", self.L.dtl_mlp_create(seed, C.byref(self.h)))
        self.n_params = self.L.dtl_mlp_param_count()
        self.n_inputs = self.L.dtl_mlp_input_count()

    def params(self):
        import numpy as np
        out = np.zeros(self.n_params, dtype=np.float32)
        _check("dtl_mlp_get_params", self.L.dtl_mlp_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check("dtl_mlp_set_params", self.L.dtl_mlp_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def _rows(self, x, t=None, w=None):
        """Checks the feature matrix (n x input_count float32 on the GPU) and
        the per-row targets and weights before any launch."""
        import torch
        _dev(x, "x", dtype=torch.float32)
        if x.dim() != 2 or x.shape[1] != self.n_inputs:
            raise ValueError("x must be n x %d (got %s)" % (self.n_inputs, tuple(x.shape)))
        n = x.shape[0]
        for v, what in ((t, "t"), (w, "w")):
            if v is not None:
                _dev(v, what, dtype=torch.float32, device=x.device)
                if v.numel() != n:
                    raise ValueError("%s must hold %d values (got %d)" % (what, n, v.numel()))
        return n

    def predict(self, x):
        import torch
        n = self._rows(x)
        y = torch.empty(n, dtype=torch.float32, device=x.device)
        _check("dtl_mlp_predict", self.L.dtl_mlp_predict(self.h, _ptr(x), n, _ptr(y), _stream(None)))
        return y

    def gradient(self, x, t, w=None):
        import torch
        if t is None:
            raise ValueError("t is required")
        self._rows(x, t, w)
        g = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        _check("dtl_mlp_gradient", self.L.dtl_mlp_gradient(self.h, _ptr(x), _ptr(t),
                                                           _ptr(w) if w is not None else None,
                                                           x.shape[0], _ptr(g), _stream(None)))
        return g

    def train(self, x, t, epochs, batch, lr, seed=0, w=None):
        if t is None:
            raise ValueError("t is required")
        self._rows(x, t, w)
        _check("dtl_mlp_train", self.L.dtl_mlp_train(self.h, _ptr(x), _ptr(t), _ptr(w) if w is not None else None,
                                                     x.shape[0], epochs, batch, lr, seed, _stream(None)))

    def close(self):
        if self.h:
            self.L.dtl_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CostModel:
    """Calibrated cost table + learned MLP for choosing filter configurations."""

    def __init__(self, handle):
        self.L = load()
        self.h = handle

    @classmethod
    def calibrate(cls, params=None):
        L = load()
        h = C.c_void_p()
        pp = C.byref(params) if params is not None else None
        _check("dtl_model_calibrate", L.dtl_model_calibrate(pp, C.byref(h), _stream(None)))
        return cls(h)

    @classmethod
    def load_file(cls, path):
        L = load()
        h = C.c_void_p()
        _check("dtl_model_load", L.dtl_model_load(path.encode(), C.byref(h)))
        return cls(h)

    def save(self, path):
        _check("dtl_model_save", self.L.dtl_model_save(self.h, path.encode()))

    def train(self, train_fraction=1.0, samples=2000000, epochs=64, seed=1):
        _check("dtl_model_train", self.L.dtl_model_train(self.h, train_fraction, samples, epochs, seed,
                                                         _stream(None)))

    def choose(self, n, tw_ns, use_mlp=True):
        ch = BbfChoice()
        _check("dtl_model_choose", self.L.dtl_model_choose(self.h, n, tw_ns, int(use_mlp), C.byref(ch),
                                                           _stream(None)))
        return ch

    def cost(self, choice, n, tw_ns):
        return self.L.dtl_model_cost(self.h, C.byref(choice), n, tw_ns)

    def config_count(self):
        return self.L.dtl_model_config_count(self.h)

    def config(self, i):
        ch = BbfChoice()
        _check("dtl_model_get_config", self.L.dtl_model_get_config(self.h, i, C.byref(ch)))
        return ch

    def lookup_ns(self, i, m):
        return self.L.dtl_model_lookup_ns(self.h, i, m)

    def fpr(self, i, bpk):
        return self.L.dtl_model_fpr(self.h, i, bpk)

    def calibration_seconds(self):
        return self.L.dtl_model_calibration_seconds(self.h)

    def close(self):
        if self.h:
            self.L.dtl_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
