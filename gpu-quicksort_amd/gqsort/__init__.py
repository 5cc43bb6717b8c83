"""Python binding of libgqsort.so (the MI355X GPU-Quicksort C ABI).

The binding mirrors the reference's entry point

    template <class T> void GPUQSort(OCLResources*, size_t size, T* d, T* dn)
    (GPU-Quicksort_OpenCL_1.2/QuicksortMain.cpp:297-298)

as ``GPUQSort(ctx, size, d, dn)``: ``d`` (numpy, host) holds the keys and is
sorted in place, ``dn`` is scratch of the same size.  Unlike the reference it
raises ``GQSError`` on failure instead of asserting.

Device-resident sorts take raw device pointers (or torch tensors) and a
stream.  There is no CPU fallback: if the shared library is missing or no GPU
is present, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgqsort.so")

U32, F32, F64 = 0, 1, 2
MODE_MULTIPIVOT, MODE_CLASSIC = 0, 1
DTYPE_CODES = {np.dtype(np.uint32): U32, np.dtype(np.float32): F32, np.dtype(np.float64): F64}
NP_DTYPES = {U32: np.uint32, F32: np.float32, F64: np.float64}
KEY_BYTES = {U32: 4, F32: 4, F64: 8}

STATUS = {0: "GQS_OK", 1: "GQS_EINVAL", 2: "GQS_ENOMEM", 3: "GQS_EDEVICE", 4: "GQS_EOVERFLOW",
          5: "GQS_EBUSY", 6: "GQS_ENOTFOUND"}

# Generators of gqs_generate.
GEN_UNIFORM, GEN_GAUSSIAN, GEN_SORTED, GEN_REVERSE, GEN_EQUAL, GEN_BUCKET, GEN_PERMUTATION, GEN_NORMAL_FP = range(8)
GENERATORS = {"uniform": 0, "gaussian": 1, "sorted": 2, "reverse": 3, "equal": 4, "bucket": 5,
              "permutation": 6, "normal": 7}

KERNEL_KINDS = ["splitters", "tileplan", "count", "scan", "plan", "scatter", "leaf", "init", "fill", "tilepart"]


class GQSError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}")


class Params(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("leaf_max", ctypes.c_uint32), ("leaf_target", ctypes.c_uint32),
                ("oversample", ctypes.c_uint32), ("max_fanout", ctypes.c_uint32), ("max_levels", ctypes.c_uint32),
                ("host_driven", ctypes.c_uint32), ("part_wgs_per_cu", ctypes.c_uint32),
                ("leaf_wgs_per_cu", ctypes.c_uint32), ("sort_threshold", ctypes.c_uint32),
                ("bitonic_cutoff", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("slot_keys", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_requests", ctypes.c_uint32), ("batch_bytes", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "batches", "max_batch_requests",
                                               "keys_sorted", "pending")]


class DeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("arch", ctypes.c_char * 256), ("pci_bus_id", ctypes.c_char * 32),
                ("driver_version", ctypes.c_int32), ("runtime_version", ctypes.c_int32),
                ("compute_units", ctypes.c_int32), ("wave_size", ctypes.c_int32),
                ("max_threads_per_block", ctypes.c_int32), ("max_block_dim", ctypes.c_int32 * 3),
                ("clock_khz", ctypes.c_int32), ("mem_clock_khz", ctypes.c_int32), ("mem_bus_bits", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("lds_per_block", ctypes.c_uint64), ("lds_per_cu", ctypes.c_uint64),
                ("l2_bytes", ctypes.c_uint64), ("global_mem_bytes", ctypes.c_uint64)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            if isinstance(v, bytes):
                v = v.decode(errors="replace")
            elif k == "max_block_dim":
                v = list(v)
            if k != "reserved":
                out[k] = v
        return out


class Stats(ctypes.Structure):
    _fields_ = [("levels_launched", ctypes.c_uint32), ("levels_used", ctypes.c_uint32),
                ("oversize_leaves", ctypes.c_uint32), ("overflow", ctypes.c_uint32),
                ("segments", ctypes.c_uint32 * 8), ("level_keys", ctypes.c_uint64 * 8)]


# Every symbol include/gqsort.h declares, with its ctypes signature.
_P = ctypes.c_void_p
_SIGS = {
    "gqs_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "gqs_destroy": (ctypes.c_int, [_P]),
    "gqs_last_error": (ctypes.c_char_p, [_P]),
    "gqs_get_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "gqs_set_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "gqs_default_params": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(Params)]),
    "gqs_check_params": (ctypes.c_int, [ctypes.POINTER(Params)]),
    "gqs_partition_level": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t, _P, _P,
                                           ctypes.POINTER(ctypes.c_uint32), _P, _P, _P,
                                           ctypes.POINTER(ctypes.c_uint32)]),
    "gqs_set_queue_limit": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gqs_set_level_scheme": (ctypes.c_int, [_P, ctypes.c_int]),
    "gqs_graph_status": (ctypes.c_int, [_P]),
    "gqs_device_report": (ctypes.c_int, [_P, ctypes.POINTER(DeviceProps)]),
    "gqs_server_resolve_config": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(ServerConfig)]),
    "gqs_sort": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t]),
    "gqs_sort_device": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t, _P]),
    "gqs_set_large_run": (ctypes.c_int, [_P, ctypes.c_size_t]),
    "gqs_set_host_pipeline": (ctypes.c_int, [_P, ctypes.c_size_t]),
    "gqs_sort_device_noalloc": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t, _P]),
    "gqs_sort_segmented": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_size_t, ctypes.c_size_t, _P]),
    "gqs_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(Stats)]),
    "gqs_set_profiling": (ctypes.c_int, [_P, ctypes.c_int]),
    "gqs_kernel_times": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_int]),
    "gqs_kernel_trace": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float),
                                         ctypes.c_int]),
    "gqs_device_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_size_t]),
    "gqs_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _P, ctypes.c_size_t]),
    "gqs_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_uint32,
                                    ctypes.POINTER(Params)]),
    "gqs_tune_save": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gqs_tune_load": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gqs_sort_each": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, ctypes.c_size_t, _P]),
    "gqs_graph_create_segmented": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P, _P, ctypes.c_size_t, ctypes.c_size_t,
                                                   ctypes.POINTER(_P)]),
    "gqs_graph_launch": (ctypes.c_int, [_P, _P]),
    "gqs_graph_destroy": (ctypes.c_int, [_P]),
    "gqs_server_default_config": (ctypes.c_int, [ctypes.POINTER(ServerConfig)]),
    "gqs_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "gqs_server_submit": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64)]),
    "gqs_server_wait": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_int64]),
    "gqs_server_poll": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]),
    "gqs_server_request_times": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "gqs_server_release": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "gqs_server_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    "gqs_server_destroy": (ctypes.c_int, [_P]),
    "gqs_server_get_config": (ctypes.c_int, [_P, ctypes.POINTER(ServerConfig)]),
    "gqs_malloc": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "gqs_free": (ctypes.c_int, [_P, _P]),
    "gqs_host_alloc": (ctypes.c_int, [_P, ctypes.c_size_t, ctypes.POINTER(_P)]),
    "gqs_host_free": (ctypes.c_int, [_P, _P]),
    "gqs_memcpy": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P]),
    "gqs_memcpy_async": (ctypes.c_int, [_P, _P, _P, ctypes.c_size_t, _P]),
    "gqs_stream_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "gqs_stream_destroy": (ctypes.c_int, [_P, _P]),
    "gqs_stream_sync": (ctypes.c_int, [_P, _P]),
    "gqs_event_create": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "gqs_event_destroy": (ctypes.c_int, [_P, _P]),
    "gqs_event_record": (ctypes.c_int, [_P, _P, _P]),
    "gqs_event_elapsed": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(ctypes.c_float)]),
}
OPTIONAL_SIGS = {}

_lib = None


def load(path=LIB_PATH):
    """Load libgqsort.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built; run gpu-quicksort_amd/build.py")
    lib = ctypes.CDLL(path)
    for name, (res, args) in {**_SIGS, **OPTIONAL_SIGS}.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(x):
    """Device address of a DeviceBuffer, a torch tensor, or an int address."""
    if x is None:
        return None
    if hasattr(x, "ptr"):
        return x.ptr
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    return int(x)


def _stream(s):
    if s is None:
        return None
    if hasattr(s, "handle") and isinstance(s, Stream):
        return s.handle
    if hasattr(s, "cuda_stream"):
        return s.cuda_stream
    return int(s)


def check_params(p):
    """True when every field of p is in range (gqs_check_params; no device needed)."""
    return load().gqs_check_params(ctypes.byref(p)) == 0


def resolve_server_config(**kw):
    """The configuration a server created with these fields runs with."""
    cfg = ServerConfig()
    _check(load().gqs_server_default_config(ctypes.byref(cfg)))
    for k, v in kw.items():
        setattr(cfg, k, v)
    out = ServerConfig()
    _check(load().gqs_server_resolve_config(ctypes.byref(cfg), ctypes.byref(out)))
    return out


def default_params(dtype=U32):
    p = Params()
    _check(load().gqs_default_params(dtype, ctypes.byref(p)))
    return p


def _check(st, ctx=None):
    if st != 0:
        msg = ""
        if ctx is not None and ctx._h:
            m = load().gqs_last_error(ctx._h)
            msg = m.decode() if m else ""
        raise GQSError(st, msg)


TUNED_TABLE = os.path.join(_HERE, "tuned_mi355x.json")


class Context:
    """Owns a gqs_ctx: device, workspace, tuned parameters.  With tuned=True
    the shipped autotuner table for MI355X is loaded (gqs_tune_load)."""

    def __init__(self, device=0, tuned=True):
        self._lib = load()
        h = _P()
        _check(self._lib.gqs_create(ctypes.byref(h), device))
        self._h = h
        if tuned and os.path.exists(TUNED_TABLE):
            self.tune_load(TUNED_TABLE)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gqs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def params(self):
        p = Params()
        _check(self._lib.gqs_get_params(self._h, ctypes.byref(p)), self)
        return p

    def set_params(self, p=None, **kw):
        q = self.params() if p is None else p
        for k, v in kw.items():
            setattr(q, k, v)
        _check(self._lib.gqs_set_params(self._h, ctypes.byref(q)), self)
        return q

    def sort_host(self, d, dn=None):
        """Reference semantics: host arrays, synchronous, d sorted in place."""
        d = np.ascontiguousarray(d) if not d.flags.c_contiguous else d
        dt = DTYPE_CODES[d.dtype]
        if dn is None:
            dn = np.empty_like(d)
        _check(self._lib.gqs_sort(self._h, dt, d.ctypes.data, dn.ctypes.data, d.size), self)
        return d

    def sort_device(self, dtype, d, dn, n, stream=None):
        _check(self._lib.gqs_sort_device(self._h, dtype, _ptr(d), _ptr(dn), n, _stream(stream)), self)

    def set_host_pipeline(self, min_keys=0):
        """Smallest n for which gqs_sort overlaps its copies with the sort
        (0: the default 2^24; 2**64 - 1: never)."""
        _check(self._lib.gqs_set_host_pipeline(self._h, min_keys), self)

    def set_large_run(self, keys=0):
        """Sorts above `keys` keys run as sorted runs of `keys` keys merged
        pairwise (0: only n >= 2^32 - 2^16, with runs of 2^31 keys)."""
        _check(self._lib.gqs_set_large_run(self._h, keys), self)

    def sort_segmented(self, dtype, d, dn, offsets, nseg, total, stream=None):
        _check(self._lib.gqs_sort_segmented(self._h, dtype, _ptr(d), _ptr(dn), _ptr(offsets), nseg, total,
                                            _stream(stream)), self)

    def autotune(self, dtype, n, gen="uniform", reps=3):
        """Tune parameters for (dtype, size class of n); returns the best Params."""
        p = Params()
        g = GENERATORS[gen] if isinstance(gen, str) else gen
        _check(self._lib.gqs_autotune(self._h, dtype, n, g, reps, ctypes.byref(p)), self)
        return p

    def tune_save(self, path):
        _check(self._lib.gqs_tune_save(self._h, os.fsencode(path)), self)

    def tune_load(self, path):
        _check(self._lib.gqs_tune_load(self._h, os.fsencode(path)), self)

    def stats(self):
        s = Stats()
        _check(self._lib.gqs_get_stats(self._h, ctypes.byref(s)), self)
        return s

    def set_profiling(self, on=True):
        _check(self._lib.gqs_set_profiling(self._h, 1 if on else 0), self)

    def kernel_times(self):
        n = len(KERNEL_KINDS)
        launches = (ctypes.c_uint32 * n)()
        ms = (ctypes.c_double * n)()
        _check(self._lib.gqs_kernel_times(self._h, launches, ms, n), self)
        return {k: (launches[i], ms[i]) for i, k in enumerate(KERNEL_KINDS)}

    def kernel_trace(self):
        """Per-launch (kind, ms) list of the last profiled sort (after kernel_times())."""
        n = self._lib.gqs_kernel_trace(self._h, None, None, 0)
        kinds = (ctypes.c_int * max(n, 1))()
        ms = (ctypes.c_float * max(n, 1))()
        self._lib.gqs_kernel_trace(self._h, kinds, ms, n)
        return [(KERNEL_KINDS[kinds[i]], ms[i]) for i in range(n)]

    def partition_level(self, dtype, d, dn, n):
        """One partition level of device keys d[0:n] into dn (gqs_partition_level).
        Returns (splitters, eq flags, bucket starts, sizes, equality flags)."""
        kd = np.uint64 if dtype == F64 else np.uint32
        spl = np.zeros(256, kd)
        eq = np.zeros(256, np.uint8)
        ns, nb = ctypes.c_uint32(), ctypes.c_uint32()
        st = np.zeros(256, np.uint64)
        sz = np.zeros(256, np.uint32)
        be = np.zeros(256, np.uint8)
        _check(self._lib.gqs_partition_level(self._h, dtype, _ptr(d), _ptr(dn), n, spl.ctypes.data, eq.ctypes.data,
                                             ctypes.byref(ns), st.ctypes.data, sz.ctypes.data, be.ctypes.data,
                                             ctypes.byref(nb)), self)
        m, b = ns.value, nb.value
        return spl[:m], eq[:m], st[:b], sz[:b], be[:b]

    def set_queue_limit(self, segments):
        """Testing hook: cap the segment queue of later sorts (0: no cap)."""
        _check(self._lib.gqs_set_queue_limit(self._h, segments), self)

    def set_level_scheme(self, scheme):
        """Testing hook: 0 runs two tile-local levels where a sort plans two
        levels and a spare (the default), 1 a contiguous level before the
        tile-local one, 2 as 0 with 64-entry run tables (the multi-table path)."""
        _check(self._lib.gqs_set_level_scheme(self._h, scheme), self)

    def device_report(self):
        """hipGetDeviceProperties report (the reference's show_CL output)."""
        p = DeviceProps()
        _check(self._lib.gqs_device_report(self._h, ctypes.byref(p)), self)
        return p.as_dict()

    def device_info(self):
        cus, lds, wave = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        name = ctypes.create_string_buffer(128)
        _check(self._lib.gqs_device_info(self._h, ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(wave),
                                         name, 128), self)
        return {"cus": cus.value, "lds_per_wg": lds.value, "wave": wave.value, "arch": name.value.decode()}


def GPUQSort(ctx, size, d, dn):
    """Drop-in for GPUQSort<T>(pOCL, size, d, dn): sorts host array d[0:size]
    ascending in place, clobbering dn (CL12/QuicksortMain.cpp:297-384)."""
    if d.dtype not in DTYPE_CODES:
        raise GQSError(1, f"unsupported key type {d.dtype}")
    if d.size < size or dn.size < size or dn.dtype != d.dtype:
        raise GQSError(1, "d and dn must hold size keys of the same type")
    dt = DTYPE_CODES[d.dtype]
    _check(load().gqs_sort(ctx.handle, dt, d.ctypes.data, dn.ctypes.data, size), ctx)


def batch_lengths(nseg, lo, hi, seed, total=None):
    """Array lengths floor(2^U), U uniform on [log2 lo, log2 hi] (seeded); with
    `total`, rescaled so they sum to about `total` (each kept >= 1)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(np.log2(lo), np.log2(hi), nseg)
    lens = np.floor(2.0 ** u)
    if total is not None:
        lens = np.maximum(1, np.floor(lens * (total / lens.sum())))
    return lens.astype(np.uint64)


def request_stream(count, seed, lo_log2=10, hi_log2=26, mix=(0.6, 0.3, 0.1)):
    """A serving workload: `count` requests with sizes floor(2^U), U uniform on
    [lo_log2, hi_log2], and key types drawn 60% uint32 / 30% float / 10% double
    (seeded).  Returns (sizes, dtypes, seeds); request i's keys are
    generate(dtypes[i], "uniform" for uint32 else "normal", seeds[i], sizes[i])."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(lo_log2, hi_log2, count)).astype(np.int64)
    kinds = rng.choice([U32, F32, F64], size=count, p=list(mix)).astype(np.int64)
    seeds = 5000 + 7919 * np.int64(seed) + np.arange(count, dtype=np.int64)
    return sizes, kinds, seeds


def request_keys(dtype, seed, n):
    """Keys of one request of request_stream()."""
    return generate(int(dtype), "uniform" if int(dtype) == U32 else "normal", int(seed), int(n))


def generate(dtype, gen, seed, n):
    """Synthetic input of one of the benchmark generators (host numpy array)."""
    if isinstance(gen, str):
        gen = GENERATORS[gen]
    a = np.empty(n, dtype=NP_DTYPES[dtype])
    _check(load().gqs_generate(dtype, gen, seed, a.ctypes.data if n else None, n))
    return a


class SegmentedGraph:
    """A captured segmented sort (gqs_graph_create_segmented), replayed with launch()."""

    def __init__(self, ctx, dtype, d, dn, offsets, nseg_dev, max_nseg, max_total_keys):
        self.ctx = ctx
        h = _P()
        _check(ctx._lib.gqs_graph_create_segmented(ctx.handle, dtype, _ptr(d), _ptr(dn), _ptr(offsets),
                                                    _ptr(nseg_dev), max_nseg, max_total_keys, ctypes.byref(h)), ctx)
        self._h = h

    def launch(self, stream=None):
        _check(self.ctx._lib.gqs_graph_launch(self._h, _stream(stream)), self.ctx)

    def status(self):
        """Waits for the last replay; raises GQSError on bad offsets or overflow."""
        _check(self.ctx._lib.gqs_graph_status(self._h), self.ctx)

    def close(self):
        if self._h:
            self.ctx._lib.gqs_graph_destroy(self._h)
            self._h = None


class DeviceBuffer:
    """Device memory from the library's own HIP runtime."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = _P()
        _check(ctx._lib.gqs_malloc(ctx.handle, nbytes, ctypes.byref(p)), ctx)
        self.ptr = p.value

    @classmethod
    def from_array(cls, ctx, a, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(self.ctx._lib.gqs_memcpy(self.ctx.handle, self.ptr, a.ctypes.data, a.nbytes, _stream(stream)),
               self.ctx)

    def download(self, dtype, count, stream=None, offset_bytes=0):
        out = np.empty(count, dtype=dtype)
        _check(self.ctx._lib.gqs_memcpy(self.ctx.handle, out.ctypes.data, self.ptr + offset_bytes, out.nbytes,
                                        _stream(stream)), self.ctx)
        return out

    def free(self):
        if self.ptr:
            self.ctx._lib.gqs_free(self.ctx.handle, self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            if self.ctx._h:
                self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream of the library's runtime."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = _P()
        _check(ctx._lib.gqs_stream_create(ctx.handle, ctypes.byref(h)), ctx)
        self.handle = h.value

    def sync(self):
        _check(self.ctx._lib.gqs_stream_sync(self.ctx.handle, self.handle), self.ctx)

    def close(self):
        if self.handle:
            self.ctx._lib.gqs_stream_destroy(self.ctx.handle, self.handle)
            self.handle = None


class Event:
    def __init__(self, ctx):
        self.ctx = ctx
        h = _P()
        _check(ctx._lib.gqs_event_create(ctx.handle, ctypes.byref(h)), ctx)
        self.handle = h.value

    def record(self, stream=None):
        _check(self.ctx._lib.gqs_event_record(self.ctx.handle, self.handle, _stream(stream)), self.ctx)

    def elapsed_ms(self, stop):
        ms = ctypes.c_float()
        _check(self.ctx._lib.gqs_event_elapsed(self.ctx.handle, self.handle, stop.handle, ctypes.byref(ms)),
               self.ctx)
        return float(ms.value)

    def close(self):
        if self.handle:
            self.ctx._lib.gqs_event_destroy(self.ctx.handle, self.handle)
            self.handle = None


def device_sort(ctx, a, stream=None):
    """Upload host array a, sort it on the device, download the result."""
    dt = DTYPE_CODES[a.dtype]
    d = DeviceBuffer.from_array(ctx, a, stream)
    dn = DeviceBuffer(ctx, a.nbytes)
    ctx.sort_device(dt, d, dn, a.size, stream)
    out = d.download(a.dtype, a.size, stream)
    d.free()
    dn.free()
    return out


class Server:
    """Continuous-batching sort server (gqs_server_*)."""

    def __init__(self, device=0, hbm_budget_bytes=None, slots=None, slot_keys=None, max_batch_requests=None,
                 batch_bytes=None):
        self._lib = load()
        cfg = ServerConfig()
        _check(self._lib.gqs_server_default_config(ctypes.byref(cfg)))
        for k, v in (("hbm_budget_bytes", hbm_budget_bytes), ("slots", slots), ("slot_keys", slot_keys),
                     ("max_batch_requests", max_batch_requests), ("batch_bytes", batch_bytes)):
            if v is not None:
                setattr(cfg, k, v)
        h = _P()
        _check(self._lib.gqs_server_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.config = ServerConfig()
        _check(self._lib.gqs_server_get_config(self._h, ctypes.byref(self.config)))

    def submit(self, dtype, in_ptr, out_ptr, n):
        """Submit a request over raw host pointers; returns its ticket."""
        t = ctypes.c_uint64()
        _check(self._lib.gqs_server_submit(self._h, dtype, in_ptr, out_ptr, n, ctypes.byref(t)))
        return t.value

    def submit_array(self, a, out=None):
        out = np.empty_like(a) if out is None else out
        return self.submit(DTYPE_CODES[a.dtype], a.ctypes.data if a.size else None,
                           out.ctypes.data if a.size else None, a.size), out

    def wait(self, ticket, timeout_us=-1):
        _check(self._lib.gqs_server_wait(self._h, ticket, timeout_us))

    def times(self, ticket):
        s, a, d = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self._lib.gqs_server_request_times(self._h, ticket, ctypes.byref(s), ctypes.byref(a),
                                                  ctypes.byref(d)))
        return s.value, a.value, d.value

    def release(self, ticket):
        _check(self._lib.gqs_server_release(self._h, ticket))

    def stats(self):
        st = ServerStats()
        _check(self._lib.gqs_server_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self._h:
            self._lib.gqs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
