"""ctypes wrapper of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")


class GpuqsortStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("levels", "blocks_total", "max_blocks", "done_records",
                                               "max_done_len", "maxseq", "invariant_violations")]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        for t in ("u32", "f32", "f64"):
            getattr(lib, f"orc_std_sort_{t}").argtypes = [P, ctypes.c_size_t]
            getattr(lib, f"orc_par_sort_{t}").argtypes = [P, ctypes.c_size_t, ctypes.c_int]
            f = getattr(lib, f"orc_gpuqsort_{t}")
            f.argtypes = [P, P, ctypes.c_size_t] + [ctypes.c_uint32] * 4 + [ctypes.POINTER(GpuqsortStats)]
            f.restype = ctypes.c_int
        lib.orc_std_sort_u64.argtypes = [P, ctypes.c_size_t]
        lib.orc_mismatch.argtypes = [P, P, ctypes.c_size_t, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        lib.orc_mismatch.restype = ctypes.c_size_t
        for f in ("orc_median_u32", "orc_host_median_u32", "orc_f32_key", "orc_f32_unkey"):
            getattr(lib, f).restype = ctypes.c_uint32
        lib.orc_median_u32.argtypes = [ctypes.c_uint32] * 3
        lib.orc_host_median_u32.argtypes = [ctypes.c_uint32] * 3
        lib.orc_f32_key.argtypes = [ctypes.c_uint32]
        lib.orc_f32_unkey.argtypes = [ctypes.c_uint32]
        lib.orc_f64_key.argtypes = [ctypes.c_uint64]
        lib.orc_f64_key.restype = ctypes.c_uint64
        lib.orc_f64_unkey.argtypes = [ctypes.c_uint64]
        lib.orc_f64_unkey.restype = ctypes.c_uint64
        for t in ("u32", "f32", "f64"):
            getattr(lib, f"orc_std_sort_segmented_{t}").argtypes = [P, P, ctypes.c_size_t]
        lib.orc_keys_f32.argtypes = [P, P, ctypes.c_size_t]
        lib.orc_keys_f64.argtypes = [P, P, ctypes.c_size_t]
        lib.orc_optp.argtypes = [ctypes.c_size_t, ctypes.c_double, ctypes.c_size_t]
        lib.orc_optp.restype = ctypes.c_size_t
        for t in ("u32", "u64"):
            f = getattr(lib, f"orc_kway_classify_{t}")
            f.argtypes = [P, ctypes.c_size_t, P, P, ctypes.c_int, P, P]
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


_SUFFIX = {np.dtype(np.uint32): "u32", np.dtype(np.float32): "f32", np.dtype(np.float64): "f64",
           np.dtype(np.uint64): "u64"}


def std_sort(a):
    """std::sort of a copy of a (the reference's own oracle)."""
    b = np.array(a, copy=True)
    getattr(load(), f"orc_std_sort_{_SUFFIX[b.dtype]}")(b.ctypes.data, b.size)
    return b


def std_sort_segmented(a, offsets):
    """std::sort of each array [offsets[i], offsets[i+1]) of a copy of a."""
    b = np.array(a, copy=True)
    o = np.ascontiguousarray(offsets, dtype=np.uint64)
    getattr(load(), f"orc_std_sort_segmented_{_SUFFIX[b.dtype]}")(b.ctypes.data, o.ctypes.data, o.size - 1)
    return b


def ordered_keys(a):
    """Keys in the sort's ordered-key representation: uint32 unchanged, float
    and double bits through the order-preserving bijection."""
    if a.dtype == np.uint32:
        return np.array(a, copy=True)
    if a.dtype == np.float32:
        src = np.ascontiguousarray(a).view(np.uint32)
        out = np.empty_like(src)
        load().orc_keys_f32(src.ctypes.data, out.ctypes.data, src.size)
        return out
    src = np.ascontiguousarray(a).view(np.uint64)
    out = np.empty_like(src)
    load().orc_keys_f64(src.ctypes.data, out.ctypes.data, src.size)
    return out


def par_sort(a, threads):
    b = np.array(a, copy=True)
    getattr(load(), f"orc_par_sort_{_SUFFIX[b.dtype]}")(b.ctypes.data, b.size, threads)
    return b


def mismatch(a, b):
    """(number of bitwise-different keys, first index) of two equal-length arrays."""
    assert a.dtype == b.dtype and a.size == b.size
    first = ctypes.c_size_t()
    n = load().orc_mismatch(a.ctypes.data, b.ctypes.data, a.size, a.itemsize, ctypes.byref(first))
    return int(n), int(first.value)


def gpuqsort(a, qbs=1728, gq_wg=128, lq_wg=128, st=256):
    """CPU restatement of the reference's GPUQSort; returns (sorted copy, stats)."""
    d = np.array(a, copy=True)
    dn = np.zeros_like(d)
    s = GpuqsortStats()
    rc = getattr(load(), f"orc_gpuqsort_{_SUFFIX[d.dtype]}")(d.ctypes.data, dn.ctypes.data, d.size, qbs, gq_wg,
                                                             lq_wg, st, ctypes.byref(s))
    if rc != 0:
        raise ValueError(f"orc_gpuqsort rc={rc}")
    return d, s


def kway_classify(keys, splitters, eq):
    """Brute-force bucket ids and counts for a k-way partition with equality buckets."""
    keys = np.ascontiguousarray(keys)
    u = np.ascontiguousarray(splitters, dtype=keys.dtype)
    e = np.ascontiguousarray(eq, dtype=np.uint8)
    out = np.empty(keys.size, dtype=np.uint16)
    counts = np.zeros(2 * len(u) + 1, dtype=np.uint64)
    t = "u64" if keys.dtype.itemsize == 8 else "u32"
    nb = getattr(load(), f"orc_kway_classify_{t}")(keys.ctypes.data, keys.size, u.ctypes.data, e.ctypes.data,
                                                   len(u), out.ctypes.data, counts.ctypes.data)
    return out, counts[:nb]


def load_ref():
    """oracle/_ref/libref.so (the reference header compiled as-is), or None."""
    if not os.path.exists(REF_LIB):
        return None
    lib = ctypes.CDLL(REF_LIB)
    lib.ref_host_median_u32.argtypes = [ctypes.c_uint32] * 3
    lib.ref_host_median_u32.restype = ctypes.c_uint32
    lib.ref_host_median_f64.argtypes = [ctypes.c_double] * 3
    lib.ref_host_median_f64.restype = ctypes.c_double
    lib.ref_record_sizes.argtypes = [ctypes.POINTER(ctypes.c_size_t)]
    lib.ref_tunables.argtypes = [ctypes.POINTER(ctypes.c_int)]
    return lib
