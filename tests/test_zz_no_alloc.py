"""The allocation-free entry point (gqs_sort_device_noalloc) against
gqs_sort_device: two calls in a row per key type, the second on a smaller
input, each bit-exact against the regular entry point's result (and
std::sort), with the device's free memory unchanged across the calls.  Sizes
above the 2^24-key maximum are refused before anything is launched."""
import ctypes

import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu

EINVAL = 1


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("dt,first,second", [
    (gqsort.U32, ("uniform", 1 << 24), ("gaussian", 300001)),
    (gqsort.F32, ("normal", 5_000_011), ("normal", 70001)),
    (gqsort.F64, ("normal", 1 << 23), ("normal", 4097)),
])
def test_noalloc_matches_sort_device(dt, first, second):
    lib = gqsort.load()
    ctx = gqsort.Context(0)
    stream = gqsort.Stream(ctx)
    try:
        n_max = first[1]
        kb = gqsort.KEY_BYTES[dt]
        d = gqsort.DeviceBuffer(ctx, n_max * kb)
        dn = gqsort.DeviceBuffer(ctx, n_max * kb)
        r = gqsort.DeviceBuffer(ctx, n_max * kb)
        rn = gqsort.DeviceBuffer(ctx, n_max * kb)
        for gen, n in (first, second):
            a = gqsort.generate(dt, gen, 44, n)
            print(f"case noalloc dtype={dt} {gen} n={n}", flush=True)
            # the regular entry point (allocates its workspace on first use)
            r.upload(a, stream)
            ctx.sort_device(dt, r, rn, n, stream)
            stream.sync()
            ref = r.download(a.dtype, n, stream)
            d.upload(a, stream)
            stream.sync()
            before = free_device_bytes()
            assert lib.gqs_sort_device_noalloc(ctx.handle, dt, d.ptr, dn.ptr, n, stream.handle) == 0, \
                lib.gqs_last_error(ctx.handle)
            stream.sync()
            after = free_device_bytes()
            got = d.download(a.dtype, n, stream)
            assert orc.mismatch(got, ref)[0] == 0, "noalloc result differs from gqs_sort_device"
            assert orc.mismatch(got, orc.std_sort(a))[0] == 0, "noalloc result differs from std::sort"
            assert after == before, f"device free memory changed by {before - after} bytes"
        # above the maximum: refused on the host
        st = lib.gqs_sort_device_noalloc(ctx.handle, dt, d.ptr, dn.ptr, (1 << 24) + 1, stream.handle)
        assert st == EINVAL and b"maximum" in lib.gqs_last_error(ctx.handle)
        for x in (d, dn, r, rn):
            x.free()
    finally:
        stream.close()
        ctx.close()
