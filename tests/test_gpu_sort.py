"""GPU parity: the HIP sort through the C ABI, bit-exact against std::sort.

The reference checks every GPU result against std::sort of the same input
(CL12/QuicksortMain.cpp:498-516); these tests do the same on seeded inputs,
plus the edge sizes around the local-sort and tile cutoffs.
"""
import numpy as np
import pytest

import gqsort
import orc
from conftest import gpu_sort

pytestmark = pytest.mark.gpu

GENS = ["uniform", "gaussian", "sorted", "reverse", "equal", "bucket", "permutation"]


def check(got, inp, case):
    exp = orc.std_sort(inp)
    bad, first = orc.mismatch(got, exp)
    assert bad == 0, f"{case}: {bad} of {inp.size} keys differ from std::sort, first at {first}"


def test_c1_uniform_2p20_seed1(gpu_ctx):
    """Plain method, small: 2^20 uniform uint32 keys from seed 1."""
    a = gqsort.generate(gqsort.U32, "uniform", 1, 1 << 20)
    print("case c1 2^20 uniform seed 1", flush=True)
    check(gpu_sort(gpu_ctx, a), a, "c1")
    s = gpu_ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("n", [2, 3, 17, 255, 4097, 8192, 8193, 16385, 65537, 300001])
def test_sizes_and_generators(gpu_ctx, gen, n):
    a = gqsort.generate(gqsort.U32, gen, 7, n)
    print(f"case u32 {gen} n={n}", flush=True)
    check(gpu_sort(gpu_ctx, a), a, f"{gen}/{n}")


@pytest.mark.parametrize("dtype,n", [(gqsort.F32, 1 << 20), (gqsort.F64, 1 << 19), (gqsort.F32, 4099),
                                     (gqsort.F64, 70001)])
def test_float_keys(gpu_ctx, dtype, n):
    a = gqsort.generate(dtype, "normal", 3, n)
    print(f"case fp dtype={dtype} n={n}", flush=True)
    check(gpu_sort(gpu_ctx, a), a, f"fp{dtype}/{n}")


def test_float_special_values(gpu_ctx):
    vals = np.array([np.inf, -np.inf, 0.0, 1.0, -1.0, np.finfo(np.float32).max, -np.finfo(np.float32).max,
                     np.finfo(np.float32).tiny, -np.finfo(np.float32).tiny, 1e-45, -1e-45], dtype=np.float32)
    rng = np.random.default_rng(5)
    a = rng.choice(vals, size=20000).astype(np.float32)
    check(gpu_sort(gpu_ctx, a), a, "f32 specials")


@pytest.mark.parametrize("host_driven", [0, 1])
@pytest.mark.parametrize("gen", ["uniform", "sorted", "equal", "bucket"])
def test_classic_mode(gpu_ctx, gen, host_driven):
    """One median-of-3 pivot per segment and a 1728-key leaf cutoff: the
    reference's schedule (CL12/Quicksort.h:91-97) on the same kernels, with
    all levels enqueued up front (host_driven=0) or with the reference's
    readback per level; nothing reaches the unbounded local sort."""
    a = gqsort.generate(gqsort.U32, gen, 11, 1 << 18)
    print(f"case classic {gen} host_driven={host_driven}", flush=True)
    try:
        got = gpu_sort(gpu_ctx, a, mode=gqsort.MODE_CLASSIC, leaf_max=1728, max_levels=40, host_driven=host_driven)
        check(got, a, f"classic/{gen}")
        s = gpu_ctx.stats()
        assert s.overflow == 0 and s.oversize_leaves == 0
    finally:
        gpu_ctx.set_params(gqsort.default_params())


def organ_pipe(n, dtype=np.uint32):
    """0, 1, .., n/2-1, n/2-1, .., 1, 0: the first, middle and last keys of
    every segment a stable single-pivot partition leaves are the segment's two
    smallest and its largest, so the median-of-3 pivot removes two keys per
    level (n/2 levels without a bound)."""
    h = np.arange(n // 2, dtype=dtype)
    return np.concatenate([h, h[::-1], np.zeros(n % 2, dtype)])


@pytest.mark.parametrize("dt,n", [(gqsort.U32, 1 << 20), (gqsort.F64, 300001)])
def test_classic_mode_adversarial(gpu_ctx, dt, n):
    """Inputs that defeat median-of-3 pivots (an organ pipe and its
    descending-then-ascending mirror) under the classic schedule with no host
    readback: the median-of-3 levels stall, the multipivot levels enqueued
    after them finish the sort, nothing reaches the unbounded local sort, and
    the level count stays at what was enqueued."""
    base = organ_pipe(n).astype(gqsort.NP_DTYPES[dt])
    try:
        for name, a in (("organ pipe", base), ("mirror", base.max() - base)):
            print(f"case classic adversarial dt={dt} n={n} {name}", flush=True)
            got = gpu_sort(gpu_ctx, a, mode=gqsort.MODE_CLASSIC, leaf_max=1728, max_levels=40, host_driven=0)
            check(got, a, f"classic adversarial {name}")
            s = gpu_ctx.stats()
            assert s.overflow == 0 and s.oversize_leaves == 0
            assert s.levels_used <= s.levels_launched <= 47
    finally:
        gpu_ctx.set_params(gqsort.default_params())


def test_tile_local_rescue(gpu_ctx):
    """Buckets of the tile-local level above leaf_max (fan-out capped at 16
    over 100,000 keys, leaf_max 2048) are gathered into contiguous segments by
    the leaf pass and partitioned again by the spare level: every key ends in a
    bounded leaf.  With fan-out 2 the spare level cannot finish either and its
    buckets take the unbounded local sort."""
    try:
        for fanout, leaf_max, bounded in ((16, 2048, True), (2, 64, False)):
            for gen in ("uniform", "bucket", "sorted"):
                a = gqsort.generate(gqsort.U32, gen, 14, 100000)
                print(f"case tile-local rescue fanout={fanout} leaf_max={leaf_max} {gen}", flush=True)
                got = gpu_sort(gpu_ctx, a, max_levels=2, leaf_max=leaf_max, leaf_target=32, max_fanout=fanout)
                check(got, a, f"rescue fanout={fanout} {gen}")
                s = gpu_ctx.stats()
                assert s.levels_launched == 2 and s.overflow == 0
                assert (s.oversize_leaves == 0) == bounded
    finally:
        gpu_ctx.set_params(gqsort.default_params())


@pytest.mark.parametrize("dt", [gqsort.F32, gqsort.F64])
def test_tile_local_rescue_float_keys(gpu_ctx, dt):
    """The rescue path with keys whose sort representation differs from their
    bits: the tile-local level's buckets above leaf_max (here below one
    work-group's capacity too) are copied together for the next level in the
    sort representation, and only sorted there, once."""
    try:
        for fanout, leaf_max in ((16, 2048), (32, 1024)):
            a = gqsort.generate(dt, "normal", 14, 100000)
            print(f"case tile-local rescue dt={dt} fanout={fanout} leaf_max={leaf_max}", flush=True)
            got = gpu_sort(gpu_ctx, a, max_levels=2, leaf_max=leaf_max, leaf_target=32, max_fanout=fanout)
            check(got, a, f"rescue dt={dt} fanout={fanout}")
            s = gpu_ctx.stats()
            assert s.levels_launched == 2 and s.levels_used == 2 and s.overflow == 0
    finally:
        gpu_ctx.set_params(gqsort.default_params())


@pytest.mark.parametrize("dt", [gqsort.U32, gqsort.F32, gqsort.F64])
@pytest.mark.parametrize("shift", [1, 3])
def test_device_pointers_off_16_byte_boundary(gpu_ctx, dt, shift):
    """d and dn anywhere a key may start (the contract only asks for the key
    type's own alignment): the sort writes exactly [d, d + n), with 16-byte
    stores only where they are aligned; the keys on either side stay."""
    n = 300001
    a = gqsort.generate(dt, "uniform" if dt == gqsort.U32 else "normal", 5, n)
    kb = a.dtype.itemsize
    lib = gpu_ctx._lib
    st = gqsort.Stream(gpu_ctx)
    d = gqsort.DeviceBuffer(gpu_ctx, (n + 8) * kb)
    dn = gqsort.DeviceBuffer(gpu_ctx, (n + 8) * kb)
    guard = np.full(n + 8, a[0], dtype=a.dtype)
    guard[:shift] = a.max()
    guard[shift + n:] = a.min()
    guard[shift:shift + n] = a
    d.upload(guard, st)
    print(f"case off-boundary pointers dt={dt} shift={shift}", flush=True)
    gpu_ctx.sort_device(dt, d.ptr + shift * kb, dn.ptr + (shift + 1) % 4 * kb, n, st)
    st.sync()
    got = d.download(a.dtype, n + 8, st)
    st.sync()
    check(got[shift:shift + n], a, f"off-boundary dt={dt} shift={shift}")
    assert np.array_equal(got[:shift].view(np.uint8), guard[:shift].view(np.uint8))
    assert np.array_equal(got[shift + n:].view(np.uint8), guard[shift + n:].view(np.uint8))


def test_unbounded_fallback(gpu_ctx):
    """With one level allowed, buckets above leaf_max take the unbounded local sort."""
    a = gqsort.generate(gqsort.U32, "uniform", 13, 100000)
    got = gpu_sort(gpu_ctx, a, max_levels=1, leaf_max=64, leaf_target=32, max_fanout=2)
    check(got, a, "fallback")
    s = gpu_ctx.stats()
    assert s.oversize_leaves > 0
    gpu_ctx.set_params(gqsort.default_params())


def test_host_pointer_entry(gpu_ctx):
    """gqs_sort with host pointers: the reference's GPUQSort(ctx, N, d, dn)."""
    n = 1 << 18
    d = gqsort.generate(gqsort.U32, "permutation", 1, n)
    inp = d.copy()
    dn = np.empty_like(d)
    gqsort.GPUQSort(gpu_ctx, n, d, dn)
    check(d, inp, "GPUQSort host")


@pytest.mark.parametrize("dt,n", [(gqsort.U32, 4 * (4 << 20) + 12345), (gqsort.F64, 3 * (2 << 20) + 7),
                                  (gqsort.F32, (4 << 20) - 1)])
def test_host_pointer_staging_chunks(gpu_ctx, dt, n):
    """Host-pointer sorts are staged through pinned 16 MiB chunks on four copy
    lanes: sizes that span several chunks and every lane, and end mid-chunk
    (and one just under a single chunk), come back bit-exact; the second call
    reuses the staging buffers."""
    gen = "uniform" if dt == gqsort.U32 else "normal"
    gpu_ctx.set_host_pipeline(NEVER)
    try:
        for rep in range(2):
            d = gqsort.generate(dt, gen, 30 + rep, n)
            inp = d.copy()
            print(f"case host staging dt={dt} n={n} rep={rep}", flush=True)
            gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
            check(d, inp, f"host staging dt={dt} n={n}")
    finally:
        gpu_ctx.set_host_pipeline(0)


NEVER = 2**64 - 1


@pytest.mark.parametrize("dt,gen,n", [(gqsort.U32, "uniform", 1 << 16), (gqsort.U32, "sorted", (1 << 16) + 1),
                                      (gqsort.U32, "reverse", 3 * (1 << 20) + 7), (gqsort.U32, "equal", 1 << 20),
                                      (gqsort.U32, "bucket", (1 << 26) + 11), (gqsort.F32, "normal", (5 << 20) + 3),
                                      (gqsort.F64, "normal", (3 << 20) + 5)])
def test_host_pointer_pipelined(gpu_ctx, dt, gen, n):
    """gqs_sort with its copies overlapping the sort (eight runs sorted as
    they arrive, merged in a tree, the final merge's eight output ranges
    downloaded in turn): bit-exact from the smallest size the path takes
    (2^16) to 2^26 keys, sizes that split unevenly into runs and merge tiles."""
    gpu_ctx.set_host_pipeline(1 << 16)
    try:
        for rep in range(2):
            d = gqsort.generate(dt, gen, 40 + rep, n)
            inp = d.copy()
            print(f"case host pipelined dt={dt} {gen} n={n} rep={rep}", flush=True)
            gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
            check(d, inp, f"host pipelined dt={dt} {gen} n={n}")
            s = gpu_ctx.stats()
            assert s.overflow == 0
    finally:
        gpu_ctx.set_host_pipeline(0)


@pytest.mark.parametrize("pipe_min", [1 << 16, NEVER])
def test_host_pointer_pinned(gpu_ctx, pipe_min):
    """Host arrays in pinned memory (gqs_host_alloc) are copied by the DMA
    engines directly, with and without the overlapped path: bit-exact."""
    import ctypes

    lib = gqsort.load()
    n = 3 * (1 << 20) + 11
    ptrs = []
    gpu_ctx.set_host_pipeline(pipe_min)
    try:
        for dt in (gqsort.U32, gqsort.F64):
            arrs = []
            for _ in range(2):
                p = ctypes.c_void_p()
                gqsort._check(lib.gqs_host_alloc(gpu_ctx.handle, n * gqsort.KEY_BYTES[dt], ctypes.byref(p)), gpu_ctx)
                ptrs.append(p)
                ct = ctypes.c_uint32 if dt == gqsort.U32 else ctypes.c_double
                arrs.append(np.ctypeslib.as_array((ct * n).from_address(p.value)))
            d, dn = arrs
            inp = gqsort.generate(dt, "uniform" if dt == gqsort.U32 else "normal", 50, n)
            d[:] = inp
            print(f"case host pinned dt={dt} pipe_min={pipe_min}", flush=True)
            gqsort.GPUQSort(gpu_ctx, n, d, dn)
            check(d, inp, f"host pinned dt={dt}")
    finally:
        gpu_ctx.set_host_pipeline(0)
        for p in ptrs:
            lib.gqs_host_free(gpu_ctx.handle, p)


@pytest.mark.parametrize("gen", ["uniform", "sorted", "equal", "bucket"])
def test_large_2p24_properties(gpu_ctx, gen):

    n = 1 << 24
    a = gqsort.generate(gqsort.U32, gen, 21, n)
    print(f"case 2^24 {gen}", flush=True)
    got = gpu_sort(gpu_ctx, a)
    check(got, a, f"2^24/{gen}")


@pytest.mark.parametrize("dt,sizes", [(gqsort.U32, [8193, 12345, 16384]), (gqsort.F32, [8193, 16384]),
                                      (gqsort.F64, [4097, 6001, 8192])])
@pytest.mark.parametrize("gen", ["uniform", "sorted", "equal", "bucket"])
def test_large_leaf_sizes(gpu_ctx, dt, sizes, gen):
    """Arrays above one 512-thread work-group's capacity (8192 keys, f64 4096)
    and up to leaf_max (twice that) are one leaf for the 1024-thread kernel."""
    gpu_ctx.set_params(gqsort.default_params(dt))
    try:
        for n in sizes:
            g = gen if dt == gqsort.U32 else ("normal" if gen in ("uniform", "bucket") else gen)
            if dt != gqsort.U32 and g in ("sorted", "equal"):
                a = np.sort(gqsort.generate(dt, "normal", 9, n)) if g == "sorted" else np.full(n, 1.5, gqsort.NP_DTYPES[dt])
            else:
                a = gqsort.generate(dt, g, 9, n)
            print(f"case large leaf dt={dt} {gen} n={n}", flush=True)
            check(gpu_sort(gpu_ctx, a), a, f"large leaf dt={dt} {gen} n={n}")
            s = gpu_ctx.stats()
            assert s.levels_launched == 0 and s.overflow == 0 and s.oversize_leaves == 0
    finally:
        gpu_ctx.set_params(gqsort.default_params())


def test_large_leaves_from_a_level(gpu_ctx):
    """Level buckets between the 512-thread capacity and leaf_max go to the
    large-leaf list (leaf_target chosen so that many buckets land there)."""
    a = gqsort.generate(gqsort.U32, "uniform", 12, 1 << 24)
    got = gpu_sort(gpu_ctx, a, leaf_target=8192, max_fanout=256)
    check(got, a, "large leaves from a level")
    s = gpu_ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0
    gpu_ctx.set_params(gqsort.default_params())


@pytest.mark.parametrize("dt", [gqsort.U32, gqsort.F32, gqsort.F64])
@pytest.mark.parametrize("distinct", [3, 17, 1000])
def test_duplicate_heavy(gpu_ctx, dt, distinct):
    """Few distinct values with skewed (Zipf-like) frequencies: repeated
    splitters give equality buckets that the leaf pass fills, and waves whose
    keys share a bucket take the aggregated histogram and rank paths."""
    n = 3 << 20
    rng = np.random.default_rng(distinct)
    w = 1.0 / np.arange(1, distinct + 1) ** 1.2
    idx = rng.choice(distinct, size=n, p=w / w.sum())
    if dt == gqsort.U32:
        vals = rng.choice(np.iinfo(np.uint32).max, size=distinct, replace=False).astype(np.uint32)
    else:
        vals = rng.standard_normal(distinct).astype(gqsort.NP_DTYPES[dt]) * 1000
    a = vals[idx]
    print(f"case duplicates dt={dt} distinct={distinct}", flush=True)
    check(gpu_sort(gpu_ctx, a), a, f"duplicates dt={dt} distinct={distinct}")
    s = gpu_ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0
