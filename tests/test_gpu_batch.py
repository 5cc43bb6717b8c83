"""Segmented (batched) sort and its captured hipGraph: every array sorted on
its own, bit-exact against std::sort of that array."""
import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu


def make_batch(dtype, lens, seed):
    total = int(lens.sum())
    gen = "normal" if dtype != gqsort.U32 else "uniform"
    keys = gqsort.generate(dtype, gen, seed, max(total, 1))[:total]
    offs = np.zeros(len(lens) + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    return keys, offs


def check_batch(got, keys, offs, case):
    for i in range(len(offs) - 1):
        lo, hi = int(offs[i]), int(offs[i + 1])
        exp = orc.std_sort(keys[lo:hi])
        bad, first = orc.mismatch(got[lo:hi], exp)
        assert bad == 0, f"{case}: array {i} [{lo},{hi}) {bad} keys differ, first at {first}"


EDGE_LENS = [0, 1, 2, 3, 17, 64, 65, 511, 512, 513, 4095, 4096, 4097, 8192, 8193, 12000, 16384, 16385, 20000,
             100000, 0, 5]


@pytest.mark.parametrize("dtype", [gqsort.U32, gqsort.F32, gqsort.F64])
def test_segmented_edges(gpu_ctx, dtype):
    lens = np.array(EDGE_LENS, np.uint64)
    keys, offs = make_batch(dtype, lens, 3)
    print(f"case segmented dtype={dtype} arrays={len(lens)}", flush=True)
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
    dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
    gpu_ctx.sort_segmented(dtype, d, dn, od, len(lens), keys.size)
    check_batch(d.download(keys.dtype, keys.size), keys, offs, f"segmented/{dtype}")


def test_segmented_log_uniform(gpu_ctx):
    lens = gqsort.batch_lengths(2000, 16, 1 << 16, seed=4)
    keys, offs = make_batch(gqsort.U32, lens, 5)
    print(f"case segmented log-uniform total={keys.size}", flush=True)
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
    dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
    gpu_ctx.sort_segmented(gqsort.U32, d, dn, od, len(lens), keys.size)
    check_batch(d.download(keys.dtype, keys.size), keys, offs, "log-uniform")


def test_sort_each_matches(gpu_ctx):
    lens = np.array([5, 9000, 300, 70000], np.uint64)
    keys, offs = make_batch(gqsort.U32, lens, 6)
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
    dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
    gqsort._check(gqsort.load().gqs_sort_each(gpu_ctx.handle, gqsort.U32, d.ptr, dn.ptr, offs.ctypes.data,
                                               len(lens), None), gpu_ctx)
    check_batch(d.download(keys.dtype, keys.size), keys, offs, "sort_each")


def test_graph_replays(gpu_ctx):
    """Capture once, replay four times with new arrays (and a smaller count on
    two replays), with eager sorts of the same context between replays; every
    replay is checked against std::sort and against the eager result."""
    max_nseg, max_total = 600, 1 << 21
    kb = 4
    d = gqsort.DeviceBuffer(gpu_ctx, max_total * kb)
    dn = gqsort.DeviceBuffer(gpu_ctx, max_total * kb)
    od = gqsort.DeviceBuffer(gpu_ctx, (max_nseg + 1) * 8)
    nd = gqsort.DeviceBuffer(gpu_ctx, 4)
    stream = gqsort.Stream(gpu_ctx)
    g = gqsort.SegmentedGraph(gpu_ctx, gqsort.U32, d, dn, od, nd, max_nseg, max_total)
    try:
        for rep, nseg in enumerate([600, 250, 600, 40]):
            lens = gqsort.batch_lengths(nseg, 16, 20000, seed=100 + rep, total=max_total * 0.9)
            keys, offs = make_batch(gqsort.U32, lens, 200 + rep)
            assert keys.size <= max_total
            print(f"case graph replay {rep} nseg={nseg} total={keys.size}", flush=True)
            d.upload(keys, stream)
            od.upload(offs, stream)
            nd.upload(np.array([nseg], np.uint32), stream)
            g.launch(stream)
            stream.sync()
            got = d.download(keys.dtype, keys.size, stream)
            check_batch(got, keys, offs, f"graph replay {rep}")
            # eager work of the same context between replays
            e = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
            en = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
            eo = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
            gpu_ctx.sort_segmented(gqsort.U32, e, en, eo, nseg, keys.size, stream)
            stream.sync()
            assert np.array_equal(e.download(keys.dtype, keys.size, stream), got)
    finally:
        g.close()
        stream.close()


@pytest.mark.parametrize("dtype", [gqsort.U32, gqsort.F32, gqsort.F64])
def test_graph_capture_in_fresh_tuned_context(dtype):
    """A graph captured on a new context that uses the shipped tuned table:
    its workspace is sized for the parameters the capture runs with (the
    tuned class of max_total_keys, which can differ from the context's own),
    so nothing is allocated while the stream captures.  Each graph is replayed
    once and checked against std::sort."""
    ctx = gqsort.Context(0)
    try:
        kb = gqsort.KEY_BYTES[dtype]
        for max_total in (1 << 16, 1 << 20, 3 << 20):
            max_nseg = 64
            d = gqsort.DeviceBuffer(ctx, max_total * kb)
            dn = gqsort.DeviceBuffer(ctx, max_total * kb)
            od = gqsort.DeviceBuffer(ctx, (max_nseg + 1) * 8)
            nd = gqsort.DeviceBuffer(ctx, 4)
            stream = gqsort.Stream(ctx)
            print(f"case fresh-context graph dtype={dtype} max_total={max_total}", flush=True)
            g = gqsort.SegmentedGraph(ctx, dtype, d, dn, od, nd, max_nseg, max_total)
            try:
                lens = gqsort.batch_lengths(max_nseg, 16, max_total // 8, seed=7, total=max_total * 0.9)
                keys, offs = make_batch(dtype, lens, 11)
                d.upload(keys, stream)
                od.upload(offs, stream)
                nd.upload(np.array([max_nseg], np.uint32), stream)
                g.launch(stream)
                stream.sync()
                check_batch(d.download(keys.dtype, keys.size, stream), keys, offs, f"fresh graph {max_total}")
            finally:
                g.close()
                stream.close()
                for b in (d, dn, od, nd):
                    b.free()
    finally:
        ctx.close()


@pytest.mark.parametrize("dtype", [gqsort.U32, gqsort.F32, gqsort.F64])
def test_segmented_random_lengths(gpu_ctx, dtype):
    """1500 arrays with random lengths across every size threshold (wave
    leaves, 512-thread leaves, large leaves, partitioned arrays), in random
    order, each checked against std::sort."""
    rng = np.random.default_rng(77 + dtype)
    lens = np.floor(2.0 ** rng.uniform(0, np.log2(40000), 1500)).astype(np.uint64)
    lens[rng.choice(lens.size, 40, replace=False)] = rng.choice(
        [0, 1, 2, 511, 512, 513, 4095, 4096, 4097, 8191, 8192, 8193, 16383, 16384, 16385], 40)
    keys, offs = make_batch(dtype, lens, 5)
    print(f"case segmented random lengths dtype={dtype} total={keys.size}", flush=True)
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
    dn = gqsort.DeviceBuffer(gpu_ctx, max(keys.nbytes, 8))
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
    stream = gqsort.Stream(gpu_ctx)
    gpu_ctx.sort_segmented(dtype, d, dn, od, lens.size, keys.size, stream)
    stream.sync()
    check_batch(d.download(keys.dtype, keys.size, stream), keys, offs, f"random lengths dtype={dtype}")
    s = gpu_ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0
    stream.close()


@pytest.mark.parametrize("dtype", [gqsort.U32, gqsort.F32, gqsort.F64])
def test_scatter_tile_blocks(gpu_ctx, dtype):
    """Arrays of 1 ... 65 partition tiles, exact and off by one: the scatter
    takes each tile's run starts from the count matrix, loaded in blocks of 32
    tiles, and work-group ranges cross array boundaries anywhere.  Every third
    array is sorted ascending and every third constant, so that waves whose
    keys share one bucket (also in the partial last tile) take the
    one-add rank path."""
    tile = 16384 if dtype != gqsort.F64 else 8192
    lens = []
    for t in (1, 2, 31, 32, 33, 63, 64, 65):
        lens += [t * tile - 1, t * tile, t * tile + 1]
    rng = np.random.default_rng(11 + dtype)
    parts = []
    for i, n in enumerate(lens):
        a = gqsort.generate(dtype, "normal" if dtype != gqsort.U32 else "uniform", 100 + i, n)
        if i % 3 == 1:
            a = np.sort(a)
        elif i % 3 == 2:
            a = np.full(n, a[rng.integers(n)], dtype=a.dtype)
        parts.append(a)
    keys = np.concatenate(parts)
    offs = np.zeros(len(lens) + 1, np.uint64)
    offs[1:] = np.cumsum(np.array(lens, np.uint64))
    print(f"case scatter tile blocks dtype={dtype} total={keys.size}", flush=True)
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
    dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
    gpu_ctx.sort_segmented(dtype, d, dn, od, len(lens), keys.size)
    check_batch(d.download(keys.dtype, keys.size), keys, offs, f"tile blocks dtype={dtype}")
    s = gpu_ctx.stats()
    assert s.overflow == 0
