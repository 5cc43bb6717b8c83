"""The boundary's contract on the GPU: stream order between calls of one
context, a thread-safe host-pointer entry point, and bound hits reported as a
status instead of unsorted keys behind GQS_OK.

The reference reports nothing: its errors are asserts that vanish under NDEBUG
(CL12/OpenCLUtils.h:41-45) and its record buffers have no bounds check
(CL20/QuicksortMain.cpp:366).  Every condition here is forced from the host
(a capped queue, bad offsets, a count above the graph's capacity); all kernels
exit normally.
"""
import os
import threading

import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu

EOVERFLOW, EINVAL = 4, 1


def check(got, inp, case):
    exp = orc.std_sort(inp)
    bad, first = orc.mismatch(got, exp)
    assert bad == 0, f"{case}: {bad} of {inp.size} keys differ from std::sort, first at {first}"


def test_two_streams_one_context(gpu_ctx):
    """Two sorts on two streams of one context, enqueued back to back without
    a host wait: they share the context's workspace, so the second waits for
    the first on the device; both results are bit-exact."""
    n1, n2 = 1 << 24, (1 << 23) + 12345
    a = gqsort.generate(gqsort.U32, "uniform", 61, n1)
    b = gqsort.generate(gqsort.U32, "bucket", 62, n2)
    s1, s2 = gqsort.Stream(gpu_ctx), gqsort.Stream(gpu_ctx)
    da = gqsort.DeviceBuffer.from_array(gpu_ctx, a)
    dna = gqsort.DeviceBuffer(gpu_ctx, a.nbytes)
    db = gqsort.DeviceBuffer.from_array(gpu_ctx, b)
    dnb = gqsort.DeviceBuffer(gpu_ctx, b.nbytes)
    try:
        for rep in range(3):
            print(f"case two streams rep {rep}", flush=True)
            da.upload(a)
            db.upload(b)
            gpu_ctx.sort_device(gqsort.U32, da, dna, n1, s1)
            gpu_ctx.sort_device(gqsort.U32, db, dnb, n2, s2)
            gpu_ctx.sort_device(gqsort.U32, da, dna, n1, s1)  # sorted input again, back on the first stream
            s1.sync()
            s2.sync()
            check(da.download(np.uint32, n1, s1), a, "stream 1")
            check(db.download(np.uint32, n2, s2), b, "stream 2")
    finally:
        for x in (da, dna, db, dnb):
            x.free()
        s1.close()
        s2.close()


def test_host_pointer_sort_from_two_threads(gpu_ctx):
    """gqs_sort holds the context's lock from upload to download: two threads
    sorting host arrays of different sizes (the larger one regrows the
    context's device copies) on one context both get exact results."""
    inputs = [gqsort.generate(gqsort.U32, "uniform", 70 + i, n) for i, n in enumerate([(1 << 20) + 3, 3 << 22])]
    inputs.append(gqsort.generate(gqsort.F64, "normal", 72, 1 << 21))
    outs = [x.copy() for x in inputs]
    errors = []

    def worker(i):
        try:
            for _ in range(3):
                d = inputs[i].copy()
                gqsort.GPUQSort(gpu_ctx, d.size, d, np.empty_like(d))
                outs[i] = d
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    print("case host sorts from three threads", flush=True)
    th = [threading.Thread(target=worker, args=(i,)) for i in range(len(inputs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, x in enumerate(inputs):
        check(outs[i], x, f"thread {i}")


@pytest.mark.parametrize("pipelined", [True, False])
def test_queue_overflow_is_reported(gpu_ctx, pipelined):
    """A segment queue capped at one entry cannot hold level 1's child
    segments: the host-pointer sort (its copies overlapping the sort, or not)
    returns GQS_EOVERFLOW, a device sort's next gqs_get_stats does, and the
    status clears with the next good sort."""
    n = 1 << 24
    a = gqsort.generate(gqsort.U32, "uniform", 81, n)
    gpu_ctx.set_queue_limit(1)
    gpu_ctx.set_host_pipeline(1 << 16 if pipelined else 2**64 - 1)
    try:
        d = a.copy()
        with pytest.raises(gqsort.GQSError) as ei:
            gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
        assert ei.value.status == EOVERFLOW
        da = gqsort.DeviceBuffer.from_array(gpu_ctx, a)
        dn = gqsort.DeviceBuffer(gpu_ctx, a.nbytes)
        gpu_ctx.sort_device(gqsort.U32, da, dn, n)
        with pytest.raises(gqsort.GQSError) as ei:
            gpu_ctx.stats()
        assert ei.value.status == EOVERFLOW
    finally:
        gpu_ctx.set_queue_limit(0)
    d = a.copy()
    gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
    gpu_ctx.set_host_pipeline(0)
    check(d, a, "after the cap is lifted")
    s = gpu_ctx.stats()
    assert s.overflow == 0


def test_large_n_overflow_survives_later_runs(gpu_ctx):
    """The large-n path sorts several runs in one call: a bound hit in the
    first run (2^23 keys: level 1 makes many child segments) is still
    reported after the second run (2^20 keys: no child segment)
    re-initialises the level counters."""
    n = (1 << 23) + (1 << 20)
    a = gqsort.generate(gqsort.U32, "uniform", 82, n)
    gpu_ctx.set_large_run(1 << 23)
    gpu_ctx.set_queue_limit(1)
    try:
        d = a.copy()
        with pytest.raises(gqsort.GQSError) as ei:
            gqsort.GPUQSort(gpu_ctx, n, d, np.empty_like(d))
        assert ei.value.status == EOVERFLOW
    finally:
        gpu_ctx.set_queue_limit(0)
        gpu_ctx.set_large_run(0)


def test_segmented_bad_offsets_are_reported(gpu_ctx):
    """An array whose end passes total_keys, or whose offsets decrease, is
    skipped and reported as GQS_EINVAL by gqs_get_stats; the other arrays are
    sorted."""
    lens = [1000, 70000, 5, 20000]
    keys = gqsort.generate(gqsort.U32, "uniform", 90, sum(lens))
    offs = np.zeros(len(lens) + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    for case, bad_offs in (("end past total", np.append(offs, np.uint64(offs[-1] + 10))),
                           ("decreasing", np.append(offs, np.uint64(offs[-1] - 100)))):
        print(f"case bad offsets: {case}", flush=True)
        d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
        dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
        od = gqsort.DeviceBuffer.from_array(gpu_ctx, bad_offs)
        gpu_ctx.sort_segmented(gqsort.U32, d, dn, od, len(lens) + 1, keys.size)
        with pytest.raises(gqsort.GQSError) as ei:
            gpu_ctx.stats()
        assert ei.value.status == EINVAL
        got = d.download(np.uint32, keys.size)
        for i in range(len(lens)):
            lo, hi = int(offs[i]), int(offs[i + 1])
            check(got[lo:hi], keys[lo:hi], f"{case}: array {i}")
        for x in (d, dn, od):
            x.free()


def test_graph_count_above_capacity_is_reported(gpu_ctx):
    """A replay whose device count exceeds the graph's max_nseg sorts the
    first max_nseg arrays and gqs_graph_status reports GQS_EINVAL; a good
    replay afterwards reports GQS_OK."""
    max_nseg, max_total = 8, 1 << 18
    lens = np.full(max_nseg, 20000, np.uint64)
    keys = gqsort.generate(gqsort.U32, "uniform", 91, int(lens.sum()))
    offs = np.zeros(max_nseg + 2, np.uint64)
    offs[1:max_nseg + 1] = np.cumsum(lens)
    offs[-1] = offs[-2]
    d = gqsort.DeviceBuffer.from_array(gpu_ctx, np.concatenate([keys, np.zeros(max_total - keys.size, np.uint32)]))
    dn = gqsort.DeviceBuffer(gpu_ctx, max_total * 4)
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
    nd = gqsort.DeviceBuffer.from_array(gpu_ctx, np.array([max_nseg + 1], np.uint32))
    stream = gqsort.Stream(gpu_ctx)
    g = gqsort.SegmentedGraph(gpu_ctx, gqsort.U32, d, dn, od, nd, max_nseg, max_total)
    try:
        g.launch(stream)
        with pytest.raises(gqsort.GQSError) as ei:
            g.status()
        assert ei.value.status == EINVAL
        got = d.download(np.uint32, keys.size, stream)
        for i in range(max_nseg):
            lo, hi = int(offs[i]), int(offs[i + 1])
            check(got[lo:hi], keys[lo:hi], f"graph array {i}")
        nd.upload(np.array([max_nseg], np.uint32), stream)
        d.upload(keys, stream)
        g.launch(stream)
        g.status()
    finally:
        g.close()
        stream.close()
        for x in (d, dn, od, nd):
            x.free()


def test_tune_load_rejects_out_of_range_entries(gpu_ctx, tmp_path):
    """A tuning table with a fan-out above the splitter table (512) or a zero
    leaf_target is refused as a whole; the context keeps sorting with its
    previous parameters."""
    for bad in ('"max_fanout": 512', '"leaf_target": 0', '"bitonic_cutoff": 1024', '"part_wgs_per_cu": 0'):
        path = os.path.join(tmp_path, "bad.json")
        with open(path, "w") as f:
            f.write('{"device": "x", "entries": [\n')
            f.write('  {"dtype": 0, "log2n": 20, "ms": 0.1, "mode": 0, "leaf_max": 8192, "leaf_target": 2048},\n')
            f.write('  {"dtype": 0, "log2n": 22, "ms": 0.1, "mode": 0, ' + bad + '}\n]}\n')
        with pytest.raises(gqsort.GQSError) as ei:
            gpu_ctx.tune_load(path)
        assert ei.value.status == EINVAL
    path = os.path.join(tmp_path, "bad_dtype.json")
    with open(path, "w") as f:
        f.write('{"entries": [\n  {"dtype": 7, "log2n": 20, "leaf_max": 8192}\n]}\n')
    with pytest.raises(gqsort.GQSError):
        gpu_ctx.tune_load(path)
    a = gqsort.generate(gqsort.U32, "uniform", 92, 1 << 20)
    check(gqsort.device_sort(gpu_ctx, a), a, "after rejected tables")


def test_device_report(gpu_ctx):
    """The show_CL report: compute units, clocks and memory of the device."""
    r = gpu_ctx.device_report()
    print(r, flush=True)
    assert r["compute_units"] == 256 and r["wave_size"] == 64 and r["arch"].startswith("gfx950")
    assert r["clock_khz"] > 0 and r["mem_clock_khz"] > 0 and r["mem_bus_bits"] > 0
    assert r["global_mem_bytes"] > 200e9 and r["lds_per_cu"] >= 160 * 1024
    assert r["max_threads_per_block"] >= 1024
