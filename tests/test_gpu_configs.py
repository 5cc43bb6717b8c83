"""Every benchmark configuration at its stated size, bit-exact against
std::sort (the reference verifies each run at the size it times,
CL12/QuicksortMain.cpp:498-516, over the grid of build_me.sh:11-20).

  C2  2^28 uint32 keys under the six generators (seed 1, the bench's inputs)
  C3  2^27 float and 2^26 double keys ~ N(0,1) with 1% wide outliers and 0.1%
      duplicates (seed 1)
  C4  65,536 uint32 arrays, lengths floor(2^U), U ~ U[4, 20], rescaled to about
      2^28 keys, in one captured segmented sort; every array checked
  C5  a serving stream: sizes 2^10 .. 2^26, 60/30/10 uint32/float/double, under
      the 64 GiB budget; every request checked

std::sort of the next input runs on a host thread while the GPU sorts the
current one (ctypes releases the interpreter lock during the oracle's call).
"""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import gqsort
import orc

pytestmark = pytest.mark.gpu

N28 = 1 << 28


def check(got, exp, case):
    bad, first = orc.mismatch(got, exp)
    assert bad == 0, f"{case}: {bad} of {exp.size} keys differ from std::sort, first at {first}"


@pytest.fixture(scope="module")
def big_buffers(gpu_ctx):
    d = gqsort.DeviceBuffer(gpu_ctx, N28 * 4)
    dn = gqsort.DeviceBuffer(gpu_ctx, N28 * 4)
    stream = gqsort.Stream(gpu_ctx)
    yield d, dn, stream
    d.free()
    dn.free()
    stream.close()


def sort_and_check(ctx, bufs, dt, a, case):
    d, dn, stream = bufs
    exp = [None]
    t = threading.Thread(target=lambda: exp.__setitem__(0, orc.std_sort(a)))
    t.start()
    d.upload(a, stream)
    ctx.sort_device(dt, d, dn, a.size, stream)
    stream.sync()
    got = d.download(a.dtype, a.size, stream)
    t.join()
    check(got, exp[0], case)
    s = ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0, case


@pytest.mark.parametrize("gen", ["uniform", "gaussian", "sorted", "reverse", "equal", "bucket"])
def test_c2_2p28_u32(gpu_ctx, big_buffers, gen):
    a = gqsort.generate(gqsort.U32, gen, 1, N28)
    print(f"case C2 2^28 u32 {gen} seed 1", flush=True)
    sort_and_check(gpu_ctx, big_buffers, gqsort.U32, a, f"C2 {gen}")


@pytest.mark.parametrize("dt,n", [(gqsort.F32, 1 << 27), (gqsort.F64, 1 << 26)])
def test_c3_float_keys(gpu_ctx, big_buffers, dt, n):
    a = gqsort.generate(dt, "normal", 1, n)
    assert not np.isnan(a).any() and not np.any((a == 0) & np.signbit(a))  # no NaN, no -0
    print(f"case C3 dtype={dt} n={n}", flush=True)
    sort_and_check(gpu_ctx, big_buffers, dt, a, f"C3 dtype={dt}")


def test_c4_65536_arrays(gpu_ctx, big_buffers):
    """The bench's batch: one captured segmented sort (and one eager) of
    65,536 arrays; every array equals std::sort of that array."""
    d, dn, stream = big_buffers
    nseg = 65536
    lens = gqsort.batch_lengths(nseg, 16, 1 << 20, seed=1, total=N28)
    total = int(lens.sum())
    assert total <= N28 and lens.min() >= 1
    offs = np.zeros(nseg + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    keys = gqsort.generate(gqsort.U32, "uniform", 2, total)
    exp = [None]
    t = threading.Thread(target=lambda: exp.__setitem__(0, orc.std_sort_segmented(keys, offs)))
    t.start()
    od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs, stream)
    nd = gqsort.DeviceBuffer.from_array(gpu_ctx, np.array([nseg], np.uint32), stream)
    print(f"case C4 {nseg} arrays, {total} keys (graph)", flush=True)
    g = gqsort.SegmentedGraph(gpu_ctx, gqsort.U32, d, dn, od, nd, nseg, total)
    try:
        d.upload(keys, stream)
        g.launch(stream)
        g.status()
        got_graph = d.download(np.uint32, total, stream)
    finally:
        g.close()
    print("case C4 eager segmented sort", flush=True)
    d.upload(keys, stream)
    gpu_ctx.sort_segmented(gqsort.U32, d, dn, od, nseg, total, stream)
    stream.sync()
    got_eager = d.download(np.uint32, total, stream)
    s = gpu_ctx.stats()
    assert s.overflow == 0 and s.oversize_leaves == 0
    t.join()
    check(got_graph, exp[0], "C4 graph")
    check(got_eager, exp[0], "C4 eager")
    od.free()
    nd.free()


def test_c5_serving_stream():
    """96 requests of the bench's stream (sizes 2^10 .. 2^26, mixed key types)
    submitted at once to a server with the 64 GiB budget; every result equals
    std::sort of its request."""
    sizes, kinds, seeds = gqsort.request_stream(96, seed=8)
    assert sizes.max() >= 1 << 24 and set(kinds.tolist()) == {gqsort.U32, gqsort.F32, gqsort.F64}
    reqs = [gqsort.request_keys(k, s, n) for n, k, s in zip(sizes, kinds, seeds)]
    print(f"case C5 {len(reqs)} requests, {int(sizes.sum())} keys", flush=True)
    with ThreadPoolExecutor(8) as ex:
        expected = ex.map(orc.std_sort, reqs)
        with gqsort.Server(0, hbm_budget_bytes=64 << 30) as srv:
            pending = [srv.submit_array(a) for a in reqs]
            for t, _ in pending:
                srv.wait(t)
            st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        for i, (exp, (_, out)) in enumerate(zip(expected, pending)):
            check(out, exp, f"C5 request {i} (n={sizes[i]}, dtype={kinds[i]})")


def test_golden_fixtures_through_hip(gpu_ctx):
    """The committed fixtures (sizes 0 .. 8193 and 65,537, every generator and
    key type, expected outputs from numpy.sort) sorted by gqs_sort_device with
    the default parameters, with the reference's classic tunables (one
    median-of-3 pivot, QUICKSORT_BLOCK_SIZE 1728 leaves, CL12/Quicksort.h:91-97)
    and packed into one segmented sort: every output equals the stored one."""
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.npz")
    fx = np.load(path)
    cases = sorted({k[:-3] for k in fx.files if k.endswith("_in")})
    assert len(cases) >= 150
    setups = [("default", {}), ("classic", dict(mode=gqsort.MODE_CLASSIC, leaf_max=1728, max_levels=40))]
    try:
        for name, params in setups:
            for c in cases:
                a, e = fx[c + "_in"], fx[c + "_out"]
                if params:
                    gpu_ctx.set_params(gqsort.default_params(gqsort.DTYPE_CODES[a.dtype]), **params)
                got = gqsort.device_sort(gpu_ctx, a) if a.size else a.copy()
                check(got, e, f"fixture {c} ({name})")
    finally:
        gpu_ctx.set_params(gqsort.default_params())
        gpu_ctx.tune_load(gqsort.TUNED_TABLE)
    for dt in (gqsort.U32, gqsort.F32, gqsort.F64):
        group = [c for c in cases if fx[c + "_in"].dtype == gqsort.NP_DTYPES[dt]]
        ins = [fx[c + "_in"] for c in group]
        offs = np.zeros(len(ins) + 1, np.uint64)
        offs[1:] = np.cumsum([x.size for x in ins])
        keys = np.concatenate(ins)
        d = gqsort.DeviceBuffer.from_array(gpu_ctx, keys)
        dn = gqsort.DeviceBuffer(gpu_ctx, keys.nbytes)
        od = gqsort.DeviceBuffer.from_array(gpu_ctx, offs)
        gpu_ctx.sort_segmented(dt, d, dn, od, len(ins), keys.size)
        got = d.download(keys.dtype, keys.size)
        for i, c in enumerate(group):
            check(got[int(offs[i]):int(offs[i + 1])], fx[c + "_out"], f"fixture {c} (segmented)")
        for x in (d, dn, od):
            x.free()
